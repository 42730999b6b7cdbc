// zsc_host.cpp -- host-side inputs of libzsc: the reference's two input files,
// the parameter derivation, a seeded bubble generator and the host initial
// order parameter.  Plain C++ (no HIP calls).
#include <errno.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/zsc.h"
#include "zsc_err.h"

using zsc::fail;

namespace {

// Whitespace tokenizer over a whole file (the reference reads with fscanf
// "%s"/"%d"/"%f", so any whitespace separates tokens).
struct Tokens {
    std::vector<std::string> t;
    size_t pos = 0;
    bool load(const char *path)
    {
        FILE *fp = fopen(path, "r");
        if (!fp) return false;
        char buf[512];
        while (fscanf(fp, "%511s", buf) == 1) t.emplace_back(buf);
        fclose(fp);
        return true;
    }
    bool skip(int k)
    {
        if (pos + k > t.size()) return false;
        pos += k;
        return true;
    }
    bool next_int(int *v)
    {
        if (pos >= t.size()) return false;
        char *end = nullptr;
        errno = 0;
        long x = strtol(t[pos].c_str(), &end, 10);
        if (!end || *end || errno == ERANGE || x < INT_MIN || x > INT_MAX) return false;
        *v = (int)x;
        ++pos;
        return true;
    }
    bool next_dbl(double *v)
    {
        if (pos >= t.size()) return false;
        char *end = nullptr;
        double x = strtod(t[pos].c_str(), &end);
        if (!end || *end) return false;
        *v = x;
        ++pos;
        return true;
    }
};

}  // namespace

extern "C" int zsc_read_inputs(const char *path, zsc_inputs *in)
{
    if (!path || !in) return fail(ZSC_ERR_ARG, "zsc_read_inputs: null argument");
    Tokens tk;
    if (!tk.load(path)) return fail(ZSC_ERR_IO, "cannot open %s", path);
    memset(in, 0, sizeof(*in));
    in->phi_star = 0.0;
    in->gravity = -1.0;
    // Token order of src/paramread.c:40-59: header words, then values.
    bool ok = tk.skip(4) && tk.next_int(&in->max_step) && tk.next_int(&in->relax_step) &&
              tk.next_int(&in->save) && tk.next_int(&in->stat) && tk.skip(3) &&
              tk.next_int(&in->xmax) && tk.next_int(&in->ymax) && tk.next_int(&in->zmax) &&
              tk.skip(2) && tk.next_int(&in->block_x) && tk.next_int(&in->block_y) &&
              tk.skip(2) && tk.next_dbl(&in->rho_l) && tk.next_dbl(&in->rho_h) && tk.skip(2) &&
              tk.next_dbl(&in->tau_rho) && tk.next_dbl(&in->tau_phi) && tk.skip(3) &&
              tk.next_dbl(&in->width) && tk.next_dbl(&in->sigma) && tk.next_dbl(&in->gamma) &&
              tk.skip(1) && tk.next_dbl(&in->eo);
    if (!ok) return fail(ZSC_ERR_IO, "%s: malformed near token %zu", path, tk.pos);
    // Optional trailing overrides: "key value" pairs.
    while (tk.pos < tk.t.size()) {
        std::string key = tk.t[tk.pos++];
        double v;
        if (!tk.next_dbl(&v)) return fail(ZSC_ERR_IO, "%s: override %s has no value", path, key.c_str());
        if (key == "phiStar")
            in->phi_star = v;
        else if (key == "gravity")
            in->gravity = v;
        else
            return fail(ZSC_ERR_IO, "%s: unknown override %s", path, key.c_str());
    }
    if (in->xmax <= 0 || in->ymax <= 0 || in->zmax <= 0)
        return fail(ZSC_ERR_ARG, "%s: domain sizes must be positive", path);
    if (in->tau_rho <= 0.5 || in->tau_phi <= 0.5)
        return fail(ZSC_ERR_ARG, "%s: tauRho and tauPhi must exceed 1/2", path);
    if (in->width <= 0.0) return fail(ZSC_ERR_ARG, "%s: width must be positive", path);
    if (in->phi_star <= 0.0 && !(in->rho_h > in->rho_l))
        return fail(ZSC_ERR_ARG, "%s: rhoH must exceed rhoL", path);
    return ZSC_OK;
}

extern "C" int zsc_read_bubbles(const char *path, double *bubbles, int capacity, int *n_out)
{
    if (!path || !n_out) return fail(ZSC_ERR_ARG, "zsc_read_bubbles: null argument");
    Tokens tk;
    if (!tk.load(path)) return fail(ZSC_ERR_IO, "cannot open %s", path);
    int nb = 0;
    // src/paramread.c:64-73: "nBubbles" <n> "xc yc zc Rc" then n rows.
    if (!(tk.skip(1) && tk.next_int(&nb) && nb >= 0 && tk.skip(4)))
        return fail(ZSC_ERR_IO, "%s: malformed header", path);
    *n_out = nb;
    if (!bubbles) return ZSC_OK;
    if (nb > capacity) return fail(ZSC_ERR_ARG, "%s: %d bubbles exceed capacity %d", path, nb, capacity);
    for (int i = 0; i < 4 * nb; ++i)
        if (!tk.next_dbl(&bubbles[i])) return fail(ZSC_ERR_IO, "%s: bubble row %d malformed", path, i / 4);
    return ZSC_OK;
}

extern "C" int zsc_write_bubbles(const char *path, const double *b, int n)
{
    if (!path || (n > 0 && !b) || n < 0) return fail(ZSC_ERR_ARG, "zsc_write_bubbles: bad argument");
    FILE *fp = fopen(path, "w");
    if (!fp) return fail(ZSC_ERR_IO, "cannot write %s", path);
    fprintf(fp, "nBubbles\n%d\n\nxc\tyc\tzc\tRc\n", n);
    for (int i = 0; i < n; ++i)
        fprintf(fp, "%.4f\t%.4f\t%.4f\t%.4f\n", b[4 * i], b[4 * i + 1], b[4 * i + 2], b[4 * i + 3]);
    fclose(fp);
    return ZSC_OK;
}

extern "C" int zsc_derive_params(const zsc_inputs *in, double r0, zsc_params *p)
{
    if (!in || !p) return fail(ZSC_ERR_ARG, "zsc_derive_params: null argument");
    memset(p, 0, sizeof(*p));
    // src/paramread.c:141-170, in FP64.
    p->phi_star = in->phi_star > 0.0 ? in->phi_star : 0.5 * (in->rho_h - in->rho_l);
    p->phi_star_sq = p->phi_star * p->phi_star;
    p->rho_mean = 0.5 * (in->rho_h + in->rho_l);
    const double inv_tau_rho = 1.0 / in->tau_rho;
    p->inv_tau_rho_one = 1.0 - inv_tau_rho;
    const double inv_tau_rho_half = 1.0 - 0.5 * inv_tau_rho;
    p->eta = 1.0 / (in->tau_phi + 0.5);
    p->eta2 = 1.0 - p->eta;
    p->inv_eta2 = 0.5 / (1.0 - p->eta);
    p->inv_tau_phi = 1.0 / in->tau_phi;
    p->inv_tau_phi_one = 1.0 - p->inv_tau_phi;
    p->alpha = 0.75 * in->sigma / (in->width * p->phi_star_sq * p->phi_star_sq);
    p->alpha4 = p->alpha * 4.0;
    p->kappa = (in->width * p->phi_star) * (in->width * p->phi_star) * p->alpha * 0.5;
    p->gamma = in->gamma;
    if (in->gravity >= 0.0) {
        p->gravity = in->gravity;
    } else {
        if (!(r0 > 0.0)) return fail(ZSC_ERR_ARG, "gravity from Eo needs a positive bubble radius");
        p->gravity = 0.25 * in->eo * in->sigma / ((in->rho_h - in->rho_l) * r0 * r0);
    }
    p->grav = 2.0 * p->phi_star * p->gravity;
    // src/constants.h:52-59 weights scaled by 1/tau and (1 - 1/(2 tau)).
    p->k0 = inv_tau_rho * (1.0 / 3.0);
    p->k1 = inv_tau_rho * (1.0 / 18.0);
    p->k2 = inv_tau_rho * (1.0 / 36.0);
    p->kc0 = inv_tau_rho_half * 1.0;
    p->kc1 = inv_tau_rho_half * (1.0 / 6.0);
    p->kc2 = inv_tau_rho_half * (1.0 / 12.0);
    p->width = in->width;
    p->sigma = in->sigma;
    return ZSC_OK;
}

namespace {
struct SplitMix64 {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};
// The value that a "%.4f" discrete.in row reads back as.
double quantize(double v)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%.4f", v);
    return strtod(buf, nullptr);
}
}  // namespace

extern "C" int zsc_generate_bubbles(uint64_t seed, int n, double r_min, double r_max, double gap,
                                    double width, int nx, int ny, int nz, int max_tries,
                                    double *b)
{
    if (n < 0 || !b || r_min <= 0.0 || r_max < r_min || max_tries <= 0)
        return fail(ZSC_ERR_ARG, "zsc_generate_bubbles: bad argument");
    SplitMix64 rng{seed};
    int placed = 0, tries = 0;
    while (placed < n) {
        if (tries++ >= max_tries)
            return fail(ZSC_ERR_ARG, "placed %d of %d bubbles in %d tries", placed, n, max_tries);
        const double r = quantize(r_min + (r_max - r_min) * rng.uniform());
        const double m = r + width + 1.0;  // keep R + W + 1 from every face
        const double ux = rng.uniform(), uy = rng.uniform(), uz = rng.uniform();
        if (nx - 1 < 2 * m || ny - 1 < 2 * m || nz - 1 < 2 * m)
            return fail(ZSC_ERR_ARG, "box too small for radius %.4f", r);
        const double xc = quantize(m + (nx - 1 - 2 * m) * ux);
        const double yc = quantize(m + (ny - 1 - 2 * m) * uy);
        const double zc = quantize(1.0 + m + (nz - 1 - 2 * m) * uz);  // z is 1-based
        bool ok = true;
        for (int j = 0; j < placed && ok; ++j) {
            const double dx = xc - b[4 * j], dy = yc - b[4 * j + 1], dz = zc - b[4 * j + 2];
            const double need = r + b[4 * j + 3] + gap;
            ok = dx * dx + dy * dy + dz * dz >= need * need;
        }
        if (!ok) continue;
        b[4 * placed + 0] = xc;
        b[4 * placed + 1] = yc;
        b[4 * placed + 2] = zc;
        b[4 * placed + 3] = r;
        ++placed;
    }
    return ZSC_OK;
}

extern "C" int zsc_init_phi_host(int nx, int ny, int nz, const double *b, int n, double phi_star,
                                 double width, double *phi)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !phi || n < 0 || (n > 0 && !b))
        return fail(ZSC_ERR_ARG, "zsc_init_phi_host: bad argument");
    const size_t total = (size_t)nx * ny * nz;
    for (size_t i = 0; i < total; ++i) phi[i] = -phi_star;
    // src/init.c:39-59: bubbles in file order, later ones overwrite; only the
    // bounding box of each bubble can satisfy r <= R + W, so visit just that.
    for (int m = 0; m < n; ++m) {
        const double xc = b[4 * m], yc = b[4 * m + 1], zc = b[4 * m + 2], rc = b[4 * m + 3];
        const double reach = rc + width;
        const int x0 = (int)fmax(0.0, floor(xc - reach)), x1 = (int)fmin(nx - 1.0, ceil(xc + reach));
        const int y0 = (int)fmax(0.0, floor(yc - reach)), y1 = (int)fmin(ny - 1.0, ceil(yc + reach));
        // stored plane z is the reference's 1-based plane z+1
        const int z0 = (int)fmax(0.0, floor(zc - reach) - 1), z1 = (int)fmin(nz - 1.0, ceil(zc + reach) - 1);
        for (int z = z0; z <= z1; ++z)
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    const double dx = (double)x - xc, dy = (double)y - yc, dz = (double)(z + 1) - zc;
                    const double r = sqrt(dx * dx + dy * dy + dz * dz);
                    if (r <= reach) phi[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)] =
                        phi_star * tanh(2.0 * (rc - r) / width);
                }
    }
    return ZSC_OK;
}

extern "C" double zsc_bytes_per_cell(int precision)
{
    // 26 populations read and written once per step; the 16-bit lattices keep
    // f_0 in FP32.
    switch (precision) {
    case ZSC_F64: return 2.0 * 26 * 8;
    case ZSC_F32:
    case ZSC_F32D: return 2.0 * 26 * 4;
    case ZSC_BF16:
    case ZSC_FP16: return 2.0 * (4 + 25 * 2);
    default: return 0.0;
    }
}
