/*
 * zsc_oracle.c -- FP64 CPU restatement of the ZSC step of
 * carlosrosales/lbm_bench_gpu.  TEST INFRASTRUCTURE ONLY (see zsc_oracle.h):
 * the checker for the HIP product path, never part of it.
 *
 * Every expression keeps the operand order of the reference's CUDA source so
 * that an FP64 build without FMA contraction (-ffp-contract=off) produces the
 * same bits as any other faithful FP64 evaluation.  This matters: the gravity
 * switch "phi > 0" (src/collision_f.cu:96) is evaluated on cells whose phi is
 * rounding noise around zero, and a flipped branch injects 2*phi*.g there.
 *
 * Ghost planes and the MPI halo exchange of the reference collapse, for one
 * rank on a periodic communicator (src/vgrid.c:31,59), to a periodic wrap in z
 * (src/pack_mpi*.cu, src/unpack_mpi*.cu); x and y wrap through the neighbour
 * tables of src/init.c:62-82.  Here every direction wraps by index arithmetic.
 */
#include "zsc_oracle.h"

#include <math.h>
#include <stddef.h>
#include <string.h>

/* Lattice tables.  Velocities follow the streaming targets of
 * src/stream_f.cu:52-59 and src/stream_g.cu:56-76; f uses the first 7. */
static const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};

/* The precision of the step's arithmetic and arrays: double (the FP64
 * oracle, liboracle.so) or float (liboracle_f32.so, built with
 * -DZO_REAL=float: the same expressions evaluated in FP32, as the reference
 * evaluates them, with every constant and coefficient rounded to FP32 first).
 * The parameter derivation, the initial profile and the statistics' sums
 * stay FP64 in both. */
typedef zo_real real;
#define RL(x) ((real)(x))

/* Weights of src/constants.h:43-59 at the step's precision. */
#define INV6 RL(1.0 / 6.0)
#define INV12 RL(1.0 / 12.0)
#define INV_CS_SQ RL(3.0)
#define CS_SQ RL(1.0 / 3.0)
#define W0 RL(1.0 / 3.0)
#define W1 RL(1.0 / 18.0)
#define W2 RL(1.0 / 36.0)
#define WC0 1.0
#define WC1 (1.0 / 6.0)
#define WC2 (1.0 / 12.0)

/* Planted mistakes for the sensitivity tests of the oracle's pins: a value
 * other than ZO_MUT_NONE changes one term of one function the way a
 * plausible slip would (zsc_oracle.h lists them).  ZO_MUT_NONE, the default,
 * is the reference's arithmetic, evaluated exactly as without the switch. */
static int g_mut = ZO_MUT_NONE;

void zo_set_mutation(int m) { g_mut = m; }

static inline int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

static inline size_t cell(int nx, int ny, int x, int y, int z)
{
    return (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z);
}

/* p[d] = phi(x + c_d) for the 19 g directions, periodic. */
static inline void gather19(int nx, int ny, int nz, const real *phi, int x,
                            int y, int z, real p[19])
{
    for (int d = 0; d < 19; ++d)
        p[d] = phi[cell(nx, ny, wrap(x + CX[d], nx), wrap(y + CY[d], ny),
                        wrap(z + CZ[d], nz))];
}

/* Laplacian of src/collision_f.cu:61-70 (same in init_f/init_g/collision_g),
 * with phi(ie,jn,k)=p7, phi(iw,js,k)=p8, ... mapped to g directions. */
static inline real lap19(const real p[19])
{
    return (p[7] + p[8] + p[9] + p[10] + p[11] + p[12] + p[13] + p[14] + p[15] +
            p[16] + p[17] + p[18] +
            RL(2.0) * (p[1] + p[2] + p[3] + p[4] + p[5] + p[6] - RL(12.0) * p[0])) *
           INV6;
}

/* Isotropic gradient bracket of src/collision_f.cu:78-94 (before mu*, *inv12). */
static inline real gx19(const real p[19])
{
    return RL(2.0) * (p[1] - p[2]) + p[7] - p[8] + p[9] - p[10] + p[11] - p[12] +
           p[13] - p[14];
}
static inline real gy19(const real p[19])
{
    return RL(2.0) * (p[3] - p[4]) + p[7] - p[8] + p[10] - p[9] + p[15] - p[16] +
           p[17] - p[18];
}
static inline real gz19(const real p[19])
{
    return RL(2.0) * (p[5] - p[6]) + p[11] - p[12] + p[14] - p[13] + p[15] - p[16] +
           p[18] - p[17];
}

void zo_derive(double rho_l, double rho_h, double tau_rho, double tau_phi,
               double width, double sigma, double gamma, double eo, double r0,
               double phi_star_override, double gravity_override, zo_coef *c)
{
    /* src/paramread.c:141-170 */
    memset(c, 0, sizeof(*c));
    c->phi_star = phi_star_override > 0.0 ? phi_star_override : 0.5 * (rho_h - rho_l);
    c->phi_star_sq = c->phi_star * c->phi_star;
    c->rho_mean = 0.5 * (rho_h + rho_l); /* src/init.c:45 */
    double inv_tau_rho = 1.0 / tau_rho;
    c->inv_tau_rho_one = 1.0 - inv_tau_rho;
    double inv_tau_rho_half = 1.0 - 0.5 * inv_tau_rho;
    c->eta = 1.0 / (tau_phi + 0.5);
    c->eta2 = 1.0 - c->eta;
    c->inv_eta2 = 0.5 / (1.0 - c->eta);
    c->inv_tau_phi = 1.0 / tau_phi;
    c->inv_tau_phi_one = 1.0 - c->inv_tau_phi;
    c->alpha = 0.75 * sigma / (width * c->phi_star_sq * c->phi_star_sq);
    c->alpha4 = c->alpha * 4.0;
    c->kappa = (width * c->phi_star) * (width * c->phi_star) * c->alpha * 0.5;
    c->gamma = gamma;
    if (gravity_override >= 0.0)
        c->gravity = gravity_override;
    else
        c->gravity = 0.25 * eo * sigma / ((rho_h - rho_l) * r0 * r0);
    c->grav = 2.0 * c->phi_star * c->gravity;
    c->k0 = inv_tau_rho * W0;
    c->k1 = inv_tau_rho * W1;
    c->k2 = inv_tau_rho * W2;
    c->kc0 = inv_tau_rho_half * WC0;
    c->kc1 = inv_tau_rho_half * WC1;
    c->kc2 = inv_tau_rho_half * WC2;
    c->width = width;
    c->sigma = sigma;
}

void zo_init_phi(int nx, int ny, int nz, const double *bubbles, int nb,
                 double phi_star, double width, double *phi)
{
    /* src/init.c:39-59; the reference's plane k (with zlg = 0 on one rank,
     * src/vgrid.c:38-52) is 1-based, so stored plane z has dz = (z+1) - zc. */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                double v = -phi_star;
                for (int m = 0; m < nb; ++m) {
                    double dx = (double)x - bubbles[4 * m + 0];
                    double dy = (double)y - bubbles[4 * m + 1];
                    double dz = (double)(z + 1) - bubbles[4 * m + 2];
                    double rc = bubbles[4 * m + 3];
                    double r = sqrt(dx * dx + dy * dy + dz * dz);
                    if (r <= rc + width)
                        v = phi_star * tanh(2.0 * (rc - r) / width);
                }
                phi[cell(nx, ny, x, y, z)] = v;
            }
}

void zo_init_f(int nx, int ny, int nz, const zo_coef *c, const real *phi,
               real *f)
{
    const size_t n = (size_t)nx * ny * nz;
    /* src/init_f.cu:42-79 */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
                real p[19];
                gather19(nx, ny, nz, phi, x, y, z, p);
                real phin = p[0];
                real lap = lap19(p);
                real mu = RL(c->alpha4) * phin * (phin * phin - RL(c->phi_star_sq)) - RL(c->kappa) * lap;
                real af = g_mut == ZO_MUT_INITF_AF_NO_HALF ? RL(c->gamma) * mu : RL(0.5) * RL(c->gamma) * mu;
                f[idx] = phin - RL(3.0) * RL(c->gamma) * mu;
                for (int i = 1; i < 7; ++i) f[i * n + idx] = af;
            }
}

void zo_init_g(int nx, int ny, int nz, const zo_coef *c, const real *phi,
               const real *rho, real *g)
{
    const size_t n = (size_t)nx * ny * nz;
    /* src/init_g.cu:46-93 */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
                real p[19];
                gather19(nx, ny, nz, phi, x, y, z, p);
                real phin = p[0];
                real rhon = rho[idx];
                real lap = lap19(p);
                real mu = RL(c->alpha4) * phin * (phin * phin - RL(c->phi_star_sq)) - RL(c->kappa) * lap;
                real ag1 = g_mut == ZO_MUT_INITG_AXIS_PHIMU ? W1 * (phin * mu + rhon)
                                                              : W1 * (RL(3.0) * phin * mu + rhon);
                real ag2 = W2 * (RL(3.0) * phin * mu + rhon);
                g[idx] = g_mut == ZO_MUT_INITG_REST_3PHIMU ? W0 * (rhon - RL(3.0) * phin * mu)
                                                           : W0 * (rhon - RL(6.0) * phin * mu);
                for (int i = 1; i < 7; ++i) g[i * n + idx] = ag1;
                for (int i = 7; i < 19; ++i) g[i * n + idx] = ag2;
            }
}

void zo_update_phi(int nx, int ny, int nz, const real *f, real *phi)
{
    const size_t n = (size_t)nx * ny * nz;
    /* src/update_phi.cu:40-44 */
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < n; ++idx)
        phi[idx] = f[idx] + f[n + idx] + f[2 * n + idx] + f[3 * n + idx] +
                   f[4 * n + idx] + f[5 * n + idx] + f[6 * n + idx];
}

void zo_update_rho(int nx, int ny, int nz, const real *g, real *rho)
{
    const size_t n = (size_t)nx * ny * nz;
    /* src/update_rho.cu:44-51: left-to-right sum g0 + ... + g18 */
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < n; ++idx) {
        real s = g[idx];
        for (int i = 1; i < 19; ++i) s = s + g[i * n + idx];
        rho[idx] = s;
    }
}

void zo_update_velocity(int nx, int ny, int nz, const real *g,
                        const real *rho, real *ux, real *uy, real *uz)
{
    const size_t n = (size_t)nx * ny * nz;
#define G(i) g[(size_t)(i) * n + idx]
    /* src/update_velocity.cu:45-61 */
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < n; ++idx) {
        real inv_rho = RL(1.0) / rho[idx];
        ux[idx] = (G(1) - G(2) + G(7) - G(8) + G(9) - G(10) + G(11) - G(12) +
                   G(13) - G(14)) * inv_rho;
        uy[idx] = (G(3) - G(4) + G(7) - G(8) - G(9) + G(10) + G(15) - G(16) +
                   G(17) - G(18)) * inv_rho;
        uz[idx] = (G(5) - G(6) + G(11) - G(12) - G(13) + G(14) + G(15) - G(16) -
                   G(17) + G(18)) * inv_rho;
    }
#undef G
}

/* Chemical potential and force of src/collision_f.cu:61-96. */
static inline void force(const zo_coef *c, int grav_on, const real p[19],
                         real *mu, real *fx, real *fy, real *fz)
{
    real phin = p[0];
    real lap = lap19(p);
    real m = RL(c->alpha4) * phin * (phin * phin - RL(c->phi_star_sq)) - RL(c->kappa) * lap;
    *mu = m;
    *fx = m * gx19(p) * INV12;
    *fy = m * gy19(p) * INV12;
    real z = m * gz19(p) * INV12;
    if (phin > RL(0.0)) z = z + (grav_on ? RL(c->grav) : RL(0.0));
    *fz = z;
}

void zo_collision_f(int nx, int ny, int nz, const zo_coef *c, int grav_on,
                    const real *phi, const real *rho, const real *ux,
                    const real *uy, const real *uz, const real *fa,
                    real *fb)
{
    const size_t n = (size_t)nx * ny * nz;
    /* src/collision_f.cu:44-116 */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
                real p[19], mu, fx, fy, fz;
                gather19(nx, ny, nz, phi, x, y, z, p);
                real phin = p[0];
                real inv_rho = RL(1.0) / rho[idx];
                force(c, grav_on, p, &mu, &fx, &fy, &fz);
                real uxn = ux[idx] + RL(0.5) * fx * inv_rho;
                real uyn = uy[idx] + RL(0.5) * fy * inv_rho;
                real uzn = uz[idx] + RL(0.5) * fz * inv_rho;
                real af = RL(0.5) * RL(c->gamma) * RL(c->inv_tau_phi) * mu;
                real cf = RL(c->inv_tau_phi) * RL(c->inv_eta2) * phin;
                real t = RL(c->inv_tau_phi_one);
                if (g_mut == ZO_MUT_AF_NO_TAU) af = RL(0.5) * RL(c->gamma) * mu;
                if (g_mut == ZO_MUT_CF_NO_TAU) cf = RL(c->inv_eta2) * phin;
                if (g_mut == ZO_MUT_CF_HALF) cf = RL(c->inv_tau_phi) * RL(0.5) * phin;
                fb[idx] = g_mut == ZO_MUT_F0_ONE_AF ? t * fa[idx] - af + RL(c->inv_tau_phi) * phin
                                                    : t * fa[idx] - RL(6.0) * af + RL(c->inv_tau_phi) * phin;
                fb[n + idx] = t * fa[n + idx] + af + cf * uxn;
                fb[2 * n + idx] = t * fa[2 * n + idx] + af - cf * uxn;
                fb[3 * n + idx] = t * fa[3 * n + idx] + af + cf * uyn;
                fb[4 * n + idx] = t * fa[4 * n + idx] + af - cf * uyn;
                fb[5 * n + idx] = t * fa[5 * n + idx] + af + cf * uzn;
                fb[6 * n + idx] = t * fa[6 * n + idx] + af - cf * uzn;
            }
}

void zo_collision_g(int nx, int ny, int nz, const zo_coef *c, int grav_on,
                    const real *phi, const real *rho, const real *ux,
                    const real *uy, const real *uz, const real *ga,
                    real *gb)
{
    const size_t n = (size_t)nx * ny * nz;
    const real K1 = RL(c->k1), K2 = RL(c->k2), KC1 = RL(c->kc1), KC2 = RL(c->kc2);
    const real t = RL(c->inv_tau_rho_one);
    /* src/collision_g.cu:49-196 */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
#define GA(i) ga[(size_t)(i) * n + idx]
#define GB(i) gb[(size_t)(i) * n + idx]
                real p[19], mu, fx, fy, fz;
                gather19(nx, ny, nz, phi, x, y, z, p);
                real phin = p[0];
                real rhon = rho[idx];
                force(c, grav_on, p, &mu, &fx, &fy, &fz);
                real uxn = ux[idx] + RL(0.5) * fx / rhon;
                real uyn = uy[idx] + RL(0.5) * fy / rhon;
                real uzn = uz[idx] + RL(0.5) * fz / rhon;
                real temp = K1 * (RL(3.0) * phin * mu +
                                    rhon * (RL(1.0) - RL(1.5) * (uxn * uxn + uyn * uyn + uzn * uzn)));
                real uf = uxn * fx + uyn * fy + uzn * fz;
                uxn = uxn * INV_CS_SQ;
                uyn = uyn * INV_CS_SQ;
                uzn = uzn * INV_CS_SQ;
                real eq1, eq2;

                GB(0) = t * GA(0) +
                        RL(c->k0) * (rhon - (g_mut == ZO_MUT_CG_REST_3PHIMU ? RL(3.0) : RL(6.0)) * phin * mu -
                                 rhon * (uxn * uxn + uyn * uyn + uzn * uzn) / RL(6.0)) -
                        RL(c->kc0) * uf;

                /* axis pairs: note the reference's K1*rhon*u*u*0.5 (x) versus
                 * K1*rhon*u*0.5*u (y, z) operand order, and eq2 = u*(K1 rho + KC1 F) */
                eq1 = temp + K1 * rhon * uxn * uxn * RL(0.5) + KC1 * (uxn * fx - uf);
                eq2 = uxn * (K1 * rhon + KC1 * fx);
                GB(1) = t * GA(1) + eq1 + eq2;
                GB(2) = t * GA(2) + eq1 - eq2;

                eq1 = temp + K1 * rhon * uyn * RL(0.5) * uyn + KC1 * (uyn * fy - uf);
                eq2 = uyn * (K1 * rhon + KC1 * fy);
                GB(3) = t * GA(3) + eq1 + eq2;
                GB(4) = t * GA(4) + eq1 - eq2;

                eq1 = temp + K1 * rhon * uzn * RL(0.5) * uzn + KC1 * (uzn * fz - uf);
                eq2 = uzn * (K1 * rhon + KC1 * fz);
                GB(5) = t * GA(5) + eq1 + eq2;
                GB(6) = t * GA(6) + eq1 - eq2;

                temp = K2 * ((g_mut == ZO_MUT_CG_DIAG_6PHIMU ? RL(6.0) : RL(3.0)) * phin * mu +
                             rhon * (RL(1.0) - (uxn * uxn + uyn * uyn + uzn * uzn) / RL(6.0)));

#define DIAG(ip, im, s, S)                                                   \
    eq1 = temp + K2 * rhon * (s) * RL(0.5) * (s) + KC2 * ((s) * (S) - uf);        \
    eq2 = K2 * rhon * (s) + KC2 * (S);                                        \
    GB(ip) = t * GA(ip) + eq1 + eq2;                                          \
    GB(im) = t * GA(im) + eq1 - eq2;

                DIAG(7, 8, (uxn + uyn), (fx + fy))
                DIAG(9, 10, (uxn - uyn), (fx - fy))
                DIAG(11, 12, (uxn + uzn), (fx + fz))
                DIAG(13, 14, (uxn - uzn), (fx - fz))
                DIAG(15, 16, (uyn + uzn), (fy + fz))
                DIAG(17, 18, (uyn - uzn), (fy - fz))
#undef DIAG
#undef GA
#undef GB
            }
}

void zo_stream_f(int nx, int ny, int nz, const zo_coef *c, const real *fb,
                 real *fa)
{
    const size_t n = (size_t)nx * ny * nz;
    /* src/stream_f.cu:52-59 restated as a pull:
     * fa_i(y) = eta * fb_i(y - c_i) + eta2 * fb_i(y), fa_0 = fb_0 */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
                fa[idx] = fb[idx];
                for (int i = 1; i < 7; ++i) {
                    size_t src = cell(nx, ny, wrap(x - CX[i], nx), wrap(y - CY[i], ny),
                                      wrap(z - CZ[i], nz));
                    fa[i * n + idx] = RL(c->eta) * fb[i * n + src] + RL(c->eta2) * fb[i * n + idx];
                }
            }
}

void zo_stream_g(int nx, int ny, int nz, const real *gb, real *ga)
{
    const size_t n = (size_t)nx * ny * nz;
    /* src/stream_g.cu:56-76 restated as a pull: ga_i(y) = gb_i(y - c_i) */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
                for (int i = 0; i < 19; ++i) {
                    size_t src = cell(nx, ny, wrap(x - CX[i], nx), wrap(y - CY[i], ny),
                                      wrap(z - CZ[i], nz));
                    ga[i * n + idx] = gb[i * n + src];
                }
            }
}

void zo_run(int nx, int ny, int nz, const zo_coef *c, int grav_on, int nsteps,
            real *f, real *g, real *work)
{
    const size_t n = (size_t)nx * ny * nz;
    real *fb = work, *gb = work + 7 * n;
    real *phi = gb + 19 * n, *rho = phi + n, *ux = rho + n, *uy = ux + n, *uz = uy + n;
    /* src/main.cu:498-566 with the halo exchanges folded into the wrap */
    for (int s = 0; s < nsteps; ++s) {
        zo_update_phi(nx, ny, nz, f, phi);
        zo_update_rho(nx, ny, nz, g, rho);
        zo_update_velocity(nx, ny, nz, g, rho, ux, uy, uz);
        zo_collision_f(nx, ny, nz, c, grav_on, phi, rho, ux, uy, uz, f, fb);
        zo_collision_g(nx, ny, nz, c, grav_on, phi, rho, ux, uy, uz, g, gb);
        zo_stream_f(nx, ny, nz, c, fb, f);
        zo_stream_g(nx, ny, nz, gb, g);
    }
}

/* Pressure of src/vtksave.c:75-119 (src/logsave.c:84-137 is the same
 * expression): alpha(phi^2(3phi^2 - 2phi*^2) - phi*^4)
 * - kappa(phi lap - |grad phi|^2/2) + Cs_sq rho, periodic stencil. */
static inline real pressure_at(const zo_coef *c, const real p[19], real rho)
{
    real phin = p[0], phin2 = phin * phin;
    real lap = lap19(p);
    real gx = gx19(p) * INV12, gy = gy19(p) * INV12, gz = gz19(p) * INV12;
    real gsq = gx * gx + gy * gy + gz * gz;
    return RL(c->alpha) * (phin2 * (RL(3.0) * phin2 - RL(2.0) * RL(c->phi_star_sq)) -
                       RL(c->phi_star_sq) * RL(c->phi_star_sq)) -
           RL(c->kappa) * (phin * lap - RL(0.5) * gsq) + CS_SQ * rho;
}

void zo_pressure(int nx, int ny, int nz, const zo_coef *c, const real *phi,
                 const real *rho, real *p_out)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
                real p[19];
                gather19(nx, ny, nz, phi, x, y, z, p);
                p_out[idx] = pressure_at(c, p, rho[idx]);
            }
}

void zo_stats(int nx, int ny, int nz, const zo_coef *c, const real *phi,
              const real *rho, const real *ux, const real *uy,
              const real *uz, double *out)
{
    double s_phi = 0, s_rho = 0, vol = 0, sx = 0, sy = 0, sz = 0, umax = 0;
    double p_in = 0, p_out = 0, vol_full = 0;
    /* Serial, fixed order: the oracle's sums are the FP64 yardstick. */
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
                s_phi += phi[idx];
                s_rho += rho[idx];
                if (phi[idx] >= RL(0.0)) vol_full += 1.0;
                /* src/stats.c:39-41 skip the first and last x and y rows */
                if (x < 1 || x > nx - 2 || y < 1 || y > ny - 2) continue;
                double usq = (double)ux[idx] * ux[idx] + (double)uy[idx] * uy[idx] +
                             (double)uz[idx] * uz[idx];
                if (usq > umax) umax = usq;
                real p[19];
                gather19(nx, ny, nz, phi, x, y, z, p);
                /* src/logsave.c:135-137 */
                double pres = pressure_at(c, p, rho[idx]);
                if (phi[idx] >= RL(0.0)) {
                    vol += 1.0;
                    sx += ux[idx];
                    sy += uy[idx];
                    sz += uz[idx];
                    p_in += pres;
                } else {
                    p_out += pres;
                }
            }
    out[0] = s_phi;
    out[1] = s_rho;
    out[2] = vol;
    out[3] = sx;
    out[4] = sy;
    out[5] = sz;
    out[6] = umax;
    out[7] = p_in;
    out[8] = p_out;
    out[9] = vol_full;
}
