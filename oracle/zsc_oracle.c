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

/* Weights of src/constants.h:43-59, at FP64. */
#define INV6 (1.0 / 6.0)
#define INV12 (1.0 / 12.0)
#define INV_CS_SQ 3.0
#define CS_SQ (1.0 / 3.0)
#define W0 (1.0 / 3.0)
#define W1 (1.0 / 18.0)
#define W2 (1.0 / 36.0)
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
static inline void gather19(int nx, int ny, int nz, const double *phi, int x,
                            int y, int z, double p[19])
{
    for (int d = 0; d < 19; ++d)
        p[d] = phi[cell(nx, ny, wrap(x + CX[d], nx), wrap(y + CY[d], ny),
                        wrap(z + CZ[d], nz))];
}

/* Laplacian of src/collision_f.cu:61-70 (same in init_f/init_g/collision_g),
 * with phi(ie,jn,k)=p7, phi(iw,js,k)=p8, ... mapped to g directions. */
static inline double lap19(const double p[19])
{
    return (p[7] + p[8] + p[9] + p[10] + p[11] + p[12] + p[13] + p[14] + p[15] +
            p[16] + p[17] + p[18] +
            2.0 * (p[1] + p[2] + p[3] + p[4] + p[5] + p[6] - 12.0 * p[0])) *
           INV6;
}

/* Isotropic gradient bracket of src/collision_f.cu:78-94 (before mu*, *inv12). */
static inline double gx19(const double p[19])
{
    return 2.0 * (p[1] - p[2]) + p[7] - p[8] + p[9] - p[10] + p[11] - p[12] +
           p[13] - p[14];
}
static inline double gy19(const double p[19])
{
    return 2.0 * (p[3] - p[4]) + p[7] - p[8] + p[10] - p[9] + p[15] - p[16] +
           p[17] - p[18];
}
static inline double gz19(const double p[19])
{
    return 2.0 * (p[5] - p[6]) + p[11] - p[12] + p[14] - p[13] + p[15] - p[16] +
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

void zo_init_f(int nx, int ny, int nz, const zo_coef *c, const double *phi,
               double *f)
{
    const size_t n = (size_t)nx * ny * nz;
    /* src/init_f.cu:42-79 */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
                double p[19];
                gather19(nx, ny, nz, phi, x, y, z, p);
                double phin = p[0];
                double lap = lap19(p);
                double mu = c->alpha4 * phin * (phin * phin - c->phi_star_sq) - c->kappa * lap;
                double af = g_mut == ZO_MUT_INITF_AF_NO_HALF ? c->gamma * mu : 0.5 * c->gamma * mu;
                f[idx] = phin - 3.0 * c->gamma * mu;
                for (int i = 1; i < 7; ++i) f[i * n + idx] = af;
            }
}

void zo_init_g(int nx, int ny, int nz, const zo_coef *c, const double *phi,
               const double *rho, double *g)
{
    const size_t n = (size_t)nx * ny * nz;
    /* src/init_g.cu:46-93 */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
                double p[19];
                gather19(nx, ny, nz, phi, x, y, z, p);
                double phin = p[0];
                double rhon = rho[idx];
                double lap = lap19(p);
                double mu = c->alpha4 * phin * (phin * phin - c->phi_star_sq) - c->kappa * lap;
                double ag1 = g_mut == ZO_MUT_INITG_AXIS_PHIMU ? W1 * (phin * mu + rhon)
                                                              : W1 * (3.0 * phin * mu + rhon);
                double ag2 = W2 * (3.0 * phin * mu + rhon);
                g[idx] = g_mut == ZO_MUT_INITG_REST_3PHIMU ? W0 * (rhon - 3.0 * phin * mu)
                                                           : W0 * (rhon - 6.0 * phin * mu);
                for (int i = 1; i < 7; ++i) g[i * n + idx] = ag1;
                for (int i = 7; i < 19; ++i) g[i * n + idx] = ag2;
            }
}

void zo_update_phi(int nx, int ny, int nz, const double *f, double *phi)
{
    const size_t n = (size_t)nx * ny * nz;
    /* src/update_phi.cu:40-44 */
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < n; ++idx)
        phi[idx] = f[idx] + f[n + idx] + f[2 * n + idx] + f[3 * n + idx] +
                   f[4 * n + idx] + f[5 * n + idx] + f[6 * n + idx];
}

void zo_update_rho(int nx, int ny, int nz, const double *g, double *rho)
{
    const size_t n = (size_t)nx * ny * nz;
    /* src/update_rho.cu:44-51: left-to-right sum g0 + ... + g18 */
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < n; ++idx) {
        double s = g[idx];
        for (int i = 1; i < 19; ++i) s = s + g[i * n + idx];
        rho[idx] = s;
    }
}

void zo_update_velocity(int nx, int ny, int nz, const double *g,
                        const double *rho, double *ux, double *uy, double *uz)
{
    const size_t n = (size_t)nx * ny * nz;
#define G(i) g[(size_t)(i) * n + idx]
    /* src/update_velocity.cu:45-61 */
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < n; ++idx) {
        double inv_rho = 1.0 / rho[idx];
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
static inline void force(const zo_coef *c, int grav_on, const double p[19],
                         double *mu, double *fx, double *fy, double *fz)
{
    double phin = p[0];
    double lap = lap19(p);
    double m = c->alpha4 * phin * (phin * phin - c->phi_star_sq) - c->kappa * lap;
    *mu = m;
    *fx = m * gx19(p) * INV12;
    *fy = m * gy19(p) * INV12;
    double z = m * gz19(p) * INV12;
    if (phin > 0.0) z = z + (grav_on ? c->grav : 0.0);
    *fz = z;
}

void zo_collision_f(int nx, int ny, int nz, const zo_coef *c, int grav_on,
                    const double *phi, const double *rho, const double *ux,
                    const double *uy, const double *uz, const double *fa,
                    double *fb)
{
    const size_t n = (size_t)nx * ny * nz;
    /* src/collision_f.cu:44-116 */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
                double p[19], mu, fx, fy, fz;
                gather19(nx, ny, nz, phi, x, y, z, p);
                double phin = p[0];
                double inv_rho = 1.0 / rho[idx];
                force(c, grav_on, p, &mu, &fx, &fy, &fz);
                double uxn = ux[idx] + 0.5 * fx * inv_rho;
                double uyn = uy[idx] + 0.5 * fy * inv_rho;
                double uzn = uz[idx] + 0.5 * fz * inv_rho;
                double af = 0.5 * c->gamma * c->inv_tau_phi * mu;
                double cf = c->inv_tau_phi * c->inv_eta2 * phin;
                double t = c->inv_tau_phi_one;
                if (g_mut == ZO_MUT_AF_NO_TAU) af = 0.5 * c->gamma * mu;
                if (g_mut == ZO_MUT_CF_NO_TAU) cf = c->inv_eta2 * phin;
                if (g_mut == ZO_MUT_CF_HALF) cf = c->inv_tau_phi * 0.5 * phin;
                fb[idx] = g_mut == ZO_MUT_F0_ONE_AF ? t * fa[idx] - af + c->inv_tau_phi * phin
                                                    : t * fa[idx] - 6.0 * af + c->inv_tau_phi * phin;
                fb[n + idx] = t * fa[n + idx] + af + cf * uxn;
                fb[2 * n + idx] = t * fa[2 * n + idx] + af - cf * uxn;
                fb[3 * n + idx] = t * fa[3 * n + idx] + af + cf * uyn;
                fb[4 * n + idx] = t * fa[4 * n + idx] + af - cf * uyn;
                fb[5 * n + idx] = t * fa[5 * n + idx] + af + cf * uzn;
                fb[6 * n + idx] = t * fa[6 * n + idx] + af - cf * uzn;
            }
}

void zo_collision_g(int nx, int ny, int nz, const zo_coef *c, int grav_on,
                    const double *phi, const double *rho, const double *ux,
                    const double *uy, const double *uz, const double *ga,
                    double *gb)
{
    const size_t n = (size_t)nx * ny * nz;
    const double K1 = c->k1, K2 = c->k2, KC1 = c->kc1, KC2 = c->kc2;
    const double t = c->inv_tau_rho_one;
    /* src/collision_g.cu:49-196 */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
#define GA(i) ga[(size_t)(i) * n + idx]
#define GB(i) gb[(size_t)(i) * n + idx]
                double p[19], mu, fx, fy, fz;
                gather19(nx, ny, nz, phi, x, y, z, p);
                double phin = p[0];
                double rhon = rho[idx];
                force(c, grav_on, p, &mu, &fx, &fy, &fz);
                double uxn = ux[idx] + 0.5 * fx / rhon;
                double uyn = uy[idx] + 0.5 * fy / rhon;
                double uzn = uz[idx] + 0.5 * fz / rhon;
                double temp = K1 * (3.0 * phin * mu +
                                    rhon * (1.0 - 1.5 * (uxn * uxn + uyn * uyn + uzn * uzn)));
                double uf = uxn * fx + uyn * fy + uzn * fz;
                uxn = uxn * INV_CS_SQ;
                uyn = uyn * INV_CS_SQ;
                uzn = uzn * INV_CS_SQ;
                double eq1, eq2;

                GB(0) = t * GA(0) +
                        c->k0 * (rhon - (g_mut == ZO_MUT_CG_REST_3PHIMU ? 3.0 : 6.0) * phin * mu -
                                 rhon * (uxn * uxn + uyn * uyn + uzn * uzn) / 6.0) -
                        c->kc0 * uf;

                /* axis pairs: note the reference's K1*rhon*u*u*0.5 (x) versus
                 * K1*rhon*u*0.5*u (y, z) operand order, and eq2 = u*(K1 rho + KC1 F) */
                eq1 = temp + K1 * rhon * uxn * uxn * 0.5 + KC1 * (uxn * fx - uf);
                eq2 = uxn * (K1 * rhon + KC1 * fx);
                GB(1) = t * GA(1) + eq1 + eq2;
                GB(2) = t * GA(2) + eq1 - eq2;

                eq1 = temp + K1 * rhon * uyn * 0.5 * uyn + KC1 * (uyn * fy - uf);
                eq2 = uyn * (K1 * rhon + KC1 * fy);
                GB(3) = t * GA(3) + eq1 + eq2;
                GB(4) = t * GA(4) + eq1 - eq2;

                eq1 = temp + K1 * rhon * uzn * 0.5 * uzn + KC1 * (uzn * fz - uf);
                eq2 = uzn * (K1 * rhon + KC1 * fz);
                GB(5) = t * GA(5) + eq1 + eq2;
                GB(6) = t * GA(6) + eq1 - eq2;

                temp = K2 * ((g_mut == ZO_MUT_CG_DIAG_6PHIMU ? 6.0 : 3.0) * phin * mu +
                             rhon * (1.0 - (uxn * uxn + uyn * uyn + uzn * uzn) / 6.0));

#define DIAG(ip, im, s, S)                                                   \
    eq1 = temp + K2 * rhon * (s) * 0.5 * (s) + KC2 * ((s) * (S) - uf);        \
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

void zo_stream_f(int nx, int ny, int nz, const zo_coef *c, const double *fb,
                 double *fa)
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
                    fa[i * n + idx] = c->eta * fb[i * n + src] + c->eta2 * fb[i * n + idx];
                }
            }
}

void zo_stream_g(int nx, int ny, int nz, const double *gb, double *ga)
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
            double *f, double *g, double *work)
{
    const size_t n = (size_t)nx * ny * nz;
    double *fb = work, *gb = work + 7 * n;
    double *phi = gb + 19 * n, *rho = phi + n, *ux = rho + n, *uy = ux + n, *uz = uy + n;
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
static inline double pressure_at(const zo_coef *c, const double p[19], double rho)
{
    double phin = p[0], phin2 = phin * phin;
    double lap = lap19(p);
    double gx = gx19(p) * INV12, gy = gy19(p) * INV12, gz = gz19(p) * INV12;
    double gsq = gx * gx + gy * gy + gz * gz;
    return c->alpha * (phin2 * (3.0 * phin2 - 2.0 * c->phi_star_sq) -
                       c->phi_star_sq * c->phi_star_sq) -
           c->kappa * (phin * lap - 0.5 * gsq) + CS_SQ * rho;
}

void zo_pressure(int nx, int ny, int nz, const zo_coef *c, const double *phi,
                 const double *rho, double *p_out)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = cell(nx, ny, x, y, z);
                double p[19];
                gather19(nx, ny, nz, phi, x, y, z, p);
                p_out[idx] = pressure_at(c, p, rho[idx]);
            }
}

void zo_stats(int nx, int ny, int nz, const zo_coef *c, const double *phi,
              const double *rho, const double *ux, const double *uy,
              const double *uz, double *out)
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
                if (phi[idx] >= 0.0) vol_full += 1.0;
                /* src/stats.c:39-41 skip the first and last x and y rows */
                if (x < 1 || x > nx - 2 || y < 1 || y > ny - 2) continue;
                double usq = ux[idx] * ux[idx] + uy[idx] * uy[idx] + uz[idx] * uz[idx];
                if (usq > umax) umax = usq;
                double p[19];
                gather19(nx, ny, nz, phi, x, y, z, p);
                /* src/logsave.c:135-137 */
                double pres = pressure_at(c, p, rho[idx]);
                if (phi[idx] >= 0.0) {
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
