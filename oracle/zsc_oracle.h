/*
 * zsc_oracle.h -- FP64 CPU restatement of the Zheng-Shu-Chew (ZSC) bubbly-flow
 * lattice-Boltzmann step of carlosrosales/lbm_bench_gpu.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product (libzsc.so) never links or calls it.
 *
 * Built with -DZO_REAL=float (liboracle_f32.so), the step functions take
 * float arrays and evaluate the same expressions in FP32, the precision of
 * the reference's own kernels: the yardstick for what FP32 arithmetic of
 * the reference's algorithm loses against FP64.
 *
 * Layout: a fully periodic nx*ny*nz box without ghost planes, cell index
 * idx = x + nx*(y + ny*z) (x fastest, as src/gridid.c:32-35), populations
 * stored pop-major: pop i of cell idx lives at [i*N + idx], N = nx*ny*nz.
 * z is stored 0-based; the reference's real plane k=1..zmax is z = k-1.
 */
#ifndef ZSC_ORACLE_H
#define ZSC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element type of the lattices and moment fields: double, or float for the
 * FP32 build (-DZO_REAL=float, liboracle_f32.so). */
#ifndef ZO_REAL
#define ZO_REAL double
#endif
typedef ZO_REAL zo_real;

/* Scalars used by the step; names follow src/constants.h:62-85. */
typedef struct zo_coef {
    double phi_star, phi_star_sq, rho_mean;
    double alpha, alpha4, kappa, gamma;
    double eta, eta2, inv_eta2;
    double inv_tau_phi, inv_tau_phi_one;
    double inv_tau_rho_one;
    double k0, k1, k2, kc0, kc1, kc2;
    double gravity, grav;      /* grav = 2*phi_star*gravity, applied when on */
    double width, sigma;
} zo_coef;

/* Planted mistakes, for tests that show a pin is sensitive: each value
 * changes one term the way a plausible slip would.  ZO_MUT_NONE (the default)
 * leaves every function as the reference computes it.  The setting is
 * process-wide; tests set it and reset it to ZO_MUT_NONE. */
enum {
    ZO_MUT_NONE = 0,
    ZO_MUT_AF_NO_TAU = 1,          /* collision_f: A_f = Gamma mu / 2, 1/tau_phi dropped */
    ZO_MUT_CF_NO_TAU = 2,          /* collision_f: C_f = phi invEta2, 1/tau_phi dropped */
    ZO_MUT_CF_HALF = 3,            /* collision_f: C_f = phi / (2 tau_phi), invEta2 -> 1/2 */
    ZO_MUT_F0_ONE_AF = 4,          /* collision_f: f_B0 subtracts A_f once, not 6 A_f */
    ZO_MUT_INITG_REST_3PHIMU = 5,  /* init_g: g_0 = W0 (rho - 3 phi mu) */
    ZO_MUT_INITG_AXIS_PHIMU = 6,   /* init_g: g_1..6 = W1 (phi mu + rho) */
    ZO_MUT_INITF_AF_NO_HALF = 7,   /* init_f: f_1..6 = Gamma mu */
    ZO_MUT_CG_REST_3PHIMU = 8,     /* collision_g: K0 (rho - 3 phi mu - ...) */
    ZO_MUT_CG_DIAG_6PHIMU = 9      /* collision_g: diagonal K2 (6 phi mu + ...) */
};
void zo_set_mutation(int m);

/* Parameter derivation of src/paramread.c:141-170 in FP64.  r0 is the radius
 * of bubble 0 (src/paramread.c:161).  phi_star_override > 0 replaces
 * 0.5*(rho_h-rho_l); gravity_override >= 0 replaces the Eo-based gravity. */
void zo_derive(double rho_l, double rho_h, double tau_rho, double tau_phi,
               double width, double sigma, double gamma, double eo, double r0,
               double phi_star_override, double gravity_override,
               zo_coef *c);

/* Initial order parameter of src/init.c:39-59 (tanh bubbles, last one wins,
 * z 1-based).  bubbles is [nb][4] = xc, yc, zc, R. */
void zo_init_phi(int nx, int ny, int nz, const double *bubbles, int nb,
                 double phi_star, double width, double *phi);

void zo_init_f(int nx, int ny, int nz, const zo_coef *c, const zo_real *phi,
               zo_real *f);
void zo_init_g(int nx, int ny, int nz, const zo_coef *c, const zo_real *phi,
               const zo_real *rho, zo_real *g);

void zo_update_phi(int nx, int ny, int nz, const zo_real *f, zo_real *phi);
void zo_update_rho(int nx, int ny, int nz, const zo_real *g, zo_real *rho);
void zo_update_velocity(int nx, int ny, int nz, const zo_real *g,
                        const zo_real *rho, zo_real *ux, zo_real *uy, zo_real *uz);

/* A -> B collisions; grav_on selects grav (evolution) or 0 (relaxation). */
void zo_collision_f(int nx, int ny, int nz, const zo_coef *c, int grav_on,
                    const zo_real *phi, const zo_real *rho, const zo_real *ux,
                    const zo_real *uy, const zo_real *uz, const zo_real *fa,
                    zo_real *fb);
void zo_collision_g(int nx, int ny, int nz, const zo_coef *c, int grav_on,
                    const zo_real *phi, const zo_real *rho, const zo_real *ux,
                    const zo_real *uy, const zo_real *uz, const zo_real *ga,
                    zo_real *gb);

/* B -> A streaming with periodic wrap. */
void zo_stream_f(int nx, int ny, int nz, const zo_coef *c, const zo_real *fb,
                 zo_real *fa);
void zo_stream_g(int nx, int ny, int nz, const zo_real *gb, zo_real *ga);

/* nsteps full steps in the reference order (src/main.cu:496-590).  f, g hold
 * the A lattice on entry and exit; work must hold 7N + 19N + 5N elements. */
void zo_run(int nx, int ny, int nz, const zo_coef *c, int grav_on, int nsteps,
            zo_real *f, zo_real *g, zo_real *work);

/* Diagnostics (src/stats.c:39-55, src/logsave.c:63-147) over moments.
 * out[0]  sum phi (all cells)        out[1]  sum rho (all cells)
 * out[2]  bubble volume (phi>=0 count over i,j in [1,n-2], all z)
 * out[3..5] sum of ux,uy,uz over those bubble cells
 * out[6]  max |u|^2 over i,j in [1,n-2]
 * out[7]  pressure sum inside (phi>=0)   out[8] pressure sum outside
 * out[9]  phi>=0 count over the full box */
/* Pressure field of src/vtksave.c:75-119 from phi and rho (periodic). */
void zo_pressure(int nx, int ny, int nz, const zo_coef *c, const zo_real *phi,
                 const zo_real *rho, zo_real *p_out);

void zo_stats(int nx, int ny, int nz, const zo_coef *c, const zo_real *phi,
              const zo_real *rho, const zo_real *ux, const zo_real *uy,
              const zo_real *uz, double *out);

#ifdef __cplusplus
}
#endif
#endif
