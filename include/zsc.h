/*
 * zsc.h -- C-ABI of the MI355X-native Zheng-Shu-Chew (ZSC) bubbly-flow
 * lattice-Boltzmann step (libzsc.so).
 *
 * The reference (carlosrosales/lbm_bench_gpu) has no FFI: its hot path is the
 * set of __global__ kernels that main() launches in a fixed order
 * (src/main.cu:496-590).  Each of those kernels has an entry point here with
 * the same name (zsc_ prefix), the same argument meaning and the same
 * structure-of-arrays population layout, so a host driver can keep its call
 * sequence and swap the launches for these calls.  The differences are:
 *   - no neighbour-table arguments (nb_east/west/north/south): the periodic
 *     wrap is done inside the kernels;
 *   - no ghost planes: a box of nx*ny*nz real cells, N = nx*ny*nz, index
 *     idx = x + nx*(y + ny*z) (x fastest, src/gridid.c:32-35);
 *   - populations are passed as arrays of 7 (f) or 19 (g) device pointers,
 *     each to 2*N elements: the A half [0,N) is post-stream, the B half
 *     [N,2N) post-collision, exactly the reference's dcol convention
 *     (src/main.cu:225-251,315);
 *   - element type is double for ZSC_F64 contexts and float for ZSC_F32;
 *   - every call returns a status (ZSC_OK or a negative code) and never
 *     aborts; zsc_last_error() describes the last failure of this thread.
 *
 * Calls enqueue asynchronously on the context's HIP stream; zsc_sync() waits.
 * A context is not re-entrant: use it from one host thread at a time.
 *
 * Beyond the drop-in kernels the library owns a fast path: a context-held
 * lattice pair advanced by zsc_step(), one fused HIP kernel per time step,
 * with FP64, FP32 (g stored as it is or as deviations from its rest value),
 * or FP32-compute/16-bit-storage precisions.
 */
#ifndef ZSC_H
#define ZSC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define ZSC_OK 0
#define ZSC_ERR_ARG (-1)       /* invalid argument or size */
#define ZSC_ERR_IO (-2)        /* missing or malformed input file */
#define ZSC_ERR_HIP (-3)       /* HIP runtime error */
#define ZSC_ERR_ALLOC (-4)     /* device or host allocation failed */
#define ZSC_ERR_STATE (-5)     /* call out of order (e.g. step before init) */
#define ZSC_ERR_DIVERGED (-6)  /* NaN/Inf found in a conserved sum */
#define ZSC_ERR_PRECISION (-7) /* call not available at this precision */

/* ---- precisions -------------------------------------------------------- */
typedef enum zsc_precision {
    ZSC_F64 = 0,  /* FP64 storage and compute (parity mode) */
    ZSC_F32 = 1,  /* FP32 storage and compute */
    ZSC_BF16 = 2, /* FP32 compute, populations stored as bf16 deviations */
    ZSC_FP16 = 3, /* FP32 compute, populations stored as fp16 deviations */
    ZSC_F32D = 4  /* FP32 storage and compute, g stored as FP32 deviations from
                     w_i * rho_mean (same bytes as ZSC_F32; for large rho_mean) */
} zsc_precision;

/* ---- inputs: Input/properties.in and Input/discrete.in ----------------- */
/* Fields of properties.in in file order (src/paramread.c:40-59), plus two
 * optional overrides that may follow as "key value" pairs at the end of the
 * file ("phiStar 1.0", "gravity 1e-6"). */
typedef struct zsc_inputs {
    int max_step, relax_step, save, stat;
    int xmax, ymax, zmax;
    int block_x, block_y; /* accepted for compatibility; unused */
    double rho_l, rho_h;
    double tau_rho, tau_phi;
    double width, sigma, gamma;
    double eo;
    double phi_star; /* > 0 overrides 0.5*(rho_h - rho_l) */
    double gravity;  /* >= 0 overrides 0.25*Eo*sigma/((rho_h-rho_l)*R0^2) */
} zsc_inputs;

/* Derived scalars (src/paramread.c:141-170), the set the reference uploads
 * with 23 cudaMemcpyToSymbol calls plus grav_d (src/main.cu:315-342,479).
 * Fields may be edited before zsc_set_params, e.g. to treat kappa and
 * alpha (= beta) as independent parameters. */
typedef struct zsc_params {
    double phi_star, phi_star_sq, rho_mean; /* rho_mean = 0.5*(rho_h+rho_l) */
    double alpha, alpha4, kappa, gamma;
    double eta, eta2, inv_eta2;
    double inv_tau_phi, inv_tau_phi_one;
    double inv_tau_rho_one;
    double k0, k1, k2, kc0, kc1, kc2;
    double gravity, grav; /* grav = 2*phi_star*gravity (src/paramread.c:162) */
    double width, sigma;
} zsc_params;

/* Reads properties.in (src/paramread.c:38-61).  Returns ZSC_ERR_IO when the
 * file is missing or a token is malformed, ZSC_ERR_ARG for values the
 * reference rejects (src/main.cu:167-173: xmax, ymax not multiples of 32 are
 * accepted here, but non-positive sizes or tau <= 1/2 are not). */
int zsc_read_inputs(const char *properties_path, zsc_inputs *out);

/* Reads discrete.in (src/paramread.c:64-73) without the reference's
 * MAX_BUBBLES=20 cap: bubbles[4*i..] = xc, yc, zc, R for i < *n_out.
 * With bubbles == NULL only the count is returned. */
int zsc_read_bubbles(const char *discrete_path, double *bubbles, int capacity,
                     int *n_out);

/* Writes a discrete.in in the reference's format. */
int zsc_write_bubbles(const char *discrete_path, const double *bubbles, int n);

/* Parameter derivation of src/paramread.c:141-170 in FP64.  r0 is the radius
 * of the first bubble (src/paramread.c:161). */
int zsc_derive_params(const zsc_inputs *in, double r0, zsc_params *out);

/* Seeded bubble generator (no reference counterpart; the reference reads
 * bubbles from discrete.in).  SplitMix64 stream from `seed`; radius uniform in
 * [r_min, r_max]; centre uniform so that the sphere of radius R + width + 1
 * stays inside the box (x, y 0-based, z 1-based as src/init.c:51-53);
 * rejection until every pair of surfaces is at least `gap` apart.
 * Returns ZSC_ERR_ARG when max_tries draws do not place n bubbles. */
int zsc_generate_bubbles(uint64_t seed, int n, double r_min, double r_max,
                         double gap, double width, int nx, int ny, int nz,
                         int max_tries, double *bubbles);

/* Host initial order parameter of src/init.c:39-59 in FP64: phi = -phi_star,
 * then for each bubble in order, phi = phi_star*tanh(2(R_c - r)/W) where
 * r <= R_c + W (last bubble wins, z 1-based). */
int zsc_init_phi_host(int nx, int ny, int nz, const double *bubbles, int n,
                      double phi_star, double width, double *phi);

/* ---- context ----------------------------------------------------------- */
typedef struct zsc_ctx zsc_ctx;

/* Creates a context for an nx*ny*nz periodic box on the current HIP device.
 * stream: a hipStream_t, or NULL for the context's own stream.
 * Replaces the allocations and dimension symbols of src/main.cu:180-251,
 * 315-320 (the context allocates only its own fast-path lattices, lazily). */
int zsc_create(int nx, int ny, int nz, int precision, void *stream,
               zsc_ctx **out);
int zsc_destroy(zsc_ctx *ctx);

/* Replaces the cudaMemcpyToSymbol calls of src/main.cu:322-342 and grav_d. */
int zsc_set_params(zsc_ctx *ctx, const zsc_params *p);

/* Gravity switch of the relaxation (off, src/main.cu:361-362) and evolution
 * (on, src/main.cu:479) phases. */
int zsc_set_gravity(zsc_ctx *ctx, int on);

int zsc_sync(zsc_ctx *ctx);
const char *zsc_last_error(void);

/* ---- drop-in kernels (reference names, no nb tables, wrap in kernel) --- */
/* Each pointer of f[7]/g[19] addresses 2*N elements (A | B).  phi, rho, ux,
 * uy, uz address N elements.  All pointers are device pointers. */

/* src/init_f.cu:27-80: f_A = equilibrium at u = 0 from phi. */
int zsc_init_f(zsc_ctx *ctx, const void *phi, void *const f[7]);
/* src/init_g.cu:27-95: g_A = equilibrium at u = 0 from phi, rho. */
int zsc_init_g(zsc_ctx *ctx, const void *phi, const void *rho,
               void *const g[19]);
/* src/update_phi.cu:28-46: phi = sum_i f_A,i. */
int zsc_update_phi(zsc_ctx *ctx, void *phi, void *const f[7]);
/* src/update_rho.cu:28-53: rho = sum_i g_A,i. */
int zsc_update_rho(zsc_ctx *ctx, void *rho, void *const g[19]);
/* src/update_velocity.cu:28-63: u = sum_i c_i g_A,i / rho (g[0] unused). */
int zsc_update_velocity(zsc_ctx *ctx, const void *rho, void *ux, void *uy,
                        void *uz, void *const g[19]);
/* src/collision_f.cu:27-118: D3Q7 BGK + forcing, A -> B. */
int zsc_collision_f(zsc_ctx *ctx, const void *phi, const void *rho,
                    const void *ux, const void *uy, const void *uz,
                    void *const f[7]);
/* src/collision_g.cu:28-198: D3Q19 BGK + forcing, A -> B. */
int zsc_collision_g(zsc_ctx *ctx, const void *phi, const void *rho,
                    const void *ux, const void *uy, const void *uz,
                    void *const g[19]);
/* src/stream_f.cu:28-63 with the halo of src/pack_mpi_f.cu/unpack_mpi_f.cu
 * and src/pack_mpi.cu/unpack_mpi.cu folded into the wrap: B -> A with the
 * ZSC smoothing f_A,i(x) = eta f_B,i(x - c_i) + (1 - eta) f_B,i(x). */
int zsc_stream_f(zsc_ctx *ctx, void *const f[7]);
/* src/stream_g.cu:28-80 (+ halo as above): g_A,i(x) = g_B,i(x - c_i). */
int zsc_stream_g(zsc_ctx *ctx, void *const g[19]);

/* ---- fast path: context-owned lattices, one fused kernel per step ------ */
/* Initial state from a host FP64 order-parameter array (N values) and an
 * optional host density array (NULL: rho = rho_mean everywhere, as
 * src/init.c:45).  Runs the init_f/init_g equilibria on the device. */
int zsc_state_init(zsc_ctx *ctx, const double *phi_host,
                   const double *rho_host);
/* Same, with phi computed on the device from bubbles (src/init.c:39-59
 * semantics, FP64 tanh), for boxes too large for a host array. */
int zsc_state_init_bubbles(zsc_ctx *ctx, const double *bubbles, int n);
/* Advances the state by nsteps reference time steps (update -> collide ->
 * stream each, src/main.cu:496-590) with the current gravity switch.  The
 * fused kernel uses 32-bit byte offsets per population: up to 2^30 cells
 * (1024^3) at FP32/BF16/FP16 and 2^29 at FP64; larger boxes return
 * ZSC_ERR_ARG. */
int zsc_step(zsc_ctx *ctx, int nsteps);
/* Same, and returns the device time of the launches between two HIP events
 * on the context stream (synchronises). */
int zsc_step_timed(zsc_ctx *ctx, int nsteps, double *ms);
/* Allocates the two lattices up to `trials` (1..16) times, times four fused
 * steps on each pair (from a rest state, with the context's parameters and
 * launch shape) and keeps the fastest pair: the step's time depends on the
 * physical pages behind the lattices, which an allocation cannot choose.
 * ms (may be NULL) receives each trial's time per step, 0 for trials not run
 * (a further pair must leave a quarter of the free device memory).  Call it
 * after zsc_set_params and before the state is set (ZSC_ERR_STATE
 * otherwise); the state set afterwards is unaffected.  No reference
 * counterpart. */
int zsc_place_lattices(zsc_ctx *ctx, int trials, double *ms);
/* Post-stream (A) populations as FP64 host arrays, pop-major: f[i*N + idx],
 * g[i*N + idx].  Storage precision is decoded.  No reference counterpart
 * (the reference keeps f and g in its own device arrays); with
 * zsc_state_set, the state transfer of the fast path. */
int zsc_state_get(zsc_ctx *ctx, double *f_host, double *g_host);
/* Replaces the state by the given A populations (FP64 host, pop-major). */
int zsc_state_set(zsc_ctx *ctx, const double *f_host, const double *g_host);
/* Moments of the A state (phi, rho, u as src/update_*.cu) to FP64 host
 * arrays of N values; any pointer may be NULL. */
int zsc_state_moments(zsc_ctx *ctx, double *phi, double *rho, double *ux,
                      double *uy, double *uz);
/* Copies the A state into / from reference-style device arrays (A half of
 * each 2N array, the f/g layout of src/main.cu:225-251; F64, F32 and F32D
 * contexts, float arrays for the last two, F32D adding w_i * rho_mean to g
 * on export and taking it off on import): moves a state between the drop-in
 * kernels and the fused path.  No reference counterpart. */
int zsc_state_export(zsc_ctx *ctx, void *const f[7], void *const g[19]);
int zsc_state_import(zsc_ctx *ctx, void *const f[7], void *const g[19]);

/* Diagnostics of src/stats.c:39-91 and src/logsave.c:63-167, reduced on the
 * device in FP64 (wavefront shuffles, then a fixed-order block pass). */
typedef struct zsc_stats {
    double sum_phi, sum_rho;       /* all cells */
    double bubble_volume;          /* phi >= 0 cells, x,y in [1,n-2] */
    double ux_sum, uy_sum, uz_sum; /* over those bubble cells */
    double umax_sq;                /* max |u|^2 over x,y in [1,n-2] */
    double p_in_sum, p_out_sum;    /* src/logsave.c:135-143 pressure sums */
    double bubble_volume_full;     /* phi >= 0 cells, whole box */
} zsc_stats;
/* Reduces the moments of the A state.  Returns ZSC_ERR_DIVERGED when a sum
 * is NaN or Inf. */
int zsc_state_stats(zsc_ctx *ctx, zsc_stats *out);
/* The same diagnostics over reference-style device moment arrays (N
 * elements each at the context's element type, as zsc_update_phi, _rho and
 * _velocity leave them): the drop-in counterpart of src/stats.c:39-91 and
 * the sums of src/logsave.c:63-143, which the reference forms on the host
 * after copying phi and u back.  F64 and F32 contexts; needs parameters
 * (the pressure sums).  The context allocates an 82 KB buffer for the
 * partial sums on first use. */
int zsc_moment_stats(zsc_ctx *ctx, const void *phi, const void *rho, const void *ux,
                     const void *uy, const void *uz, zsc_stats *out);
/* zsc_moment_stats without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial sums live in a library-owned device array of 82 KB per GPU
 * ((1024 + 1) * 10 doubles, held until the process ends, its address looked
 * up once per device by zsc_create).  The reduction's grid is fixed, so
 * every box a context accepts fits; each call overwrites all of the array
 * it reads, so nothing needs clearing between calls.  Calls that overlap on
 * one GPU (several host threads, or contexts on several streams) share that
 * array and race: serialise them.  Results equal zsc_moment_stats' bit for
 * bit (same kernels, same order). */
int zsc_moment_stats_noalloc(zsc_ctx *ctx, const void *phi, const void *rho, const void *ux,
                             const void *uy, const void *uz, zsc_stats *out);

/* ---- output and restart ------------------------------------------------ */
/* Pressure of src/vtksave.c:75-119 (periodic 19-point stencil of phi, at the
 * context's compute precision) for the A state, to an FP64 host array. */
int zsc_state_pressure(zsc_ctx *ctx, double *p_host);
/* Legacy VTK STRUCTURED_POINTS file of the A state in the layout of
 * src/vtksave.c:29-141: scalars Phi and Pressure, then Vectors Velocity, x
 * fastest, ORIGIN 1 1 1.  binary = 0 writes ASCII "%f" values like the
 * reference; binary = 1 the legacy BINARY form (big-endian float32), for
 * boxes where text is impractical.  Replaces vtkSave's host loops. */
int zsc_vtk_save(zsc_ctx *ctx, const char *path, int binary);
/* Restart file (no reference counterpart): the lattice exactly as held (post-
 * stream or post-collision, storage precision), the box, precision,
 * parameters, gravity switch and a caller step counter.  zsc_state_load into
 * a context of the same box and precision restores all of them, so the run
 * continues bit for bit as if it had not stopped; it returns ZSC_ERR_ARG for
 * another box or precision and ZSC_ERR_IO for a damaged file. */
int zsc_state_save(zsc_ctx *ctx, const char *path, int64_t step);
int zsc_state_load(zsc_ctx *ctx, const char *path, int64_t *step);
/* Same, with up to ZSC_RESTART_USER_MAX caller values stored beside the
 * step counter (e.g. a driver's reference bubble volume, so that a resumed
 * run reports the same phase conservation as one that did not stop).
 * zsc_state_load_ex copies min(capacity, stored) values to user and the
 * stored count to *n_user (may be NULL).  Files without caller values load
 * with *n_user = 0.  Parameters in the file are validated as by
 * zsc_set_params (ZSC_ERR_IO otherwise, before the state is touched). */
#define ZSC_RESTART_USER_MAX 8
int zsc_state_save_ex(zsc_ctx *ctx, const char *path, int64_t step, const double *user,
                      int n_user);
int zsc_state_load_ex(zsc_ctx *ctx, const char *path, int64_t *step, double *user,
                      int capacity, int *n_user);

/* Launch shape of the fused kernel: ty wavefront rows of 64 cells per block
 * (1, 2, 4 or 8), zchunk planes marched per block (0: the whole box) and
 * flags: 1 non-temporal output stores, 2 (with 1) non-temporal loads, 4
 * prefetch the next plane's g populations (software pipelining: g_0..g_5 for
 * FP32 storage, all 19 for 16-bit storage), 8 no
 * XCD-contiguous tile order, 16 (FP32-compute precisions, ty 4 or 8, nx a
 * multiple of 128; ignored otherwise) two cells per lane, 32 (with 1 and
 * without 2; ignored for ty 1) keep f_5 / f_6, which consecutive planes
 * share, in LDS (one cell per lane) or registers (two cells per lane)
 * instead of loading them twice, and bits
 * 8..15 a number of KiB (<= 128) of unused LDS per block that limits how
 * many blocks share a CU.  Of bits 1-16, FP64 contexts use 1 and 32 only.  No reference counterpart; a
 * tuning knob of this implementation. */
int zsc_set_fused_config(zsc_ctx *ctx, int ty, int zchunk, int flags);
/* The launch shape in use (the per-precision default until set); any
 * pointer may be NULL. */
int zsc_get_fused_config(zsc_ctx *ctx, int *ty, int *zchunk, int *flags);

/* ---- adjoint (no reference counterpart: the reference has no derivatives) */
/* Hand-written reverse mode of the time step for F64, F32 and F32D
 * contexts.  All pointers are device pointers; lattices are plain pop-major
 * (26, N) arrays of the context's element type (f_0..f_6 then g_0..g_18,
 * population q at element q*N; float for F32D, whose g_i are stored as
 * g_i - w_i * rho_mean); phi arrays hold N elements; grads is a device
 * double[16] that the calls accumulate into (index: 0 alpha4, 1 kappa,
 * 2 inv_tau_phi_one, 3 inv_tau_phi, 4 inv_eta2, 5 eta, 6 inv_tau_rho_one,
 * 7-9 k0..k2, 10-12 kc0..kc2, 13 grav, 14 the loss).  Over F32D states the
 * sums 6-9 leave out the rest state's part (rho_mean times the lambda sums
 * weighted by w_i, resp. over each k's population group), which cancels in
 * every derivative through tau_g; that cancellation is what FP32 states lose
 * accuracy to, and F32D states keep tau_g's gradient within 1e-4.  The gravity switch applies as in
 * zsc_step.  A context's scratch buffer of zsc_adj_scratch_bytes() is passed
 * in by the caller, which owns all memory. */
size_t zsc_adj_scratch_bytes(zsc_ctx *ctx);
/* Storage of the adjoint lattices (lam arguments of zsc_adj_seed,
 * zsc_adj_backward, zsc_adj_init_backward): the context's own precision
 * (default; ZSC_F32 and ZSC_F32D both mean FP32 adjoint lattices over an
 * F32D context), or ZSC_F32 over an F64 context -- FP32 adjoint lattices over
 * FP64 states, with the per-cell reverse arithmetic and the coefficient
 * sums in FP64.  That mode keeps the FP64 forward's accuracy where FP32
 * states cannot: the tau_g derivative pairs sum_i lam_i g_i against
 * equilibria built from rho = sum_i g_i, whose FP32 rounding is the same in
 * every alike bulk cell and so adds up.  Set it before zsc_adj_scratch_bytes
 * is read; ZSC_ERR_PRECISION for other combinations. */
int zsc_adj_set_lambda_precision(zsc_ctx *ctx, int precision);
/* a0 = init_f/init_g equilibria of phi0 (src/init_f.cu, src/init_g.cu). */
int zsc_adj_init(zsc_ctx *ctx, const void *phi0, void *a0);
/* a_out = one time step of a_in (update, collide, stream). */
int zsc_adj_forward(zsc_ctx *ctx, const void *a_in, void *a_out, void *scratch);
/* Loss L = sum (phi(a_n) - phi_tgt)^2 into grads[14]; lam = dL/da_n. */
int zsc_adj_seed(zsc_ctx *ctx, const void *a_n, const void *phi_tgt, void *lam, void *scratch,
                 double *grads);
/* b = collision of a (post-collision lattice), phi = phi(a). */
int zsc_adj_collide(zsc_ctx *ctx, const void *a, void *b, void *phi);
/* b_out = one fused step from the post-collision lattice b_in (stream, then
 * update + collide); phi (may be NULL) receives phi of the post-stream state
 * in between. */
int zsc_adj_fused(zsc_ctx *ctx, const void *b_in, void *b_out, void *phi);
/* a = streaming of the post-collision lattice b. */
int zsc_adj_stream(zsc_ctx *ctx, const void *b, void *a);
/* lam_out = J(a_n)^T lam_next for the step a_n -> a_{n+1}; grads += the
 * step's coefficient derivatives.  state is a_n itself (pulled = 0; phi_n may
 * be NULL and is then computed) or the previous post-collision lattice b_{n-1}
 * from which a_n is streamed on the fly (pulled = 1; phi_n = phi(a_n)
 * required). */
int zsc_adj_backward(zsc_ctx *ctx, const void *state, int pulled, const void *phi_n,
                     const void *lam_next, void *lam_out, void *scratch, double *grads);
/* k backward steps in one call: lam_out = dL/da_{n0} from lam_next =
 * dL/da_{n0+k}, the steps n0+k-1 .. n0 in turn.  states[j] and phis[j] are
 * the state and phi of step n0+j as zsc_adj_backward takes them: pulled
 * (phi required), except states[0] when first_pulled == 0 (a_{n0} itself;
 * phis[0] may then be NULL).  lam_tmp is a third adjoint lattice (needed
 * for k >= 2); lam_next is not modified.  The result and the coefficient
 * sums (grads as zsc_adj_backward) equal those of k zsc_adj_backward calls
 * bit for bit, but every step except the last leaves the stencil part of
 * phi's adjoint (lap[a_lap] - div[a_grad], which the single call adds to
 * the seven f adjoints in a second pass) as one field in the scratch
 * buffer, and the next step adds it as it loads those adjoints: 20 instead
 * of 72 bytes per cell and step for that part. */
int zsc_adj_backward_steps(zsc_ctx *ctx, int k, const void *const *states,
                           const void *const *phis, int first_pulled, const void *lam_next,
                           void *lam_out, void *lam_tmp, void *scratch, double *grads);
/* With grads == NULL, zsc_adj_backward keeps the step's coefficient sums in
 * the scratch buffer; zsc_adj_flush adds everything kept since the last
 * flush (or zsc_adj_seed, which drops it) to grads (a no-op when nothing is
 * kept).  Saves one reduction launch per step.  The sums live in the scratch
 * buffer: keep passing the same one until the flush; zsc_adj_forward, which
 * reuses the buffer, fails with ZSC_ERR_STATE while sums are kept. */
int zsc_adj_flush(zsc_ctx *ctx, void *scratch, double *grads);
/* grads += derivatives of the initial equilibria (alpha4, kappa) given
 * lam0 = dL/da0. */
int zsc_adj_init_backward(zsc_ctx *ctx, const void *phi0, const void *lam0, void *scratch,
                          double *grads);

/* The practical read + write ceiling of the memory behind the context's two
 * lattices: device time (ms per copy, mean of five) of plain copies from one
 * lattice to the other, each moving 2 x the lattice bytes.  ms[0]: a linear
 * copy (each lane moves 64 bytes, loads before stores); ms[1]: the fused
 * step's access pattern without its neighbour pulls or arithmetic (tiles
 * marching z, the 26 populations of each cell; 0 when the x extent does not
 * suit the 16-bit two-cell pattern).  The step's time depends on the
 * physical pages behind the lattices, so this is measured on the same ones:
 * call it after zsc_set_params / zsc_place_lattices and before the state is
 * set (it overwrites the lattices; ZSC_ERR_STATE otherwise).  No reference
 * counterpart: roofline accounting. */
int zsc_copy_ceiling(zsc_ctx *ctx, double ms[2]);

/* Bytes per lattice update the fused step must move (26 populations read
 * and written once at the storage precision).  No reference counterpart:
 * roofline accounting. */
double zsc_bytes_per_cell(int precision);

#ifdef __cplusplus
}
#endif
#endif /* ZSC_H */
