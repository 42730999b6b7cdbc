// zsc_launch.h -- host-side launchers of libzsc's kernels, one explicit
// instantiation set per storage type (zsc_f64.hip: double; zsc_f32.hip:
// float, bf16s, f16s).  Internal to the library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/zsc.h"
#include "zsc_store.h"

namespace zsc {

struct Geo {
    int nx, ny, nz;
    int64_t n;
};

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte layout of one context lattice: f_0 | f_1..6 | g_0..18, packed.  For
// FP64 and FP32 this is exactly a pop-major (26, n) array, population q at
// element q*n, which the adjoint entry points exchange with callers.
template <class S>
struct LatLayout {
    using S0 = typename Traits<S>::S0;
    using SF = typename Traits<S>::SF;
    using SG = typename Traits<S>::SG;
    static size_t f_off(int64_t n) { return (size_t)n * sizeof(S0); }
    static size_t g_off(int64_t n) { return f_off(n) + (size_t)n * 6 * sizeof(SF); }
    static size_t bytes(int64_t n) { return g_off(n) + (size_t)n * 19 * sizeof(SG); }
    static Lattice<S> view(void *base, int64_t n)
    {
        char *b = (char *)base;
        return {(S0 *)b, (SF *)(b + f_off(n)), (SG *)(b + g_off(n))};
    }
};

// Launch configuration of the fused kernel (tunable).
enum : int {
    kFusedNtStore = 1,      // non-temporal stores of the output lattice
    kFusedNtLoad = 2,       // non-temporal loads (only together with kFusedNtStore)
    kFusedPrefetch = 4,     // load (part of) the next plane's g during this plane
    kFusedIdentityMap = 8,  // no XCD-contiguous tile remapping
    kFusedWide = 16,        // two cells per lane (nx % 128 == 0, ty 4 or 8; not FP64)
    kFusedZStash = 32,      // f_5 / f_6 shared by consecutive planes kept (LDS or registers)
};
struct FusedCfg {
    int ty;      // tile rows (waves per block); tile is 64 x ty cells
    int zchunk;  // planes marched per block (0: whole box)
    int flags;   // kFused* bits
};

template <class S>
struct Ops {
    using R = typename Traits<S>::R;
    // A <- equilibria from phi (R[n]) and rho (R[n] or nullptr: rho_mean).
    static hipError_t lat_init(const Geo &, const zsc_params &, hipStream_t, const void *phi,
                               const void *rho, void *A);
    // phi (R[n]) <- sum f of A.
    static hipError_t lat_phi(const Geo &, hipStream_t, const void *A, void *phi);
    // B <- collide(A) using the phi array for the stencil.
    static hipError_t lat_collide(const Geo &, const zsc_params &, int grav_on, hipStream_t,
                                  const void *A, const void *phi, void *B);
    // A <- stream(B).
    static hipError_t lat_stream(const Geo &, const zsc_params &, hipStream_t, const void *B,
                                 void *A);
    // Bout <- collide(stream(Bin)): one fused time step; phi (R[n], may be
    // nullptr) receives phi of the intermediate post-stream state.
    static hipError_t lat_fused(const Geo &, const zsc_params &, int grav_on, hipStream_t,
                                const FusedCfg &, const void *Bin, void *Bout, void *phi = nullptr);
    // moments of A into R arrays (any may be nullptr).
    static hipError_t lat_moments(const Geo &, const zsc_params &, hipStream_t, const void *A,
                                  void *phi, void *rho, void *ux, void *uy, void *uz);
    // population q (0..6 f, 7..25 g) of a lattice <-> FP64 device array.
    static hipError_t pop_get(const Geo &, const zsc_params &, hipStream_t, const void *L, int q,
                              double *dst);
    static hipError_t pop_set(const Geo &, const zsc_params &, hipStream_t, void *L, int q,
                              const double *src);
    // R array <-> FP64 device array.
    static hipError_t to_f64(hipStream_t, int64_t n, const void *src, double *dst);
    static hipError_t from_f64(hipStream_t, int64_t n, const double *src, void *dst);
    // phi (R[n]) from bubbles (device FP64 [nb][4]).
    static hipError_t init_phi_bubbles(const Geo &, hipStream_t, const double *bubbles, int nb,
                                       double phi_star, double width, void *phi);
    // rho_p (R[n]) <- pressure from phi and rho_p (the density on entry).
    static hipError_t pressure(const Geo &, const zsc_params &, hipStream_t, const void *phi,
                               void *rho_p);
    // FP64 reductions of the moment arrays; out (device) gets 10 doubles in
    // zsc_stats order, partial needs stats_partials() * 10 doubles.
    static hipError_t stats(const Geo &, const zsc_params &, hipStream_t, const void *phi,
                            const void *rho, const void *ux, const void *uy, const void *uz,
                            double *partial, double *out);
};

constexpr int kStatsBlocks = 1024;

// Reference-layout kernels (each pop pointer -> 2n elements, A | B).
template <class T>
struct DropIn {
    static hipError_t init_f(const Geo &, const zsc_params &, hipStream_t, const T *phi,
                             T *const f[7]);
    static hipError_t init_g(const Geo &, const zsc_params &, hipStream_t, const T *phi,
                             const T *rho, T *const g[19]);
    static hipError_t update_phi(const Geo &, hipStream_t, T *phi, T *const f[7]);
    static hipError_t update_rho(const Geo &, hipStream_t, T *rho, T *const g[19]);
    static hipError_t update_velocity(const Geo &, hipStream_t, const T *rho, T *ux, T *uy,
                                      T *uz, T *const g[19]);
    static hipError_t collision_f(const Geo &, const zsc_params &, int grav_on, hipStream_t,
                                  const T *phi, const T *rho, const T *ux, const T *uy,
                                  const T *uz, T *const f[7]);
    static hipError_t collision_g(const Geo &, const zsc_params &, int grav_on, hipStream_t,
                                  const T *phi, const T *rho, const T *ux, const T *uy,
                                  const T *uz, T *const g[19]);
    static hipError_t stream_f(const Geo &, const zsc_params &, hipStream_t, T *const f[7]);
    static hipError_t stream_g(const Geo &, hipStream_t, T *const g[19]);
    // A half of reference arrays <-> context lattice A; dev (non-null): the
    // lattice keeps g as deviations from w_i rho_mean of these parameters.
    static hipError_t export_state(const Geo &, hipStream_t, const void *A, T *const f[7],
                                   T *const g[19], const zsc_params *dev = nullptr);
    static hipError_t import_state(const Geo &, hipStream_t, void *A, T *const f[7],
                                   T *const g[19], const zsc_params *dev = nullptr);
};

// Reverse mode of the step on plain (26, n) lattices (zsc_adjoint.h): states
// and arithmetic at R, adjoint lattices stored at L.
template <class R, class L = R>
struct Adj;

}  // namespace zsc
