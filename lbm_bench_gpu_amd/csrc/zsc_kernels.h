// zsc_kernels.h -- HIP kernels of libzsc (gfx950), templated on the
// population storage type.  Included by exactly one translation unit per
// precision family, which instantiates Ops<S> / DropIn<T>.
//
// Layout: fully periodic nx*ny*nz box, idx = x + nx*(y + ny*z), x fastest
// (src/gridid.c:32-35), structure-of-arrays populations.  Every kernel wraps
// neighbour indices itself; this replaces the neighbour tables of
// src/init.c:62-82 and the z ghost planes + halo exchange of src/main.cu:515-589.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>

#include "zsc_cell.h"
#include "zsc_launch.h"
#include "zsc_store.h"

#ifndef ZSC_OPAQUE_STRIDE
#define ZSC_OPAQUE_STRIDE 1
#endif

namespace zsc {

template <class R>
inline Coef<R> make_coef(const zsc_params &p, int grav_on)
{
    Coef<R> c;
    c.phi_star_sq = (R)p.phi_star_sq;
    c.alpha4 = (R)p.alpha4;
    c.kappa = (R)p.kappa;
    c.gamma = (R)p.gamma;
    c.eta = (R)p.eta;
    c.eta2 = (R)p.eta2;
    c.inv_eta2 = (R)p.inv_eta2;
    c.inv_tau_phi = (R)p.inv_tau_phi;
    c.inv_tau_phi_one = (R)p.inv_tau_phi_one;
    c.inv_tau_rho_one = (R)p.inv_tau_rho_one;
    c.k0 = (R)p.k0;
    c.k1 = (R)p.k1;
    c.k2 = (R)p.k2;
    c.kc0 = (R)p.kc0;
    c.kc1 = (R)p.kc1;
    c.kc2 = (R)p.kc2;
    c.grav = grav_on ? (R)p.grav : R(0);
    c.rho_mean = (R)p.rho_mean;
    if constexpr (sizeof(R) == 4) {
        // Rounding each FP32 coefficient on its own leaves t + 1/tau,
        // eta + eta2 and t_g + K0 + 6 K1 + 12 K2 a few ulps off one: a
        // systematic per-step drift of sum(phi) and sum(rho).  Form them so
        // the sums are exact: 1 - a is exact for a in [1/2, 2] (Sterbenz);
        // K2 keeps 22 significant bits so that 24 K2 is exact, and
        // K0 = 1/tau - 24 K2 is exact again.  (FP64 keeps the host values.)
        c.inv_tau_phi_one = R(1) - c.inv_tau_phi;
        c.eta2 = R(1) - c.eta;
        const float a = (float)(1.0 - p.inv_tau_rho_one);
        c.inv_tau_rho_one = 1.0f - a;
        float k2 = (float)p.k2;
        uint32_t u;
        __builtin_memcpy(&u, &k2, 4);
        u = (u + 2u) & ~3u;  // round to 22 significant bits
        __builtin_memcpy(&k2, &u, 4);
        c.k2 = k2;
        c.k1 = 2.0f * k2;
        c.k0 = a - 24.0f * k2;
    }
    return c;
}

__device__ inline int wrap_dec(int v, int n) { return v == 0 ? n - 1 : v - 1; }
__device__ inline int wrap_inc(int v, int n) { return v == n - 1 ? 0 : v + 1; }

// Neighbour offsets of one cell: cols[1+dx] + rows[1+dy] + planes[1+dz] is
// the linear index of (x+dx, y+dy, z+dz), periodic.
struct Nbr {
    int64_t cols[3], rows[3], planes[3];
    __device__ Nbr(int x, int y, int z, int nx, int ny, int nz)
    {
        const int64_t nxy = (int64_t)nx * ny;
        cols[0] = wrap_dec(x, nx);
        cols[1] = x;
        cols[2] = wrap_inc(x, nx);
        rows[0] = (int64_t)wrap_dec(y, ny) * nx;
        rows[1] = (int64_t)y * nx;
        rows[2] = (int64_t)wrap_inc(y, ny) * nx;
        planes[0] = (int64_t)wrap_dec(z, nz) * nxy;
        planes[1] = (int64_t)z * nxy;
        planes[2] = (int64_t)wrap_inc(z, nz) * nxy;
    }
    // index of x + (dx, dy, dz)
    __device__ int64_t at(int dx, int dy, int dz) const
    {
        return cols[1 + dx] + rows[1 + dy] + planes[1 + dz];
    }
};

template <class R>
__device__ inline void gather19(const R *__restrict__ phi, const Nbr &nb, R *p)
{
#pragma unroll
    for (int d = 0; d < 19; ++d) p[d] = phi[nb.at(cx(d), cy(d), cz(d))];
}

__device__ inline void cell_xyz(int64_t idx, int nx, int ny, int &x, int &y, int &z)
{
    const int64_t t = idx / nx;
    x = (int)(idx - t * nx);
    y = (int)(t % ny);
    z = (int)(t / ny);
}

// Same for cell indices below 2^32 (the fused and adjoint kernels' range):
// 32-bit divisions instead of 64-bit ones.
__device__ inline void cell_xyz32(uint32_t idx, int nx, int ny, int &x, int &y, int &z)
{
    const uint32_t t = idx / (uint32_t)nx;
    x = (int)(idx - t * (uint32_t)nx);
    const uint32_t zz = t / (uint32_t)ny;
    y = (int)(t - zz * (uint32_t)ny);
    z = (int)zz;
}

#define ZSC_GRID_STRIDE(idx, n) \
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (n); \
         idx += (int64_t)gridDim.x * blockDim.x)

inline dim3 grid_for(int64_t n, int block = 256)
{
    int64_t b = (n + block - 1) / block;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

// ===========================================================================
// Context-lattice kernels
// ===========================================================================

// Memory-access flags of the fused kernel: streaming (non-temporal) hints.
enum : int { kNtStore = 1, kNtLoad = 2, kZStash = 4 };

template <int FL, class S>
__device__ inline S ld_raw(const S *p)
{
    if constexpr ((FL & kNtLoad) != 0) {
        if constexpr (std::is_class<S>::value) {  // bf16s, f16s, f32d: one scalar member
            S v;
            v.v = __builtin_nontemporal_load(&p->v);
            return v;
        } else {
            return __builtin_nontemporal_load(p);
        }
    } else {
        return *p;
    }
}

template <int FL, class S>
__device__ inline void st_raw(S *p, S v)
{
    if constexpr ((FL & kNtStore) != 0) {
        if constexpr (std::is_class<S>::value)
            __builtin_nontemporal_store(v.v, &p->v);
        else
            __builtin_nontemporal_store(v, p);
    } else {
        *p = v;
    }
}

template <int FL = 0, class S, class R>
__device__ inline R ldg(const S *p, int64_t i, R rest)
{
    return Codec<S>::dec(ld_raw<FL>(p + i), rest);
}
template <int FL = 0, class S, class R>
__device__ inline void stg(S *p, int64_t i, R v, R rest)
{
    st_raw<FL>(p + i, Codec<S>::enc(v, rest));
}

template <class R>
__device__ inline R g_rest(const Coef<R> &c, int i)
{
    return g_weight<R>(i) * c.rho_mean;
}

// Decoded A populations of one cell of a lattice.
template <class S, class R>
__device__ inline void load_local(const Lattice<S> &L, int64_t n, int64_t idx, const Coef<R> &c,
                                  R *f, R *g)
{
    f[0] = ldg(L.f0, idx, R(0));
#pragma unroll
    for (int i = 1; i < 7; ++i) f[i] = ldg(L.f, (int64_t)(i - 1) * n + idx, R(0));
#pragma unroll
    for (int i = 0; i < 19; ++i) g[i] = ldg(L.g, (int64_t)i * n + idx, g_rest(c, i));
}

template <class S, class R, int FL = 0>
__device__ inline void store_local(const Lattice<S> &L, int64_t n, int64_t idx, const Coef<R> &c,
                                   const R *f, const R *g)
{
    stg<FL>(L.f0, idx, f[0], R(0));
#pragma unroll
    for (int i = 1; i < 7; ++i) stg<FL>(L.f, (int64_t)(i - 1) * n + idx, f[i], R(0));
#pragma unroll
    for (int i = 0; i < 19; ++i) stg<FL>(L.g, (int64_t)i * n + idx, g[i], g_rest(c, i));
}

// load_local / store_local with g kept as stored deviations from w_i
// rho_mean (deviation codecs), never added to and taken off the rest value.
template <class S, class R>
__device__ inline void load_local_dev(const Lattice<S> &L, int64_t n, int64_t idx, R *f, R *dg)
{
    f[0] = ldg(L.f0, idx, R(0));
#pragma unroll
    for (int i = 1; i < 7; ++i) f[i] = ldg(L.f, (int64_t)(i - 1) * n + idx, R(0));
#pragma unroll
    for (int i = 0; i < 19; ++i) dg[i] = Codec<typename Traits<S>::SG>::dec0(L.g[(int64_t)i * n + idx]);
}

template <class S, class R>
__device__ inline void store_local_dev(const Lattice<S> &L, int64_t n, int64_t idx, const R *f,
                                       const R *dg)
{
    stg(L.f0, idx, f[0], R(0));
#pragma unroll
    for (int i = 1; i < 7; ++i) stg(L.f, (int64_t)(i - 1) * n + idx, f[i], R(0));
#pragma unroll
    for (int i = 0; i < 19; ++i) L.g[(int64_t)i * n + idx] = Codec<typename Traits<S>::SG>::enc0(dg[i]);
}

// Post-stream f of one cell, pulled from a B lattice (src/stream_f.cu:52-59
// as a pull): f_A,i(y) = eta f_B,i(y - c_i) + eta2 f_B,i(y).
template <class S, class R, int FL = 0>
__device__ inline void pull_f(const Lattice<S> &B, int64_t n, const Nbr &nb, const Coef<R> &c,
                              R *f)
{
    const int64_t own = nb.at(0, 0, 0);
    f[0] = ldg<FL>(B.f0, own, R(0));
#pragma unroll
    for (int i = 1; i < 7; ++i) {
        const auto *p = B.f + (int64_t)(i - 1) * n;
        f[i] = c.eta * ldg<FL>(p, nb.at(-cx(i), -cy(i), -cz(i)), R(0)) + c.eta2 * ldg<FL>(p, own, R(0));
    }
}

// Post-stream g of one cell (src/stream_g.cu:56-76 as a pull).
template <class S, class R, int FL = 0>
__device__ inline void pull_g(const Lattice<S> &B, int64_t n, const Nbr &nb, const Coef<R> &c,
                              R *g)
{
#pragma unroll
    for (int i = 0; i < 19; ++i)
        g[i] = ldg<FL>(B.g + (int64_t)i * n, nb.at(-cx(i), -cy(i), -cz(i)), g_rest(c, i));
}

// ---- 32-bit-offset access for the fused kernel.  Every population array
// of a lattice is at most 4 GiB (checked at launch), so a byte offset within
// it fits 32 bits; with a uniform per-population base the loads and stores
// take the scalar-base + 32-bit vector-offset form.
struct Nbr32 {
    uint32_t cols[3], rows[3], planes[3];
    __device__ Nbr32(int x, int y, int z, int nx, int ny, int nz)
    {
        const uint32_t nxy = (uint32_t)nx * (uint32_t)ny;
        cols[0] = (uint32_t)wrap_dec(x, nx);
        cols[1] = (uint32_t)x;
        cols[2] = (uint32_t)wrap_inc(x, nx);
        rows[0] = (uint32_t)wrap_dec(y, ny) * (uint32_t)nx;
        rows[1] = (uint32_t)y * (uint32_t)nx;
        rows[2] = (uint32_t)wrap_inc(y, ny) * (uint32_t)nx;
        planes[0] = (uint32_t)wrap_dec(z, nz) * nxy;
        planes[1] = (uint32_t)z * nxy;
        planes[2] = (uint32_t)wrap_inc(z, nz) * nxy;
    }
    __device__ uint32_t at(int dx, int dy, int dz) const
    {
        return cols[1 + dx] + rows[1 + dy] + planes[1 + dz];
    }
    __device__ uint32_t off(int dx, int dy, int dz) const { return at(dx, dy, dz); }
    __device__ static constexpr int xstep(int) { return 0; }
};

// Neighbour offsets for a cell whose x-neighbours do not wrap.  off() gives
// the element offset of the (dy, dz) row, x step excluded; the x step is
// added to the pointer afterwards, where it folds into the instruction's
// immediate offset.  The nine row offsets are wrapped once per cell.
struct NbrRow {
    uint32_t o[3][3];
    __device__ NbrRow(int x, int y, int z, int nx, int ny, int nz)
    {
        const int32_t nxy = nx * ny;
        const uint32_t own = (uint32_t)x + (uint32_t)nx * ((uint32_t)y + (uint32_t)ny * (uint32_t)z);
        const int32_t ro[3] = {y == 0 ? (ny - 1) * nx : -nx, 0, y == ny - 1 ? -(ny - 1) * nx : nx};
        const int32_t po[3] = {z == 0 ? (nz - 1) * nxy : -nxy, 0, z == nz - 1 ? -(nz - 1) * nxy : nxy};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) o[a][b] = own + (uint32_t)(ro[a] + po[b]);
    }
    __device__ uint32_t off(int, int dy, int dz) const { return o[1 + dy][1 + dz]; }
    __device__ static constexpr int xstep(int dx) { return dx; }
    __device__ uint32_t at(int dx, int dy, int dz) const { return off(dx, dy, dz) + (uint32_t)dx; }
};

template <int FL, class S, class R>
__device__ inline void st32(S *base, uint32_t i, R v, R rest)
{
    S *p = (S *)((char *)base + i * (uint32_t)sizeof(S));
    st_raw<FL>(p, Codec<S>::enc(v, rest));
}

// zero rest value
template <int FL, class S>
__device__ inline auto ld32z(const S *base, uint32_t i)
{
    const S *p = (const S *)((const char *)base + i * (uint32_t)sizeof(S));
    return Codec<S>::dec0(ld_raw<FL>(p));
}

template <int FL, class S, class R>
__device__ inline void st32z(S *base, uint32_t i, R v)
{
    S *p = (S *)((char *)base + i * (uint32_t)sizeof(S));
    st_raw<FL>(p, Codec<S>::enc0(v));
}

// element (off + xs) of base, with xs added after the 32-bit offset
template <int FL, class S>
__device__ inline S ldx(const S *base, uint32_t off, int xs)
{
    const S *p = (const S *)((const char *)base + (uint64_t)(off * (uint32_t)sizeof(S))) + xs;
    return ld_raw<FL>(p);
}

template <class S, class R, int FL, class NB>
__device__ inline void pull_f32(const Lattice<S> &B, int64_t n, const NB &nb, const Coef<R> &c,
                                R *f)
{
    const uint32_t own = nb.at(0, 0, 0);
    f[0] = ld32z<FL>(B.f0, own);
#pragma unroll
    for (int i = 1; i < 7; ++i) {
        const auto *p = B.f + (int64_t)(i - 1) * n;
        using SF = typename Traits<S>::SF;
        const R up = Codec<SF>::dec0(
            ldx<FL>(p, nb.off(-cx(i), -cy(i), -cz(i)), NB::xstep(-cx(i))));
        f[i] = c.eta * up + c.eta2 * ld32z<FL>(p, own);
    }
}

// pull_f32 for a cell whose previous plane was pulled by the same thread:
// f_5's upwind value (plane z-1) is that plane's own f_5, and f_6's own value
// is that plane's upwind f_6 (plane z); both come from the thread's stash
// (s5, s6) when `use`, and the stash is refilled for the next plane.
template <class S, class R, int FL, class NB>
__device__ inline void pull_f32_zs(const Lattice<S> &B, int64_t n, const NB &nb, const Coef<R> &c,
                                   R *f, R &s5, R &s6, bool use)
{
    using SF = typename Traits<S>::SF;
    const uint32_t own = nb.at(0, 0, 0);
    f[0] = ld32z<FL>(B.f0, own);
#pragma unroll
    for (int i = 1; i < 5; ++i) {
        const auto *p = B.f + (int64_t)(i - 1) * n;
        const R up = Codec<SF>::dec0(
            ldx<FL>(p, nb.off(-cx(i), -cy(i), -cz(i)), NB::xstep(-cx(i))));
        f[i] = c.eta * up + c.eta2 * ld32z<FL>(p, own);
    }
    {
        const auto *p = B.f + (int64_t)4 * n;
        const R own5 = ld32z<FL>(p, own);
        const R up5 = use ? s5 : Codec<SF>::dec0(ldx<FL>(p, nb.off(0, 0, -1), 0));
        f[5] = c.eta * up5 + c.eta2 * own5;
        s5 = own5;
    }
    {
        const auto *p = B.f + (int64_t)5 * n;
        const R up6 = Codec<SF>::dec0(ldx<FL>(p, nb.off(0, 0, 1), 0));
        const R own6 = use ? s6 : ld32z<FL>(p, own);
        f[6] = c.eta * up6 + c.eta2 * own6;
        s6 = up6;
    }
}

// Pulled g of one cell; for deviation storage the values stay deviations
// (the fused kernel collides them in that form).
template <class S, class R, int FL, class NB, int LO = 0>
__device__ inline void pull_g32(const Lattice<S> &B, int64_t n, const NB &nb, const Coef<R> &c,
                                R *g)
{
    using SG = typename Traits<S>::SG;
#pragma unroll
    for (int i = LO; i < 19; ++i) {
        const SG v = ldx<FL>(B.g + (int64_t)i * n, nb.off(-cx(i), -cy(i), -cz(i)), NB::xstep(-cx(i)));
        if constexpr (Codec<SG>::deviation)
            g[i] = Codec<SG>::dec0(v);
        else
            g[i] = Codec<SG>::dec(v, g_rest(c, i));
    }
}

// The 19 pulled g of one cell as stored (no decode), so that a load issued
// a plane ahead is not waited for until its value is used.
template <class S, int FL, int HI = 19, class NB>
__device__ inline void pull_g32_raw(const Lattice<S> &B, int64_t n, const NB &nb,
                                    typename Traits<S>::SG *g)
{
#pragma unroll
    for (int i = 0; i < HI; ++i)
        g[i] = ldx<FL>(B.g + (int64_t)i * n, nb.off(-cx(i), -cy(i), -cz(i)), NB::xstep(-cx(i)));
}

template <class S, class R, int FL>
__device__ inline void store32(const Lattice<S> &L, int64_t n, uint32_t idx, const Coef<R> &c,
                               const R *f, const R *g)
{
    using SG = typename Traits<S>::SG;
    st32z<FL>(L.f0, idx, f[0]);
#pragma unroll
    for (int i = 1; i < 7; ++i) st32z<FL>(L.f + (int64_t)(i - 1) * n, idx, f[i]);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        if constexpr (Codec<SG>::deviation)
            st32z<FL>(L.g + (int64_t)i * n, idx, g[i]);  // g holds deviations
        else
            st32<FL>(L.g + (int64_t)i * n, idx, g[i], g_rest(c, i));
    }
}

template <class S>
__global__ void __launch_bounds__(256)
    k_lat_init(int nx, int ny, int nz, int64_t n, Coef<typename Traits<S>::R> c,
               const typename Traits<S>::R *__restrict__ phi,
               const typename Traits<S>::R *__restrict__ rho, Lattice<S> A)
{
    using R = typename Traits<S>::R;
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        Nbr nb(x, y, z, nx, ny, nz);
        R p[19], f[7], g[19];
        gather19(phi, nb, p);
        init_f_cell(c, p, f);
        if constexpr (Codec<typename Traits<S>::SG>::deviation) {
            init_g_cell_dev(c, p, rho ? rho[idx] - c.rho_mean : R(0), g);
            store_local_dev(A, n, idx, f, g);
        } else {
            init_g_cell(c, p, rho ? rho[idx] : c.rho_mean, g);
            store_local(A, n, idx, c, f, g);
        }
    }
}

template <class S>
__global__ void __launch_bounds__(256)
    k_lat_phi(int64_t n, Coef<typename Traits<S>::R> c, Lattice<S> A,
              typename Traits<S>::R *__restrict__ phi)
{
    using R = typename Traits<S>::R;
    ZSC_GRID_STRIDE(idx, n)
    {
        R f[7];
        f[0] = ldg(A.f0, idx, R(0));
#pragma unroll
        for (int i = 1; i < 7; ++i) f[i] = ldg(A.f, (int64_t)(i - 1) * n + idx, R(0));
        phi[idx] = sum_f(f);
    }
}

template <class S>
__global__ void __launch_bounds__(256)
    k_lat_collide(int nx, int ny, int nz, int64_t n, Coef<typename Traits<S>::R> c, Lattice<S> A,
                  const typename Traits<S>::R *__restrict__ phi, Lattice<S> B)
{
    using R = typename Traits<S>::R;
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        Nbr nb(x, y, z, nx, ny, nz);
        R p[19], fa[7], ga[19], fb[7], gb[19];
        gather19(phi, nb, p);
        if constexpr (Codec<typename Traits<S>::SG>::deviation) {
            load_local_dev(A, n, idx, fa, ga);
            collide_cell_dev(c, p, fa, ga, fb, gb);
            store_local_dev(B, n, idx, fb, gb);
        } else {
            load_local(A, n, idx, c, fa, ga);
            collide_cell(c, p, fa, ga, fb, gb);
            store_local(B, n, idx, c, fb, gb);
        }
    }
}

template <class S>
__global__ void __launch_bounds__(256)
    k_lat_stream(int nx, int ny, int nz, int64_t n, Coef<typename Traits<S>::R> c, Lattice<S> B,
                 Lattice<S> A)
{
    using R = typename Traits<S>::R;
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        Nbr nb(x, y, z, nx, ny, nz);
        R f[7], g[19];
        pull_f(B, n, nb, c, f);
        if constexpr (Codec<typename Traits<S>::SG>::deviation) {
            // g moves as stored
#pragma unroll
            for (int i = 0; i < 19; ++i)
                A.g[(int64_t)i * n + idx] = B.g[(int64_t)i * n + nb.at(-cx(i), -cy(i), -cz(i))];
            stg(A.f0, idx, f[0], R(0));
#pragma unroll
            for (int i = 1; i < 7; ++i) stg(A.f, (int64_t)(i - 1) * n + idx, f[i], R(0));
        } else {
            pull_g(B, n, nb, c, g);
            store_local(A, n, idx, c, f, g);
        }
    }
}

template <class S>
__global__ void __launch_bounds__(256)
    k_lat_moments(int64_t n, Coef<typename Traits<S>::R> c, Lattice<S> A,
                  typename Traits<S>::R *phi, typename Traits<S>::R *rho,
                  typename Traits<S>::R *ux, typename Traits<S>::R *uy,
                  typename Traits<S>::R *uz)
{
    using R = typename Traits<S>::R;
    ZSC_GRID_STRIDE(idx, n)
    {
        R f[7], g[19], r;
        if constexpr (Codec<typename Traits<S>::SG>::deviation) {
            // rho = rho_mean + sum of deviations; the rest state carries no
            // momentum (sum_i c_i w_i = 0), so u comes from the deviations
            load_local_dev(A, n, idx, f, g);
            r = c.rho_mean + sum_g(g);
        } else {
            load_local(A, n, idx, c, f, g);
            r = sum_g(g);
        }
        R vx, vy, vz;
        velocity(g, r, vx, vy, vz);
        if (phi) phi[idx] = sum_f(f);
        if (rho) rho[idx] = r;
        if (ux) ux[idx] = vx;
        if (uy) uy[idx] = vy;
        if (uz) uz[idx] = vz;
    }
}

template <class S0, class T, class R>
__global__ void __launch_bounds__(256)
    k_pop_get(int64_t n, const S0 *__restrict__ f0, const T *__restrict__ p, int is_f0, R rest,
              double *dst)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        if (is_f0)
            dst[idx] = (double)ldg(f0, idx, R(0));
        else if constexpr (Codec<T>::deviation)
            dst[idx] = (double)Codec<T>::dec0(p[idx]) + (double)rest;  // exact in FP64
        else
            dst[idx] = (double)ldg(p, idx, rest);
    }
}

template <class S0, class T, class R>
__global__ void __launch_bounds__(256)
    k_pop_set(int64_t n, S0 *__restrict__ f0, T *__restrict__ p, int is_f0, R rest,
              const double *src)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        if (is_f0)
            stg(f0, idx, (R)src[idx], R(0));
        else if constexpr (Codec<T>::deviation)
            p[idx] = Codec<T>::enc0((R)(src[idx] - (double)rest));  // deviation rounded once
        else
            stg(p, idx, (R)src[idx], rest);
    }
}

template <class R>
__global__ void __launch_bounds__(256) k_to_f64(int64_t n, const R *src, double *dst)
{
    ZSC_GRID_STRIDE(idx, n) dst[idx] = (double)src[idx];
}

// dst = src converted; src == nullptr fills zeros (a kernel, so that the
// clear can sit in a captured graph).
template <class R>
__global__ void __launch_bounds__(256) k_from_f64(int64_t n, const double *src, R *dst)
{
    ZSC_GRID_STRIDE(idx, n) dst[idx] = src ? (R)src[idx] : R(0);
}

// Initial phi of src/init.c:39-59 in FP64 (last bubble wins, z 1-based).
template <class R>
__global__ void __launch_bounds__(256)
    k_init_phi_bubbles(int nx, int ny, int64_t n, const double *__restrict__ bub, int nb,
                       double phi_star, double width, R *phi)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        double v = -phi_star;
        for (int m = 0; m < nb; ++m) {
            const double rc = bub[4 * m + 3];
            const double reach = rc + width;
            const double dx = (double)x - bub[4 * m + 0];
            if (dx > reach || dx < -reach) continue;
            const double dy = (double)y - bub[4 * m + 1];
            if (dy > reach || dy < -reach) continue;
            const double dz = (double)(z + 1) - bub[4 * m + 2];
            if (dz > reach || dz < -reach) continue;
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r <= reach) v = phi_star * tanh(2.0 * (rc - r) / width);
        }
        phi[idx] = (R)v;
    }
}

// ---------------------------------------------------------------------------
// The fused step: Bout = collide(stream(Bin)).
//
// A block owns a 64 x TY column tile (one wavefront per row, lanes along x)
// and marches a chunk of z planes.  For plane k+1 it pulls the 7 f
// populations of every tile cell and of a 1-cell ring around the tile,
// sums them to phi and keeps the plane in an LDS ring of 4 planes; the own
// cell's pulled f waits for the next iteration (in LDS for FP32 compute, in
// registers for FP64).  Plane k is then
// collided with its 19-point phi stencil read from LDS and 19 pulled g, and
// written once, coalesced, to Bout.  Four LDS slots make one barrier per plane
// sufficient: a wave writing slot (k+2)&3 can only run ahead of a wave that
// still reads slots k-1..k+1 by one iteration.
// ---------------------------------------------------------------------------
// Without the prefetch, the FP32-compute variants fit in 80 vector
// registers (6 waves per SIMD); the attribute keeps them there.
//
// Prefetch (PF): g populations 0..NP-1 of plane k+1 are loaded during plane
// k, issued before plane k+1's f so that both sets of loads are in flight
// together; the rest load in their own plane.  FP32 storage prefetches 6
// within a 96-register budget (2 blocks of 512 per CU, like the 117
// registers of a full prefetch, but faster: 5.55 against 5.70 ms per 512^3
// step); 16-bit storage prefetches all 19 as raw words.
template <class S>
constexpr int prefetch_pops()
{
    return sizeof(typename Traits<S>::SG) == 4 ? 6 : 19;
}
template <class S, bool PF>
constexpr int fused_waves()
{
    if constexpr (sizeof(S) == 8) return 1;
    if constexpr (!PF) return 6;
    return prefetch_pops<S>() < 19 ? 5 : 1;
}

template <class S, int TY, int FL, bool PF>
__global__ void __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(fused_waves<S, PF>())))
    k_fused(int nx, int ny, int nz, int64_t n, int zchunk, int nbx, int nby, int nblk, int remap,
            Coef<typename Traits<S>::R> c, Lattice<S> in, Lattice<S> out,
            typename Traits<S>::R *__restrict__ phi_out)
{
    using R = typename Traits<S>::R;
    __shared__ R sp[4][TY + 2][66];
    // FP32 compute: each cell's own pulled f of planes k and k+1 wait in LDS
    // (thread-private slots, no barrier) rather than in 14 vector registers
    // held across the z loop; this is what lets the kernel fit in 80.
    constexpr bool kStash = sizeof(R) == 4;
    __shared__ R fsh[kStash ? 2 : 1][kStash ? 7 : 1][kStash ? TY : 1][64];
    // f_5 / f_6 values shared by consecutive planes (thread-private slots)
    constexpr bool kZS = (FL & kZStash) != 0;
    constexpr int kRing = 2 * 64 + 2 * (TY + 2);
    __shared__ R zs[kZS ? 2 : 1][kZS ? TY : 1][64];
    __shared__ R zsr[kZS ? 2 : 1][kZS ? kRing : 1];

    // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous range of tiles so that neighbouring rows share its L2.
    const int b = blockIdx.x;
    int l = b;
    if (remap && (nblk & 7) == 0) l = (b & 7) * (nblk >> 3) + (b >> 3);
    const int bx = l % nbx;
    const int by = (l / nbx) % nby;
    const int bz = l / (nbx * nby);
    const int x0 = bx * 64, y0 = by * TY, z0 = bz * zchunk;
    const int z1 = min(nz, z0 + zchunk);
    const int tw = min(64, nx - x0), th = min(TY, ny - y0);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const bool own = tx < tw && ty < th;
    const int x = x0 + min(tx, tw - 1), y = y0 + min(ty, th - 1);
    const int nh = 2 * tw + 2 * (th + 2);  // ring cells per plane
    // block-uniform: the tile's x-neighbours need no wrap
    const bool xint = tw == 64 && x0 > 0 && x0 + 64 < nx;
    // the ring cells (x0-1 .. x0+64) pull from x0-2 .. x0+65
    const bool xint_ring = tw == 64 && x0 >= 2 && x0 + 66 <= nx;
    const int tid = ty * 64 + tx;

    // phi of plane zz (tile + ring) into LDS slot; own pulled f into fown.
    auto plane_phi = [&](int zz, R *fown, bool use) {
        const int zc = zz < 0 ? zz + nz : (zz >= nz ? zz - nz : zz);
        const int slot = (zz + 4) & 3;
        if (own) {
            if constexpr (kZS) {
                if (xint)
                    pull_f32_zs<S, R, FL>(in, n, NbrRow(x, y, zc, nx, ny, nz), c, fown, zs[0][ty][tx],
                                          zs[1][ty][tx], use);
                else
                    pull_f32_zs<S, R, FL>(in, n, Nbr32(x, y, zc, nx, ny, nz), c, fown, zs[0][ty][tx],
                                          zs[1][ty][tx], use);
            } else {
                if (xint)
                    pull_f32<S, R, FL>(in, n, NbrRow(x, y, zc, nx, ny, nz), c, fown);
                else
                    pull_f32<S, R, FL>(in, n, Nbr32(x, y, zc, nx, ny, nz), c, fown);
            }
            sp[slot][ty + 1][tx + 1] = sum_f(fown);
            if constexpr (kStash) {
#pragma unroll
                for (int i = 0; i < 7; ++i) fsh[(zz + 2) & 1][i][ty][tx] = fown[i];
            }
        }
        for (int h = tid; h < nh; h += 64 * TY) {
            int hx, hy, r, col;
            if (h < tw) {  // row below the tile
                col = h + 1;
                r = 0;
                hx = x0 + h;
                hy = y0 - 1;
            } else if (h < 2 * tw) {  // row above
                col = h - tw + 1;
                r = th + 1;
                hx = x0 + h - tw;
                hy = y0 + th;
            } else if (h < 2 * tw + th + 2) {  // column left, corners included
                col = 0;
                r = h - 2 * tw;
                hx = x0 - 1;
                hy = y0 - 1 + r;
            } else {  // column right
                col = tw + 1;
                r = h - 2 * tw - th - 2;
                hx = x0 + tw;
                hy = y0 - 1 + r;
            }
            hx = hx < 0 ? hx + nx : (hx >= nx ? hx - nx : hx);
            hy = hy < 0 ? hy + ny : (hy >= ny ? hy - ny : hy);
            R f[7];
            if constexpr (kZS) {
                if (xint_ring)
                    pull_f32_zs<S, R, FL>(in, n, NbrRow(hx, hy, zc, nx, ny, nz), c, f, zsr[0][h],
                                          zsr[1][h], use);
                else
                    pull_f32_zs<S, R, FL>(in, n, Nbr32(hx, hy, zc, nx, ny, nz), c, f, zsr[0][h],
                                          zsr[1][h], use);
            } else {
                if (xint_ring)
                    pull_f32<S, R, FL>(in, n, NbrRow(hx, hy, zc, nx, ny, nz), c, f);
                else
                    pull_f32<S, R, FL>(in, n, Nbr32(hx, hy, zc, nx, ny, nz), c, f);
            }
            sp[slot][r][col] = sum_f(f);
        }
    };

    R fcur[7], fnext[kStash ? 1 : 7];
    // Prefetched g of the next plane, held as stored: a decode right after
    // the loads (a shift or conversion for 16-bit storage) would make the
    // wave wait for them in the plane that issues them.
    using SG = typename Traits<S>::SG;
    constexpr int NP = PF ? prefetch_pops<S>() : 0;
    SG gnext[NP > 0 ? NP : 1];
    auto prefetch = [&](int zz) {
        if (xint)
            pull_g32_raw<S, FL, NP>(in, n, NbrRow(x, y, zz, nx, ny, nz), gnext);
        else
            pull_g32_raw<S, FL, NP>(in, n, Nbr32(x, y, zz, nx, ny, nz), gnext);
    };
    if constexpr (PF) {
        if (own) prefetch(z0);
    }
    // Wave priority: a wave issuing the plane's loads (raised) is preferred
    // by the CU's arbiter over waves in their collision arithmetic
    // (default), so loads leave early and more bytes are in flight.
    __builtin_amdgcn_s_setprio(2);
    plane_phi(z0 - 1, fcur, false);  // phi only; its own f is not collided
    plane_phi(z0, fcur, true);
    for (int k = z0; k < z1; ++k) {
#if ZSC_OPAQUE_STRIDE
        // Keep the 52 per-population base addresses out of loop-invariant
        // hoisting (they would exceed the scalar registers and spill); the
        // scalar unit recomputes them each plane for free.
        {
            uint32_t lo = (uint32_t)n;
            asm volatile("" : "+s"(lo));
            n = (int64_t)lo;
        }
#endif
        __builtin_amdgcn_s_setprio(2);
        R ga[19];
        const uint32_t idx = (uint32_t)x + (uint32_t)nx * ((uint32_t)y + (uint32_t)ny * (uint32_t)k);
        if constexpr (PF) {
            // software pipeline: plane k's first NP g arrived during the
            // previous iteration; load the rest of plane k, then issue plane
            // k+1's first NP, to land during this one
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                if constexpr (Codec<SG>::deviation)
                    ga[i] = Codec<SG>::dec0(gnext[i]);
                else
                    ga[i] = Codec<SG>::dec(gnext[i], g_rest(c, i));
            }
            if constexpr (NP < 19) {
                if (own) {
                    if (xint)
                        pull_g32<S, R, FL, NbrRow, NP>(in, n, NbrRow(x, y, k, nx, ny, nz), c, ga);
                    else
                        pull_g32<S, R, FL, Nbr32, NP>(in, n, Nbr32(x, y, k, nx, ny, nz), c, ga);
                }
            }
            if (own && k + 1 < z1) prefetch(k + 1);
        } else {
            if (own) {
                if (xint)
                    pull_g32<S, R, FL>(in, n, NbrRow(x, y, k, nx, ny, nz), c, ga);
                else
                    pull_g32<S, R, FL>(in, n, Nbr32(x, y, k, nx, ny, nz), c, ga);
            }
        }
        if constexpr (kStash) {
            R ftmp[7];
            plane_phi(k + 1, ftmp, true);
        } else {
            plane_phi(k + 1, fnext, true);
        }
        __syncthreads();
        __builtin_amdgcn_s_setprio(0);
        if (own) {
            if constexpr (kStash) {
#pragma unroll
                for (int i = 0; i < 7; ++i) fcur[i] = fsh[(k + 2) & 1][i][ty][tx];
            }
            const int sm = (k + 3) & 3, s0 = k & 3, spl = (k + 1) & 3;
            const int r = ty + 1, q = tx + 1;
            R p[19];
            p[0] = sp[s0][r][q];
            p[1] = sp[s0][r][q + 1];
            p[2] = sp[s0][r][q - 1];
            p[3] = sp[s0][r + 1][q];
            p[4] = sp[s0][r - 1][q];
            p[5] = sp[spl][r][q];
            p[6] = sp[sm][r][q];
            p[7] = sp[s0][r + 1][q + 1];
            p[8] = sp[s0][r - 1][q - 1];
            p[9] = sp[s0][r - 1][q + 1];
            p[10] = sp[s0][r + 1][q - 1];
            p[11] = sp[spl][r][q + 1];
            p[12] = sp[sm][r][q - 1];
            p[13] = sp[sm][r][q + 1];
            p[14] = sp[spl][r][q - 1];
            p[15] = sp[spl][r + 1][q];
            p[16] = sp[sm][r - 1][q];
            p[17] = sp[sm][r + 1][q];
            p[18] = sp[spl][r - 1][q];
            R fb[7], gb[19];
            if constexpr (Codec<typename Traits<S>::SG>::deviation)
                collide_cell_dev(c, p, fcur, ga, fb, gb);
            else
                collide_cell(c, p, fcur, ga, fb, gb);
            store32<S, R, FL>(out, n, idx, c, fb, gb);
            if (phi_out) phi_out[idx] = p[0];  // phi of the post-stream state
        }
        if constexpr (!kStash) {
#pragma unroll
            for (int i = 0; i < 7; ++i) fcur[i] = fnext[i];
        }
    }
}

// ---------------------------------------------------------------------------
// The fused step for 16-bit storage with two cells per lane.
//
// A wave covers 128 consecutive cells of a row, lane l the cells x0 + 2l and
// x0 + 2l + 1, so every population is read and written as aligned 32-bit
// words holding a lane's two values: each wave load or store moves 256 bytes,
// as in the FP32 kernel, instead of 128.  A pull across x is assembled in
// registers: the neighbouring lane's word comes through a DPP wave shift, the
// lane at the tile edge takes the neighbouring tile's word instead, and
// v_alignbit joins the two halves.  Tiles are 128 x TY cells (TY = 4 or 8)
// and need nx % 128 == 0.  The cell arithmetic is the one of k_fused.
// ---------------------------------------------------------------------------

typedef float f32x2 __attribute__((ext_vector_type(2)));

// 32-bit word at (even) element e of a 16-bit population array.
template <class T>
__device__ inline uint32_t ld_w(const T *base, uint32_t e)
{
    return *(const uint32_t *)((const char *)base + (uint64_t)(e * 2u));
}
template <int FL, class T>
__device__ inline void st_w(T *base, uint32_t e, uint32_t v)
{
    uint32_t *p = (uint32_t *)((char *)base + (uint64_t)(e * 2u));
    if constexpr ((FL & kNtStore) != 0)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// The word of a lane's two cells pulled from x - cx, given the word w of
// the same two cells' own x positions (in the source row) and the word
// `edge` of the neighbouring tile (x0 - 2 for cx = +1, x0 + 128 for cx = -1).
template <int CX>
__device__ inline uint32_t shift_word(uint32_t w, uint32_t edge)
{
    if constexpr (CX == 0) {
        return w;
    } else if constexpr (CX == 1) {
        // lane l - 1's word (the edge word in lane 0): wave_shr:1
        const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)w, 0x138, 0xF, 0xF, false);
        return __builtin_amdgcn_alignbit(w, prev, 16);
    } else {
        // lane l + 1's word (the edge word in lane 63): wave_shl:1
        const uint32_t next = (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)w, 0x130, 0xF, 0xF, false);
        return __builtin_amdgcn_alignbit(next, w, 16);
    }
}

template <class T>
__device__ inline T half_of(uint32_t w, int j)
{
    T v;
    v.v = (uint16_t)(j ? (w >> 16) : w);
    return v;
}

// A lane's two consecutive elements of a population array: one 32-bit word
// for 16-bit storage, a float pair for FP32 storage.
template <class T>
struct Pair {
    using type = uint32_t;
};
template <>
struct Pair<float> {
    using type = f32x2;
};
template <>
struct Pair<f32d> {
    using type = f32x2;
};

template <class T>
__device__ inline typename Pair<T>::type ld_pair(const T *base, uint32_t e)
{
    if constexpr (sizeof(T) == 2)
        return ld_w(base, e);
    else
        return *(const f32x2 *)((const char *)base + (uint64_t)(e * 4u));
}
template <int FL, class T>
__device__ inline void st_pair(T *base, uint32_t e, typename Pair<T>::type v)
{
    if constexpr (sizeof(T) == 2) {
        st_w<FL>(base, e, v);
    } else {
        f32x2 *p = (f32x2 *)((char *)base + (uint64_t)(e * 4u));
        if constexpr ((FL & kNtStore) != 0)
            __builtin_nontemporal_store(v, p);
        else
            *p = v;
    }
}

// The pair pulled from x - cx (see shift_word); for FP32 only one element
// crosses lanes.  (The elements are copied to scalars before the bit casts:
// a bit cast of a vector element expression reads the vector's first
// element.)
template <int CX>
__device__ inline f32x2 shift_word(f32x2 w, f32x2 edge)
{
    const float wx = w.x, wy = w.y;
    if constexpr (CX == 0) {
        return w;
    } else if constexpr (CX == 1) {
        const float ey = edge.y;
        const float prev = __builtin_bit_cast(
            float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, ey), __builtin_bit_cast(int, wy),
                                               0x138, 0xF, 0xF, false));
        f32x2 r;
        r.x = prev;
        r.y = wx;
        return r;
    } else {
        const float ex = edge.x;
        const float next = __builtin_bit_cast(
            float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, ex), __builtin_bit_cast(int, wx),
                                               0x130, 0xF, 0xF, false));
        f32x2 r;
        r.x = wy;
        r.y = next;
        return r;
    }
}

// Element j of a pair, as stored (16-bit: still encoded).
template <class T>
__device__ inline auto elem(typename Pair<T>::type w, int j)
{
    if constexpr (sizeof(T) == 2)
        return half_of<T>(w, j);
    else
        return j ? w.y : w.x;
}
// A cell's output encoded for storage, and two cells' outputs joined.
template <class T>
__device__ inline auto enc_out(float v)
{
    if constexpr (sizeof(T) == 2)
        return (uint32_t)Codec<T>::enc0(v).v;
    else
        return v;
}
template <class T, class E>
__device__ inline typename Pair<T>::type join(E a, E b)
{
    if constexpr (sizeof(T) == 2) {
        return a | (b << 16);
    } else {
        f32x2 r;
        r.x = a;
        r.y = b;
        return r;
    }
}

// Start elements (x = 0) of the 3 x 3 rows (y + dy, z + dz), periodic.
struct RowStarts {
    uint32_t r[3][3];
    __device__ RowStarts(int y, int z, int nx, int ny, int nz)
    {
        const int ys[3] = {wrap_dec(y, ny), y, wrap_inc(y, ny)};
        const int zs[3] = {wrap_dec(z, nz), z, wrap_inc(z, nz)};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
                r[a][b] = (uint32_t)nx * ((uint32_t)ys[a] + (uint32_t)ny * (uint32_t)zs[b]);
    }
    // row of the source x - c: (y - cy, z - cz)
    __device__ uint32_t src(int cy_, int cz_) const { return r[1 - cy_][1 - cz_]; }
};

// Pulled, decoded f of a lane's two cells (src/stream_f.cu:52-59 as a pull,
// the operand order of pull_f32): f[j][i] for the cell x0 + 2l + j.
template <class S, int FL>
__device__ inline void pull_f_w(const Lattice<S> &B, int64_t n, const RowStarts &rs, uint32_t xo,
                                uint32_t left, uint32_t right, const Coef<float> &c, float f[2][7],
                                typename Pair<typename Traits<S>::SF>::type *s5 = nullptr,
                                typename Pair<typename Traits<S>::SF>::type *s6 = nullptr,
                                bool use = false)
{
    using SF = typename Traits<S>::SF;
    const f32x2 f0 = *(const f32x2 *)((const char *)B.f0 + (uint64_t)((rs.r[1][1] + xo) * 4u));
    f[0][0] = f0.x;
    f[1][0] = f0.y;
#pragma unroll
    for (int i = 1; i < 7; ++i) {
        const SF *p = B.f + (int64_t)(i - 1) * n;
        const uint32_t row = rs.src(cy(i), cz(i));
        // with a stash (s5, s6), f_5's upwind word is the previous plane's own
        // word and f_6's own word the previous plane's upwind one (`use`)
        typename Pair<SF>::type own, up;
        if (s6 != nullptr && i == 6 && use)
            own = *s6;
        else
            own = ld_pair(p, rs.r[1][1] + xo);
        if (s5 != nullptr && i == 5 && use)
            up = *s5;
        else if (cx(i) == 0)
            up = ld_pair(p, row + xo);
        else if (cx(i) == 1)
            up = shift_word<1>(ld_pair(p, row + xo), ld_pair(p, row + left));
        else
            up = shift_word<-1>(ld_pair(p, row + xo), ld_pair(p, row + right));
        if (s5 != nullptr && i == 5) *s5 = own;
        if (s6 != nullptr && i == 6) *s6 = up;
#pragma unroll
        for (int j = 0; j < 2; ++j)
            f[j][i] = c.eta * Codec<SF>::dec0(elem<SF>(up, j)) +
                      c.eta2 * Codec<SF>::dec0(elem<SF>(own, j));
    }
}

// LDS holds 3 blocks of 4 rows (12 waves) or 2 of 8 rows (16 waves) per CU;
// the register budget matches.
template <class S, int TY, int FL>
__global__ void __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(TY == 4 ? 3 : 4)))
    k_fused_w(int nx, int ny, int nz, int64_t n, int zchunk, int nbx, int nby, int nblk, int remap,
              Coef<float> c, Lattice<S> in, Lattice<S> out, float *__restrict__ phi_out)
{
    using R = float;
    using SF = typename Traits<S>::SF;
    using SG = typename Traits<S>::SG;
    static_assert(sizeof(SG) == sizeof(SF) && TY >= 3, "FP32 or 16-bit storage, >= 3 rows");
    // phi ring: [slot][row][cell parity][lane + 1]; x0 - 1 is [1][0], x0 + 128 is [0][65]
    __shared__ R sp[4][TY + 2][2][66];
    // each lane's pulled f of planes k and k+1 (thread-private slots)
    __shared__ R fsh[2][7][TY][2][64];

    const int b = blockIdx.x;
    int l = b;
    if (remap && (nblk & 7) == 0) l = (b & 7) * (nblk >> 3) + (b >> 3);
    const int bx = l % nbx;
    const int by = (l / nbx) % nby;
    const int bz = l / (nbx * nby);
    const int x0 = bx * 128, y0 = by * TY, z0 = bz * zchunk;
    const int z1 = min(nz, z0 + zchunk);
    const int th = min(TY, ny - y0);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const bool own = ty < th;
    const int y = y0 + min(ty, th - 1);
    const uint32_t xo = (uint32_t)(x0 + 2 * tx);
    const uint32_t left = (uint32_t)(x0 == 0 ? nx - 2 : x0 - 2);
    const uint32_t right = (uint32_t)(x0 + 128 == nx ? 0 : x0 + 128);

    // f_5 / f_6 words shared by consecutive planes (own row, ring row)
    constexpr bool kZS = (FL & kZStash) != 0;
    using FW = typename Pair<SF>::type;
    FW zs5, zs6, zr5, zr6;

    auto plane_phi = [&](int zz, bool use) {
        const int zc = zz < 0 ? zz + nz : (zz >= nz ? zz - nz : zz);
        const int slot = (zz + 4) & 3;
        if (own) {
            float f[2][7];
            pull_f_w<S, FL>(in, n, RowStarts(y, zc, nx, ny, nz), xo, left, right, c, f,
                            kZS ? &zs5 : nullptr, kZS ? &zs6 : nullptr, use);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                sp[slot][ty + 1][j][tx + 1] = sum_f(f[j]);
#pragma unroll
                for (int i = 0; i < 7; ++i) fsh[(zz + 2) & 1][i][ty][j][tx] = f[j][i];
            }
        }
        // ring rows y0 - 1 (wave 0) and y0 + th (wave 1): whole rows
        if (ty < 2) {
            const int r = ty == 0 ? 0 : th + 1;
            int hy = ty == 0 ? y0 - 1 : y0 + th;
            hy = hy < 0 ? hy + ny : (hy >= ny ? hy - ny : hy);
            float f[2][7];
            pull_f_w<S, FL>(in, n, RowStarts(hy, zc, nx, ny, nz), xo, left, right, c, f,
                            kZS ? &zr5 : nullptr, kZS ? &zr6 : nullptr, use);
#pragma unroll
            for (int j = 0; j < 2; ++j) sp[slot][r][j][tx + 1] = sum_f(f[j]);
        }
        // ring columns x0 - 1 and x0 + 128, rows y0 - 1 .. y0 + th (wave 2)
        if (ty == 2 && tx < 2 * (th + 2)) {
            const int side = tx / (th + 2), r = tx % (th + 2);
            int hx = side == 0 ? x0 - 1 : x0 + 128;
            hx = hx < 0 ? hx + nx : (hx >= nx ? hx - nx : hx);
            int hy = y0 - 1 + r;
            hy = hy < 0 ? hy + ny : (hy >= ny ? hy - ny : hy);
            R f[7];
            pull_f32<S, R, FL>(in, n, Nbr32(hx, hy, zc, nx, ny, nz), c, f);
            if (side == 0)
                sp[slot][r][1][0] = sum_f(f);
            else
                sp[slot][r][0][65] = sum_f(f);
        }
    };

    __builtin_amdgcn_s_setprio(2);
    plane_phi(z0 - 1, false);
    plane_phi(z0, true);
    for (int k = z0; k < z1; ++k) {
#if ZSC_OPAQUE_STRIDE
        {
            uint32_t lo = (uint32_t)n;
            asm volatile("" : "+s"(lo));
            n = (int64_t)lo;
        }
#endif
        __builtin_amdgcn_s_setprio(2);
        // plane k's pulled g as stored words (shifted at the collision)
        typename Pair<SG>::type gw[19], ge[19];
        if (own) {
            const RowStarts rs(y, k, nx, ny, nz);
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const SG *p = in.g + (int64_t)i * n;
                const uint32_t row = rs.src(cy(i), cz(i));
                gw[i] = ld_pair(p, row + xo);
                if (cx(i) == 1) ge[i] = ld_pair(p, row + left);
                if (cx(i) == -1) ge[i] = ld_pair(p, row + right);
            }
        }
        plane_phi(k + 1, true);
        __syncthreads();
        __builtin_amdgcn_s_setprio(0);
        if (own) {
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                if (cx(i) == 1) gw[i] = shift_word<1>(gw[i], ge[i]);
                if (cx(i) == -1) gw[i] = shift_word<-1>(gw[i], ge[i]);
            }
            const int sm = (k + 3) & 3, s0 = k & 3, spl = (k + 1) & 3;
            const int r = ty + 1;
            const uint32_t orow = (uint32_t)nx * ((uint32_t)y + (uint32_t)ny * (uint32_t)k) + xo;
            using OE = decltype(enc_out<SG>(0.f));
            OE lo16[25];  // cell 0's outputs as stored: f1..6, g0..18
            float f0lo = 0.f;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                // x offsets of the stencil in [parity][lane + 1] form
                const int jm = j ^ 1, qm = tx + j, jp = j ^ 1, qp = tx + 1 + j, q = tx + 1;
                R p[19];
                p[0] = sp[s0][r][j][q];
                p[1] = sp[s0][r][jp][qp];
                p[2] = sp[s0][r][jm][qm];
                p[3] = sp[s0][r + 1][j][q];
                p[4] = sp[s0][r - 1][j][q];
                p[5] = sp[spl][r][j][q];
                p[6] = sp[sm][r][j][q];
                p[7] = sp[s0][r + 1][jp][qp];
                p[8] = sp[s0][r - 1][jm][qm];
                p[9] = sp[s0][r - 1][jp][qp];
                p[10] = sp[s0][r + 1][jm][qm];
                p[11] = sp[spl][r][jp][qp];
                p[12] = sp[sm][r][jm][qm];
                p[13] = sp[sm][r][jp][qp];
                p[14] = sp[spl][r][jm][qm];
                p[15] = sp[spl][r + 1][j][q];
                p[16] = sp[sm][r - 1][j][q];
                p[17] = sp[sm][r + 1][j][q];
                p[18] = sp[spl][r - 1][j][q];
                R fcur[7], ga[19], fb[7], gb[19];
#pragma unroll
                for (int i = 0; i < 7; ++i) fcur[i] = fsh[(k + 2) & 1][i][ty][j][tx];
#pragma unroll
                for (int i = 0; i < 19; ++i) {
                    if constexpr (Codec<SG>::deviation)
                        ga[i] = Codec<SG>::dec0(elem<SG>(gw[i], j));
                    else
                        ga[i] = Codec<SG>::dec(elem<SG>(gw[i], j), g_rest(c, i));
                }
                if constexpr (Codec<SG>::deviation)
                    collide_cell_dev(c, p, fcur, ga, fb, gb);
                else
                    collide_cell(c, p, fcur, ga, fb, gb);
                if (phi_out) phi_out[orow + j] = p[0];
                if (j == 0) {
                    f0lo = fb[0];
#pragma unroll
                    for (int i = 1; i < 7; ++i) lo16[i - 1] = enc_out<SF>(fb[i]);
#pragma unroll
                    for (int i = 0; i < 19; ++i) lo16[6 + i] = enc_out<SG>(gb[i]);
                } else {
                    f32x2 f0v;
                    f0v.x = f0lo;
                    f0v.y = fb[0];
                    f32x2 *f0p = (f32x2 *)((char *)out.f0 + (uint64_t)(orow * 4u));
                    if constexpr ((FL & kNtStore) != 0)
                        __builtin_nontemporal_store(f0v, f0p);
                    else
                        *f0p = f0v;
#pragma unroll
                    for (int i = 1; i < 7; ++i)
                        st_pair<FL>(out.f + (int64_t)(i - 1) * n, orow,
                                    join<SF>(lo16[i - 1], enc_out<SF>(fb[i])));
#pragma unroll
                    for (int i = 0; i < 19; ++i)
                        st_pair<FL>(out.g + (int64_t)i * n, orow,
                                    join<SG>(lo16[6 + i], enc_out<SG>(gb[i])));
                }
            }
        }
    }
}

// FP64 reductions for zsc_stats (src/stats.c:39-55, src/logsave.c:63-143).
template <class R>
__global__ void __launch_bounds__(256)
    k_stats(int nx, int ny, int nz, int64_t n, double alpha, double kappa, double phi_star_sq,
            const R *__restrict__ phi, const R *__restrict__ rho, const R *__restrict__ ux,
            const R *__restrict__ uy, const R *__restrict__ uz, double *partial)
{
    double a[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        const double ph = (double)phi[idx];
        a[0] += ph;
        a[1] += (double)rho[idx];
        if (ph >= 0.0) a[9] += 1.0;
        if (x < 1 || x > nx - 2 || y < 1 || y > ny - 2) continue;
        const double vx = ux[idx], vy = uy[idx], vz = uz[idx];
        const double usq = vx * vx + vy * vy + vz * vz;
        a[6] = usq > a[6] ? usq : a[6];
        Nbr nb(x, y, z, nx, ny, nz);
        double p[19];
#pragma unroll
        for (int d = 0; d < 19; ++d) p[d] = (double)phi[nb.at(cx(d), cy(d), cz(d))];
        const double pres = pressure19(alpha, phi_star_sq, kappa, p, (double)rho[idx]);
        if (ph >= 0.0) {
            a[2] += 1.0;
            a[3] += vx;
            a[4] += vy;
            a[5] += vz;
            a[7] += pres;
        } else {
            a[8] += pres;
        }
    }
    // wavefront reduction (64 lanes), then across the block's 4 waves
    __shared__ double red[4][10];
#pragma unroll
    for (int j = 0; j < 10; ++j) {
        double v = a[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(v, off, 64);
            v = (j == 6) ? (o > v ? o : v) : v + o;
        }
        a[j] = v;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int j = 0; j < 10; ++j) red[wv][j] = a[j];
    __syncthreads();
    if (threadIdx.x < 10) {
        const int j = threadIdx.x;
        double v = red[0][j];
        for (int w = 1; w < 4; ++w) v = (j == 6) ? (red[w][j] > v ? red[w][j] : v) : v + red[w][j];
        partial[(int64_t)blockIdx.x * 10 + j] = v;
    }
}

// Pressure field at precision R, written over the density array (each cell
// reads only its own rho, so the update in place is race-free).
template <class R>
__global__ void __launch_bounds__(256)
    k_pressure(int nx, int ny, int nz, int64_t n, R alpha, R phi_star_sq, R kappa,
               const R *__restrict__ phi, R *rho_p)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        R p[19];
        gather19(phi, Nbr(x, y, z, nx, ny, nz), p);
        rho_p[idx] = pressure19(alpha, phi_star_sq, kappa, p, rho_p[idx]);
    }
}

template <class R>
__global__ void __launch_bounds__(256) k_stats_final(int nblocks, const double *partial, double *out)
{
    // strided partial sums in a fixed order, then a fixed shuffle tree
    double a[10];
    for (int j = 0; j < 10; ++j) a[j] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256)
        for (int j = 0; j < 10; ++j) {
            const double p = partial[(int64_t)b * 10 + j];
            a[j] = (j == 6) ? (p > a[j] ? p : a[j]) : a[j] + p;
        }
    __shared__ double red[4][10];
    for (int j = 0; j < 10; ++j) {
        double v = a[j];
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(v, off, 64);
            v = (j == 6) ? (o > v ? o : v) : v + o;
        }
        a[j] = v;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int j = 0; j < 10; ++j) red[wv][j] = a[j];
    __syncthreads();
    if (threadIdx.x < 10) {
        const int j = threadIdx.x;
        double v = red[0][j];
        for (int w = 1; w < 4; ++w) v = (j == 6) ? (red[w][j] > v ? red[w][j] : v) : v + red[w][j];
        out[j] = v;
    }
}

// ---------------------------------------------------------------------------
// Ops<S> launchers
// ---------------------------------------------------------------------------
template <class S>
hipError_t Ops<S>::lat_init(const Geo &g, const zsc_params &p, hipStream_t s, const void *phi,
                            const void *rho, void *A)
{
    hipLaunchKernelGGL(k_lat_init<S>, grid_for(g.n), dim3(256), 0, s, g.nx, g.ny, g.nz, g.n,
                       make_coef<R>(p, 0), (const R *)phi, (const R *)rho,
                       LatLayout<S>::view(A, g.n));
    return hipGetLastError();
}

template <class S>
hipError_t Ops<S>::lat_phi(const Geo &g, hipStream_t s, const void *A, void *phi)
{
    zsc_params p{};
    hipLaunchKernelGGL(k_lat_phi<S>, grid_for(g.n), dim3(256), 0, s, g.n, make_coef<R>(p, 0),
                       LatLayout<S>::view(const_cast<void *>(A), g.n), (R *)phi);
    return hipGetLastError();
}

template <class S>
hipError_t Ops<S>::lat_collide(const Geo &g, const zsc_params &p, int grav_on, hipStream_t s,
                               const void *A, const void *phi, void *B)
{
    hipLaunchKernelGGL(k_lat_collide<S>, grid_for(g.n), dim3(256), 0, s, g.nx, g.ny, g.nz, g.n,
                       make_coef<R>(p, grav_on), LatLayout<S>::view(const_cast<void *>(A), g.n),
                       (const R *)phi, LatLayout<S>::view(B, g.n));
    return hipGetLastError();
}

template <class S>
hipError_t Ops<S>::lat_stream(const Geo &g, const zsc_params &p, hipStream_t s, const void *B,
                              void *A)
{
    hipLaunchKernelGGL(k_lat_stream<S>, grid_for(g.n), dim3(256), 0, s, g.nx, g.ny, g.nz, g.n,
                       make_coef<R>(p, 0), LatLayout<S>::view(const_cast<void *>(B), g.n),
                       LatLayout<S>::view(A, g.n));
    return hipGetLastError();
}

template <class S, int TY, int FL, bool PF>
static hipError_t launch_fused_v(const Geo &g, const Coef<typename Traits<S>::R> &c,
                                 hipStream_t s, const FusedCfg &cfg, const void *Bin, void *Bout,
                                 void *phi_out)
{
    int zchunk = cfg.zchunk > 0 ? cfg.zchunk : g.nz;
    const int nbx = (g.nx + 63) / 64, nby = (g.ny + TY - 1) / TY;
    // small boxes: shorter z chunks until there are enough blocks for 256 CUs
    while ((int64_t)nbx * nby * ((g.nz + zchunk - 1) / zchunk) < 2048 && zchunk > 4) zchunk /= 2;
    const int nbz = (g.nz + zchunk - 1) / zchunk;
    const int64_t nblk = (int64_t)nbx * nby * nbz;
    if (nblk > 0x7fffffff) return hipErrorInvalidValue;
    // 32-bit byte offsets within each population array
    if ((uint64_t)g.n * sizeof(typename Traits<S>::S0) > ((uint64_t)1 << 32) ||
        (uint64_t)g.n * sizeof(typename Traits<S>::SF) > ((uint64_t)1 << 32) ||
        (uint64_t)g.n * sizeof(typename Traits<S>::SG) > ((uint64_t)1 << 32))
        return hipErrorInvalidValue;
    // optional dynamic LDS that only limits how many blocks share a CU
    const unsigned pad = (unsigned)((cfg.flags >> 8) & 0xff) * 1024u;
    hipLaunchKernelGGL((k_fused<S, TY, FL, PF>), dim3((unsigned)nblk), dim3(64, TY), pad, s,
                       g.nx, g.ny, g.nz, g.n, zchunk, nbx, nby, (int)nblk,
                       (cfg.flags & kFusedIdentityMap) ? 0 : 1, c,
                       LatLayout<S>::view(const_cast<void *>(Bin), g.n),
                       LatLayout<S>::view(Bout, g.n), (typename Traits<S>::R *)phi_out);
    return hipGetLastError();
}

template <class S, int TY, int FL>
static hipError_t launch_fused_w(const Geo &g, const Coef<typename Traits<S>::R> &c,
                                 hipStream_t s, const FusedCfg &cfg, const void *Bin, void *Bout,
                                 void *phi_out)
{
    int zchunk = cfg.zchunk > 0 ? cfg.zchunk : g.nz;
    const int nbx = g.nx / 128, nby = (g.ny + TY - 1) / TY;
    while ((int64_t)nbx * nby * ((g.nz + zchunk - 1) / zchunk) < 2048 && zchunk > 4) zchunk /= 2;
    const int nbz = (g.nz + zchunk - 1) / zchunk;
    const int64_t nblk = (int64_t)nbx * nby * nbz;
    if (nblk > 0x7fffffff || g.nx % 128 != 0) return hipErrorInvalidValue;
    if ((uint64_t)g.n * sizeof(typename Traits<S>::S0) > ((uint64_t)1 << 32)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_fused_w<S, TY, FL>), dim3((unsigned)nblk), dim3(64, TY), 0, s, g.nx, g.ny,
                       g.nz, g.n, zchunk, nbx, nby, (int)nblk,
                       (cfg.flags & kFusedIdentityMap) ? 0 : 1, c,
                       LatLayout<S>::view(const_cast<void *>(Bin), g.n),
                       LatLayout<S>::view(Bout, g.n), (float *)phi_out);
    return hipGetLastError();
}

template <class S, int TY>
static hipError_t launch_fused_ty(const Geo &g, const Coef<typename Traits<S>::R> &c,
                                  hipStream_t s, const FusedCfg &cfg, const void *Bin, void *Bout,
                                  void *phi)
{
    if constexpr (sizeof(S) == 8) {
        // FP64: streaming-store hint and f_5 / f_6 stash only (the other bits
        // tune FP32 compute); one-row tiles ignore the stash
        if constexpr (TY > 1)
            if ((cfg.flags & kFusedZStash) && (cfg.flags & kNtStore))
                return launch_fused_v<S, TY, kNtStore | kZStash, false>(g, c, s, cfg, Bin, Bout, phi);
        if (cfg.flags & kNtStore) return launch_fused_v<S, TY, 1, false>(g, c, s, cfg, Bin, Bout, phi);
        return launch_fused_v<S, TY, 0, false>(g, c, s, cfg, Bin, Bout, phi);
    } else {
        if constexpr (TY == 4 || TY == 8) {
            // two cells per lane (rows of 128 cells)
            if ((cfg.flags & kFusedWide) && g.nx % 128 == 0) {
                if ((cfg.flags & kFusedZStash) && (cfg.flags & kNtStore))
                    return launch_fused_w<S, TY, kNtStore | kZStash>(g, c, s, cfg, Bin, Bout, phi);
                if (cfg.flags & kNtStore) return launch_fused_w<S, TY, 1>(g, c, s, cfg, Bin, Bout, phi);
                return launch_fused_w<S, TY, 0>(g, c, s, cfg, Bin, Bout, phi);
            }
        }
        const int fl = cfg.flags & (kNtStore | kNtLoad);
        const bool pf = (cfg.flags & kFusedPrefetch) != 0;
        if constexpr (TY > 1) {  // one-row tiles ignore the stash
            if ((cfg.flags & kFusedZStash) && fl == kNtStore) {
                if (pf) return launch_fused_v<S, TY, kNtStore | kZStash, true>(g, c, s, cfg, Bin, Bout, phi);
                return launch_fused_v<S, TY, kNtStore | kZStash, false>(g, c, s, cfg, Bin, Bout, phi);
            }
        }
        switch (fl + (pf ? 4 : 0)) {
        case 0: return launch_fused_v<S, TY, 0, false>(g, c, s, cfg, Bin, Bout, phi);
        case 1: return launch_fused_v<S, TY, 1, false>(g, c, s, cfg, Bin, Bout, phi);
        case 3: return launch_fused_v<S, TY, 3, false>(g, c, s, cfg, Bin, Bout, phi);
        case 4: return launch_fused_v<S, TY, 0, true>(g, c, s, cfg, Bin, Bout, phi);
        case 5: return launch_fused_v<S, TY, 1, true>(g, c, s, cfg, Bin, Bout, phi);
        case 7: return launch_fused_v<S, TY, 3, true>(g, c, s, cfg, Bin, Bout, phi);
        default: return hipErrorInvalidValue;
        }
    }
}

template <class S>
hipError_t Ops<S>::lat_fused(const Geo &g, const zsc_params &p, int grav_on, hipStream_t s,
                             const FusedCfg &cfg, const void *Bin, void *Bout, void *phi)
{
    const Coef<R> c = make_coef<R>(p, grav_on);
    switch (cfg.ty) {
    case 1: return launch_fused_ty<S, 1>(g, c, s, cfg, Bin, Bout, phi);
    case 2: return launch_fused_ty<S, 2>(g, c, s, cfg, Bin, Bout, phi);
    case 4: return launch_fused_ty<S, 4>(g, c, s, cfg, Bin, Bout, phi);
    case 8: return launch_fused_ty<S, 8>(g, c, s, cfg, Bin, Bout, phi);
    default: return hipErrorInvalidValue;
    }
}

template <class S>
hipError_t Ops<S>::lat_moments(const Geo &g, const zsc_params &p, hipStream_t s, const void *A,
                               void *phi, void *rho, void *ux, void *uy, void *uz)
{
    hipLaunchKernelGGL(k_lat_moments<S>, grid_for(g.n), dim3(256), 0, s, g.n, make_coef<R>(p, 0),
                       LatLayout<S>::view(const_cast<void *>(A), g.n), (R *)phi, (R *)rho,
                       (R *)ux, (R *)uy, (R *)uz);
    return hipGetLastError();
}

template <class S>
hipError_t Ops<S>::pop_get(const Geo &g, const zsc_params &p, hipStream_t s, const void *L, int q,
                           double *dst)
{
    const Lattice<S> v = LatLayout<S>::view(const_cast<void *>(L), g.n);
    const Coef<R> c = make_coef<R>(p, 0);
    if (q < 7) {
        const auto *ptr = q == 0 ? v.f : v.f + (int64_t)(q - 1) * g.n;
        hipLaunchKernelGGL((k_pop_get<typename Traits<S>::S0, typename Traits<S>::SF, R>),
                           grid_for(g.n), dim3(256), 0, s, g.n, v.f0, ptr, q == 0 ? 1 : 0, R(0), dst);
    } else {
        const auto *ptr = v.g + (int64_t)(q - 7) * g.n;
        hipLaunchKernelGGL((k_pop_get<typename Traits<S>::S0, typename Traits<S>::SG, R>),
                           grid_for(g.n), dim3(256), 0, s, g.n, v.f0, ptr, 0,
                           g_weight<R>(q - 7) * c.rho_mean, dst);
    }
    return hipGetLastError();
}

template <class S>
hipError_t Ops<S>::pop_set(const Geo &g, const zsc_params &p, hipStream_t s, void *L, int q,
                           const double *src)
{
    const Lattice<S> v = LatLayout<S>::view(L, g.n);
    const Coef<R> c = make_coef<R>(p, 0);
    if (q < 7) {
        auto *ptr = q == 0 ? v.f : v.f + (int64_t)(q - 1) * g.n;
        hipLaunchKernelGGL((k_pop_set<typename Traits<S>::S0, typename Traits<S>::SF, R>),
                           grid_for(g.n), dim3(256), 0, s, g.n, v.f0, ptr, q == 0 ? 1 : 0, R(0), src);
    } else {
        auto *ptr = v.g + (int64_t)(q - 7) * g.n;
        hipLaunchKernelGGL((k_pop_set<typename Traits<S>::S0, typename Traits<S>::SG, R>),
                           grid_for(g.n), dim3(256), 0, s, g.n, v.f0, ptr, 0,
                           g_weight<R>(q - 7) * c.rho_mean, src);
    }
    return hipGetLastError();
}

template <class S>
hipError_t Ops<S>::to_f64(hipStream_t s, int64_t n, const void *src, double *dst)
{
    hipLaunchKernelGGL(k_to_f64<R>, grid_for(n), dim3(256), 0, s, n, (const R *)src, dst);
    return hipGetLastError();
}

template <class S>
hipError_t Ops<S>::from_f64(hipStream_t s, int64_t n, const double *src, void *dst)
{
    hipLaunchKernelGGL(k_from_f64<R>, grid_for(n), dim3(256), 0, s, n, src, (R *)dst);
    return hipGetLastError();
}

template <class S>
hipError_t Ops<S>::init_phi_bubbles(const Geo &g, hipStream_t s, const double *bubbles, int nb,
                                    double phi_star, double width, void *phi)
{
    hipLaunchKernelGGL(k_init_phi_bubbles<R>, grid_for(g.n), dim3(256), 0, s, g.nx, g.ny, g.n,
                       bubbles, nb, phi_star, width, (R *)phi);
    return hipGetLastError();
}

template <class S>
hipError_t Ops<S>::stats(const Geo &g, const zsc_params &p, hipStream_t s, const void *phi,
                         const void *rho, const void *ux, const void *uy, const void *uz,
                         double *partial, double *out)
{
    hipLaunchKernelGGL(k_stats<R>, dim3(kStatsBlocks), dim3(256), 0, s, g.nx, g.ny, g.nz, g.n,
                       p.alpha, p.kappa, p.phi_star_sq, (const R *)phi, (const R *)rho,
                       (const R *)ux, (const R *)uy, (const R *)uz, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_stats_final<R>, dim3(1), dim3(256), 0, s, kStatsBlocks, partial, out);
    return hipGetLastError();
}

template <class S>
hipError_t Ops<S>::pressure(const Geo &g, const zsc_params &p, hipStream_t s, const void *phi,
                            void *rho_p)
{
    hipLaunchKernelGGL(k_pressure<R>, grid_for(g.n), dim3(256), 0, s, g.nx, g.ny, g.nz, g.n,
                       (R)p.alpha, (R)p.phi_star_sq, (R)p.kappa, (const R *)phi, (R *)rho_p);
    return hipGetLastError();
}

// ===========================================================================
// Reference-layout (drop-in) kernels: each population pointer addresses
// 2n elements, A = [0, n) post-stream, B = [n, 2n) post-collision.
// ===========================================================================
template <class T, int Q>
struct Pops {
    T *p[Q];
};

template <class T, int Q>
inline Pops<T, Q> pops(T *const *a)
{
    Pops<T, Q> r;
    for (int i = 0; i < Q; ++i) r.p[i] = a[i];
    return r;
}

// src/init_f.cu:42-79
template <class T>
__global__ void __launch_bounds__(256)
    k_init_f(int nx, int ny, int nz, int64_t n, Coef<T> c, const T *__restrict__ phi, Pops<T, 7> f)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        Nbr nb(x, y, z, nx, ny, nz);
        T p[19], v[7];
        gather19(phi, nb, p);
        init_f_cell(c, p, v);
#pragma unroll
        for (int i = 0; i < 7; ++i) f.p[i][idx] = v[i];
    }
}

// src/init_g.cu:46-93
template <class T>
__global__ void __launch_bounds__(256)
    k_init_g(int nx, int ny, int nz, int64_t n, Coef<T> c, const T *__restrict__ phi,
             const T *__restrict__ rho, Pops<T, 19> g)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        Nbr nb(x, y, z, nx, ny, nz);
        T p[19], v[19];
        gather19(phi, nb, p);
        init_g_cell(c, p, rho[idx], v);
#pragma unroll
        for (int i = 0; i < 19; ++i) g.p[i][idx] = v[i];
    }
}

// src/update_phi.cu:40-44
template <class T>
__global__ void __launch_bounds__(256) k_update_phi(int64_t n, T *phi, Pops<T, 7> f)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        T v[7];
#pragma unroll
        for (int i = 0; i < 7; ++i) v[i] = f.p[i][idx];
        phi[idx] = sum_f(v);
    }
}

// src/update_rho.cu:44-51
template <class T>
__global__ void __launch_bounds__(256) k_update_rho(int64_t n, T *rho, Pops<T, 19> g)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        T v[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) v[i] = g.p[i][idx];
        rho[idx] = sum_g(v);
    }
}

// src/update_velocity.cu:45-61 (g_0 is not read)
template <class T>
__global__ void __launch_bounds__(256)
    k_update_velocity(int64_t n, const T *rho, T *ux, T *uy, T *uz, Pops<T, 19> g)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        T v[19];
        v[0] = T(0);
#pragma unroll
        for (int i = 1; i < 19; ++i) v[i] = g.p[i][idx];
        T a, b, d;
        velocity(v, rho[idx], a, b, d);
        ux[idx] = a;
        uy[idx] = b;
        uz[idx] = d;
    }
}

// src/collision_f.cu:44-116
template <class T>
__global__ void __launch_bounds__(256)
    k_collision_f(int nx, int ny, int nz, int64_t n, Coef<T> c, const T *__restrict__ phi,
                  const T *__restrict__ rho, const T *__restrict__ ux, const T *__restrict__ uy,
                  const T *__restrict__ uz, Pops<T, 7> f)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        Nbr nb(x, y, z, nx, ny, nz);
        T p[19], fa[7], fb[7], mu, fx, fy, fz;
        gather19(phi, nb, p);
        force19(c, p, mu, fx, fy, fz);
#pragma unroll
        for (int i = 0; i < 7; ++i) fa[i] = f.p[i][idx];
        collide_f(c, p[0], rho[idx], ux[idx], uy[idx], uz[idx], mu, fx, fy, fz, fa, fb);
#pragma unroll
        for (int i = 0; i < 7; ++i) f.p[i][n + idx] = fb[i];
    }
}

// src/collision_g.cu:49-196
template <class T>
__global__ void __launch_bounds__(256)
    k_collision_g(int nx, int ny, int nz, int64_t n, Coef<T> c, const T *__restrict__ phi,
                  const T *__restrict__ rho, const T *__restrict__ ux, const T *__restrict__ uy,
                  const T *__restrict__ uz, Pops<T, 19> g)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        Nbr nb(x, y, z, nx, ny, nz);
        T p[19], ga[19], gb[19], mu, fx, fy, fz;
        gather19(phi, nb, p);
        force19(c, p, mu, fx, fy, fz);
#pragma unroll
        for (int i = 0; i < 19; ++i) ga[i] = g.p[i][idx];
        collide_g(c, p[0], rho[idx], ux[idx], uy[idx], uz[idx], mu, fx, fy, fz, ga, gb);
#pragma unroll
        for (int i = 0; i < 19; ++i) g.p[i][n + idx] = gb[i];
    }
}

// src/stream_f.cu:40-61 as a pull, halo of src/pack_mpi_f.cu / unpack_mpi_f.cu
// and src/pack_mpi.cu / unpack_mpi.cu folded into the periodic wrap.
template <class T>
__global__ void __launch_bounds__(256)
    k_stream_f(int nx, int ny, int nz, int64_t n, T eta, T eta2, Pops<T, 7> f)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        Nbr nb(x, y, z, nx, ny, nz);
        f.p[0][idx] = f.p[0][n + idx];
#pragma unroll
        for (int i = 1; i < 7; ++i)
            f.p[i][idx] = eta * f.p[i][n + nb.at(-cx(i), -cy(i), -cz(i))] + eta2 * f.p[i][n + idx];
    }
}

// src/stream_g.cu:44-78 as a pull (+ halo, as above).
template <class T>
__global__ void __launch_bounds__(256) k_stream_g(int nx, int ny, int nz, int64_t n, Pops<T, 19> g)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        Nbr nb(x, y, z, nx, ny, nz);
#pragma unroll
        for (int i = 0; i < 19; ++i) g.p[i][idx] = g.p[i][n + nb.at(-cx(i), -cy(i), -cz(i))];
    }
}

// The rest values w_i rho_mean of the 19 g populations (kernel argument).
template <class T>
struct GRest {
    T v[19];
};

// DEV: the lattice's g are deviations from the rest values (f32d), which
// export adds back and import takes off.
template <class T, bool DEV>
__global__ void __launch_bounds__(256)
    k_export(int64_t n, Lattice<T> A, Pops<T, 7> f, Pops<T, 19> g, GRest<T> rest)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        f.p[0][idx] = A.f0[idx];
#pragma unroll
        for (int i = 1; i < 7; ++i) f.p[i][idx] = A.f[(int64_t)(i - 1) * n + idx];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            const T v = A.g[(int64_t)i * n + idx];
            g.p[i][idx] = DEV ? v + rest.v[i] : v;
        }
    }
}

template <class T, bool DEV>
__global__ void __launch_bounds__(256)
    k_import(int64_t n, Lattice<T> A, Pops<T, 7> f, Pops<T, 19> g, GRest<T> rest)
{
    ZSC_GRID_STRIDE(idx, n)
    {
        A.f0[idx] = f.p[0][idx];
#pragma unroll
        for (int i = 1; i < 7; ++i) A.f[(int64_t)(i - 1) * n + idx] = f.p[i][idx];
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            const T v = g.p[i][idx];
            A.g[(int64_t)i * n + idx] = DEV ? v - rest.v[i] : v;
        }
    }
}

// w_i rho_mean as the fused path forms it (g_rest), or zeros.
template <class T>
inline GRest<T> grest_of(const zsc_params *p)
{
    GRest<T> r{};
    if (p) {
        const Coef<T> c = make_coef<T>(*p, 0);
        for (int i = 0; i < 19; ++i) r.v[i] = g_weight<T>(i) * c.rho_mean;
    }
    return r;
}

#define ZSC_LAUNCH(kern, ...)                                                     \
    do {                                                                          \
        hipLaunchKernelGGL(kern, grid_for(g.n), dim3(256), 0, s, __VA_ARGS__);   \
        return hipGetLastError();                                                 \
    } while (0)

template <class T>
hipError_t DropIn<T>::init_f(const Geo &g, const zsc_params &p, hipStream_t s, const T *phi,
                             T *const f[7])
{
    ZSC_LAUNCH(k_init_f<T>, g.nx, g.ny, g.nz, g.n, make_coef<T>(p, 0), phi, pops<T, 7>(f));
}

template <class T>
hipError_t DropIn<T>::init_g(const Geo &g, const zsc_params &p, hipStream_t s, const T *phi,
                             const T *rho, T *const gp[19])
{
    ZSC_LAUNCH(k_init_g<T>, g.nx, g.ny, g.nz, g.n, make_coef<T>(p, 0), phi, rho, pops<T, 19>(gp));
}

template <class T>
hipError_t DropIn<T>::update_phi(const Geo &g, hipStream_t s, T *phi, T *const f[7])
{
    ZSC_LAUNCH(k_update_phi<T>, g.n, phi, pops<T, 7>(f));
}

template <class T>
hipError_t DropIn<T>::update_rho(const Geo &g, hipStream_t s, T *rho, T *const gp[19])
{
    ZSC_LAUNCH(k_update_rho<T>, g.n, rho, pops<T, 19>(gp));
}

template <class T>
hipError_t DropIn<T>::update_velocity(const Geo &g, hipStream_t s, const T *rho, T *ux, T *uy,
                                      T *uz, T *const gp[19])
{
    ZSC_LAUNCH(k_update_velocity<T>, g.n, rho, ux, uy, uz, pops<T, 19>(gp));
}

template <class T>
hipError_t DropIn<T>::collision_f(const Geo &g, const zsc_params &p, int grav_on, hipStream_t s,
                                  const T *phi, const T *rho, const T *ux, const T *uy,
                                  const T *uz, T *const f[7])
{
    ZSC_LAUNCH(k_collision_f<T>, g.nx, g.ny, g.nz, g.n, make_coef<T>(p, grav_on), phi, rho, ux,
               uy, uz, pops<T, 7>(f));
}

template <class T>
hipError_t DropIn<T>::collision_g(const Geo &g, const zsc_params &p, int grav_on, hipStream_t s,
                                  const T *phi, const T *rho, const T *ux, const T *uy,
                                  const T *uz, T *const gp[19])
{
    ZSC_LAUNCH(k_collision_g<T>, g.nx, g.ny, g.nz, g.n, make_coef<T>(p, grav_on), phi, rho, ux,
               uy, uz, pops<T, 19>(gp));
}

template <class T>
hipError_t DropIn<T>::stream_f(const Geo &g, const zsc_params &p, hipStream_t s, T *const f[7])
{
    ZSC_LAUNCH(k_stream_f<T>, g.nx, g.ny, g.nz, g.n, (T)p.eta, (T)p.eta2, pops<T, 7>(f));
}

template <class T>
hipError_t DropIn<T>::stream_g(const Geo &g, hipStream_t s, T *const gp[19])
{
    ZSC_LAUNCH(k_stream_g<T>, g.nx, g.ny, g.nz, g.n, pops<T, 19>(gp));
}

template <class T>
hipError_t DropIn<T>::export_state(const Geo &g, hipStream_t s, const void *A, T *const f[7],
                                   T *const gp[19], const zsc_params *dev)
{
    const Lattice<T> v = LatLayout<T>::view(const_cast<void *>(A), g.n);
    if (dev) ZSC_LAUNCH((k_export<T, true>), g.n, v, pops<T, 7>(f), pops<T, 19>(gp), grest_of<T>(dev));
    ZSC_LAUNCH((k_export<T, false>), g.n, v, pops<T, 7>(f), pops<T, 19>(gp), grest_of<T>(nullptr));
}

template <class T>
hipError_t DropIn<T>::import_state(const Geo &g, hipStream_t s, void *A, T *const f[7],
                                   T *const gp[19], const zsc_params *dev)
{
    const Lattice<T> v = LatLayout<T>::view(A, g.n);
    if (dev) ZSC_LAUNCH((k_import<T, true>), g.n, v, pops<T, 7>(f), pops<T, 19>(gp), grest_of<T>(dev));
    ZSC_LAUNCH((k_import<T, false>), g.n, v, pops<T, 7>(f), pops<T, 19>(gp), grest_of<T>(nullptr));
}

#undef ZSC_LAUNCH

}  // namespace zsc
