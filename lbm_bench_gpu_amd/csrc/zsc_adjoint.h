// zsc_adjoint.h -- hand-written reverse mode of the ZSC time step (no
// reference counterpart: the reference has no derivatives).
//
// Forward step (A-state form): A_{n+1} = S(C(A_n; theta)), with C the cell
// collision of src/collision_f.cu / src/collision_g.cu, fed by the moments of
// src/update_*.cu and the 19-point phi stencil, and S the streaming of
// src/stream_f.cu / src/stream_g.cu.  The backward step maps
// lam_{n+1} = dL/dA_{n+1} to lam_n = J(A_n)^T lam_{n+1}:
//   1. S^T (a gather from downwind): lam_B,i(x) = lam_i(x + c_i) for g;
//      eta lam_i(x + c_i) + eta2 lam_i(x) for f_1..6.
//   2. C^T per cell: the transposed BGK + forcing Jacobian with respect to
//      the local populations (through phi, rho, u) and to mu and grad(phi).
//   3. The adjoint of mu = 4 beta phi (phi^2 - phi*^2) - kappa lap(phi):
//      lap is symmetric, so its transpose is itself; the isotropic gradient
//      is antisymmetric, so its transpose is minus itself.  A second pass
//      adds L[a_lap] - div(a_grad) to every f population's adjoint.
// The phi > 0 gravity switch (src/collision_f.cu:96) is a step function; its
// derivative with respect to phi is taken as zero.
// Alongside, dL/d(step coefficients) accumulates in FP64; the host maps those
// to kappa, beta, tau_f, tau_g and gravity by the chain rule of the
// parameter derivation (src/paramread.c:141-170).
#pragma once

#include "zsc_kernels.h"

namespace zsc {

// Coefficient slots of the parameter adjoint.
enum : int {
    kGAlpha4 = 0,
    kGKappa,
    kGInvTauPhiOne,
    kGInvTauPhi,
    kGInvEta2,
    kGEta,
    kGInvTauRhoOne,
    kGK0,
    kGK1,
    kGK2,
    kGKC0,
    kGKC1,
    kGKC2,
    kGGrav,
    kGLoss,  // the loss itself (seed kernel)
    kNumGrad = 16
};

// Reverse mode of one cell's update + collision.  Inputs: the phi stencil p,
// the post-stream populations fa/ga, the adjoints lf/lg of the
// post-collision populations and lfd (the streaming's upwind-minus-own
// adjoint differences).  Outputs: the local adjoints afa/aga of fa/ga (afa
// without the phi-stencil coupling), the adjoint of phi(x) through its local
// uses, of lap(phi)(x) and of grad(phi)(x), and deta = sum_i fb_i lfd_i (the
// eta derivative of the streaming, fb the forward post-collision f).  Each
// intermediate is formed as late, and dropped as early, as the data flow
// allows, which keeps the kernel's register count down.
//
// P is the precision of the states, the forward recompute and the
// coefficient derivatives ap (with c); R that of the adjoint quantities
// (with cr, the same coefficients at R).  With P = R every expression is the
// plain one at that precision.  FP32 adjoints over FP64 states (P = double,
// R = float) keep in FP64 what cancels: the coefficient sums pair large
// primal values (sum_i lam_i g_i against lambda-weighted equilibria) whose
// difference is ~1e-4 of either.
//
// DEV: ga holds deviations g_i - w_i rho_mean (FP32 states stored so, as
// the f32d lattice keeps them).  rho = rho_mean + sum of deviations, the
// momenta are the same sums (sum_i c_i w_i = 0), and the coefficient sums
// of inv_tau_rho_one (sum_i g_i lam_i) and of k0..k2 (the "rho * 1" terms
// of the equilibria) leave out the rest state's part, rho_mean sum_i w_i
// lam_i: through tau_g (inv_tau_rho_one = 1 - 1/tau_g, k_i = w_i/tau_g) the
// two parts cancel exactly, and without them the sums keep the bits that
// FP32 would lose to that cancellation.
template <class P, class R, bool DEV = false>
__device__ inline void adj_cell(const Coef<P> &c, const Coef<R> &cr, const P *p, const P *fa,
                                const P *ga, const R *lf, const P *lfd, const R *lg, R *afa,
                                R *aga, R &aphi, R &alap, R *agr, P &deta, P *ap, int grav_on)
{
    // ---- forward recompute
    const P phi = p[0];
    const P gsum = sum_g(ga);
    const P rho = DEV ? c.rho_mean + gsum : gsum;
    const P jx = ga[1] - ga[2] + ga[7] - ga[8] + ga[9] - ga[10] + ga[11] - ga[12] + ga[13] - ga[14];
    const P jy = ga[3] - ga[4] + ga[7] - ga[8] - ga[9] + ga[10] + ga[15] - ga[16] + ga[17] - ga[18];
    const P jz = ga[5] - ga[6] + ga[11] - ga[12] - ga[13] + ga[14] + ga[15] - ga[16] - ga[17] + ga[18];
    // quotients by rho: FP32 by the hardware reciprocal, FP64 exact
    using Qp = ::zsc::Q<P>;
    using Qr = ::zsc::Q<R>;
    const P u[3] = {Qp::div(jx, rho), Qp::div(jy, rho), Qp::div(jz, rho)};
    const P lap = lap19(p);
    const P G[3] = {
        (P(2) * (p[1] - p[2]) + p[7] - p[8] + p[9] - p[10] + p[11] - p[12] + p[13] - p[14]) *
            W<P>::inv12,
        (P(2) * (p[3] - p[4]) + p[7] - p[8] + p[10] - p[9] + p[15] - p[16] + p[17] - p[18]) *
            W<P>::inv12,
        (P(2) * (p[5] - p[6]) + p[11] - p[12] + p[14] - p[13] + p[15] - p[16] + p[18] - p[17]) *
            W<P>::inv12};
    const P phi3 = phi * (phi * phi - c.phi_star_sq);
    const P mu = c.alpha4 * phi3 - c.kappa * lap;
    const P H = phi > P(0) ? P(1) : P(0);
    P F[3] = {mu * G[0], mu * G[1], mu * G[2] + H * c.grav};
    P v[3];
    for (int a = 0; a < 3; ++a) v[a] = u[a] + Qp::div(P(0.5) * F[a], rho);
    const P hg = P(0.5) * c.gamma * c.inv_tau_phi;
    const P cfk = c.inv_tau_phi * c.inv_eta2;
    const P Af = hg * mu, Cf = cfk * phi;
    {
        P fb[7];
        fb[0] = c.inv_tau_phi_one * fa[0] - P(6) * Af + c.inv_tau_phi * phi;
        for (int a = 0; a < 3; ++a) {
            fb[1 + 2 * a] = c.inv_tau_phi_one * fa[1 + 2 * a] + Af + Cf * v[a];
            fb[2 + 2 * a] = c.inv_tau_phi_one * fa[2 + 2 * a] + Af - Cf * v[a];
        }
        deta = P(0);
#pragma unroll
        for (int i = 1; i < 7; ++i) deta += fb[i] * lfd[i];
    }
    // the primal values the reverse sweep uses, at R
    const R phi_r = (R)phi, mu_r = (R)mu, rho_r = (R)rho;

    // ---- f collision (src/collision_f.cu:105-114), reversed
    R a_phi = R(0), a_mu = R(0), a_rho = R(0), a_v[3] = {0, 0, 0}, a_F[3] = {0, 0, 0};
    const R tf = cr.inv_tau_phi_one;
    P s_tf = P(0);
    for (int i = 0; i < 7; ++i) s_tf += fa[i] * (P)lf[i];
    ap[kGInvTauPhiOne] += s_tf;
    R a_Af = R(-6) * lf[0];
    for (int i = 1; i < 7; ++i) a_Af += lf[i];
    a_phi += cr.inv_tau_phi * lf[0];
    P a_Cf = P(0);
    for (int a = 0; a < 3; ++a) {
        const P d = (P)lf[1 + 2 * a] - (P)lf[2 + 2 * a];
        a_Cf += v[a] * d;
        a_v[a] += (R)(Cf * d);
    }
    a_mu += (R)hg * a_Af;
    a_phi += (R)(cfk * a_Cf);
    ap[kGInvTauPhi] += phi * (P)lf[0] + P(0.5) * c.gamma * mu * (P)a_Af + c.inv_eta2 * phi * a_Cf;
    ap[kGInvEta2] += c.inv_tau_phi * phi * a_Cf;

    // ---- g collision (src/collision_g.cu:105-194), reversed
    const R tg = cr.inv_tau_rho_one;
    P s_tg = P(0);
    for (int i = 0; i < 19; ++i) s_tg += ga[i] * (P)lg[i];
    ap[kGInvTauRhoOne] += s_tg;
    const P ut[3] = {P(3) * v[0], P(3) * v[1], P(3) * v[2]};
    const P UF = v[0] * F[0] + v[1] * F[1] + v[2] * F[2];
    const P Q = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    P X;
    if constexpr (DEV)
        X = P(3) * phi * mu + gsum - P(1.5) * rho * Q;  // drho for rho * 1
    else
        X = P(3) * phi * mu + rho * (P(1) - P(1.5) * Q);
    const R Q_r = (R)Q;
    const R utr[3] = {(R)ut[0], (R)ut[1], (R)ut[2]}, Fr[3] = {(R)F[0], (R)F[1], (R)F[2]};
    R a_UF = R(0), a_Q = R(0), a_X = R(0), a_ut[3] = {0, 0, 0};
    P g_k0, g_k1 = P(0), g_k2 = P(0), g_kc0, g_kc1 = P(0), g_kc2 = P(0);
    {
        const R D = lg[0];
        // the "rho * 1" term: drho (= gsum) for deviation states
        g_k0 = (gsum - P(6) * phi * mu - P(1.5) * rho * Q) * (P)D;
        a_rho += cr.k0 * (R(1) - R(1.5) * Q_r) * D;
        a_phi += R(-6) * cr.k0 * mu_r * D;
        a_mu += R(-6) * cr.k0 * phi_r * D;
        a_Q += R(-1.5) * cr.k0 * rho_r * D;
        g_kc0 = -UF * (P)D;
        a_UF += -cr.kc0 * D;
    }
    for (int a = 0; a < 3; ++a) {
        const P Sp = (P)lg[1 + 2 * a] + (P)lg[2 + 2 * a], Dl = (P)lg[1 + 2 * a] - (P)lg[2 + 2 * a];
        const R Spr = (R)Sp, Dlr = (R)Dl;
        const P w = ut[a];
        const R wr = utr[a];
        a_X += cr.k1 * Spr;
        g_k1 += X * Sp + P(0.5) * rho * w * w * Sp + w * rho * Dl;
        a_rho += R(0.5) * cr.k1 * wr * wr * Spr + wr * cr.k1 * Dlr;
        a_ut[a] += (cr.k1 * rho_r * wr + cr.kc1 * Fr[a]) * Spr + (cr.k1 * rho_r + cr.kc1 * Fr[a]) * Dlr;
        a_F[a] += cr.kc1 * wr * Spr + wr * cr.kc1 * Dlr;
        g_kc1 += (w * F[a] - UF) * Sp + w * F[a] * Dl;
        a_UF += -cr.kc1 * Spr;
    }
    const int da[6] = {0, 0, 0, 0, 1, 1}, db[6] = {1, 1, 2, 2, 2, 2};
    const P sg[6] = {P(1), P(-1), P(1), P(-1), P(1), P(-1)};
    const R sgr[6] = {R(1), R(-1), R(1), R(-1), R(1), R(-1)};
    for (int d = 0; d < 6; ++d) {
        const int ip = 7 + 2 * d, im = 8 + 2 * d, a = da[d], b = db[d];
        const P Sp = (P)lg[ip] + (P)lg[im], Dl = (P)lg[ip] - (P)lg[im];
        const R Spr = (R)Sp, Dlr = (R)Dl;
        const P s = ut[a] + sg[d] * ut[b], S = F[a] + sg[d] * F[b];
        const R sr = (R)s, Sr = (R)S;
        a_X += cr.k2 * Spr;
        g_k2 += X * Sp + P(0.5) * rho * s * s * Sp + rho * s * Dl;
        a_rho += R(0.5) * cr.k2 * sr * sr * Spr + cr.k2 * sr * Dlr;
        const R a_s = (cr.k2 * rho_r * sr + cr.kc2 * Sr) * Spr + cr.k2 * rho_r * Dlr;
        const R a_S = cr.kc2 * sr * Spr + cr.kc2 * Dlr;
        g_kc2 += (s * S - UF) * Sp + S * Dl;
        a_UF += -cr.kc2 * Spr;
        a_ut[a] += a_s;
        a_ut[b] += sgr[d] * a_s;
        a_F[a] += a_S;
        a_F[b] += sgr[d] * a_S;
    }
    ap[kGK0] += g_k0;
    ap[kGK1] += g_k1;
    ap[kGK2] += g_k2;
    ap[kGKC0] += g_kc0;
    ap[kGKC1] += g_kc1;
    ap[kGKC2] += g_kc2;
    // X = 3 phi mu + rho (1 - 1.5 Q)
    a_phi += R(3) * mu_r * a_X;
    a_mu += R(3) * phi_r * a_X;
    a_rho += (R(1) - R(1.5) * Q_r) * a_X;
    a_Q += R(-1.5) * rho_r * a_X;
    const R vr[3] = {(R)v[0], (R)v[1], (R)v[2]};
    for (int a = 0; a < 3; ++a) {
        a_v[a] += R(3) * a_ut[a] + R(2) * vr[a] * a_Q + Fr[a] * a_UF;
        a_F[a] += vr[a] * a_UF;
    }
    // v = u + F / (2 rho)
    R a_u[3];
    for (int a = 0; a < 3; ++a) {
        a_u[a] = a_v[a];
        a_F[a] += Qr::div(a_v[a] * R(0.5), rho_r);
        a_rho += Qr::div(-Fr[a] * a_v[a] * R(0.5), rho_r * rho_r);
    }
    // F = mu grad(phi) + grav H(phi) e_z
    a_mu += (R)G[0] * a_F[0] + (R)G[1] * a_F[1] + (R)G[2] * a_F[2];
    for (int a = 0; a < 3; ++a) agr[a] = mu_r * a_F[a];
    if (grav_on) ap[kGGrav] += H * (P)a_F[2];
    // mu = alpha4 phi (phi^2 - phi*^2) - kappa lap
    a_phi += cr.alpha4 * (R(3) * phi_r * phi_r - cr.phi_star_sq) * a_mu;
    ap[kGAlpha4] += phi3 * (P)a_mu;
    ap[kGKappa] += -lap * (P)a_mu;
    alap = -cr.kappa * a_mu;
    // u = j / rho, rho = sum g, j = sum c g
    R a_j[3];
    for (int a = 0; a < 3; ++a) {
        a_j[a] = Qr::div(a_u[a], rho_r);
        a_rho += Qr::div(-(R)u[a] * a_u[a], rho_r);
    }
#pragma unroll
    for (int i = 0; i < 7; ++i) afa[i] = tf * lf[i];
#pragma unroll
    for (int i = 0; i < 19; ++i)
        aga[i] = tg * lg[i] + (a_rho + R(cx(i)) * a_j[0] + R(cy(i)) * a_j[1] + R(cz(i)) * a_j[2]);
    aphi = a_phi;
}

// Wavefront + block reduction of kNumGrad doubles into partial[block].
template <int NT>
__device__ inline void block_reduce_grads(double *ap, double *partial, bool accumulate = false)
{
    __shared__ double red[NT / 64][kNumGrad];
#pragma unroll
    for (int j = 0; j < kNumGrad; ++j) {
        double v = ap[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        ap[j] = v;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int j = 0; j < kNumGrad; ++j) red[wv][j] = ap[j];
    __syncthreads();
    if (threadIdx.x < kNumGrad) {
        double v = 0.0;
        for (int w = 0; w < NT / 64; ++w) v += red[w][threadIdx.x];
        double *dst = partial + (int64_t)blockIdx.x * kNumGrad + threadIdx.x;
        *dst = accumulate ? *dst + v : v;
    }
}

// Blocks of the adjoint kernels: kAdjCellsPerThread cells per thread (the
// per-block reduction of the coefficient sums, ~15% of pass 1's VALU work
// for one cell, is paid once per thread; 2 times 3.6% faster than 1 and
// 1.5% faster than 4, whose last round of blocks runs part-empty; a grid of
// one round, the blocks the GPU holds at once with about 11 cells per
// thread, is 2-4% slower), a multiple of 8 blocks for the XCD-contiguous
// order, more cells per thread beyond kAdjMaxBlocks blocks (per-thread sums
// at precision R, block and global sums in FP64).  The partial buffers hold
// kAdjMaxBlocks rows.
constexpr int kAdjMaxBlocks = 65536;
constexpr int kAdjCellsPerThread = 2;
inline int adj_blocks(int64_t n)
{
    constexpr int cpt = kAdjCellsPerThread;
    int64_t b = (n + 256 * cpt - 1) / (256 * cpt);
    b = (b + 7) & ~(int64_t)7;
    return (int)(b < 8 ? 8 : (b > kAdjMaxBlocks ? kAdjMaxBlocks : b));
}

// Cells of a pass 1 thread.  Blocks are dispatched round-robin over the 8
// XCDs; with `remap` the blocks of XCD j (blockIdx % 8 == j) sweep the j-th
// contiguous eighth of the cells in order, so the rows and planes next to
// the cells being updated were mostly fetched into that XCD's L2, except at
// the eighth's two ends (the per-block partial sums stay indexed by
// blockIdx).
struct AdjRange {
    int64_t first, end, stride;
};
__device__ inline AdjRange adj_range(int64_t n, int remap)
{
    const int64_t nb = gridDim.x;
    if (remap && (nb & 7) == 0) {
        const int64_t j = blockIdx.x & 7, lb = blockIdx.x >> 3, per = (n + 7) >> 3;
        const int64_t e = (j + 1) * per;
        return {j * per + lb * blockDim.x + threadIdx.x, e < n ? e : n, (nb >> 3) * blockDim.x};
    }
    return {(int64_t)blockIdx.x * blockDim.x + threadIdx.x, n, nb * blockDim.x};
}
#define ZSC_ADJ_RANGE(idx, n, remap)                                             \
    for (AdjRange rg_ = adj_range((n), (remap)), *rp_ = &rg_; rp_; rp_ = nullptr) \
        for (int64_t idx = rg_.first; idx < rg_.end; idx += rg_.stride)

// Pass 1: lam_out (local part, including phi's local uses), the stencil-
// adjoint fields a_lap, a_grad (4 fields) and per-block coefficient
// partials.  A and lam are plain (26, n) lattices.  The state A, phi and all
// arithmetic are at precision R; the adjoint lattices and the fields are
// stored at precision L (L = R, or FP32 adjoints over FP64 states).
// FP64 states need more registers than the file holds at two waves per SIMD;
// the bound keeps two resident (spilling what does not fit) rather than one.
// FP32 states are held to three (168 registers): left free, the compiler
// takes a few more and runs two, 16% slower per adjoint step at 128^3.
template <class R>
constexpr int adj_waves()
{
    return sizeof(R) == 8 ? 2 : 3;
}

//
// adef (may be null): the next step's stencil term a = lap[a_lap] -
// div[a_grad], not yet added to lam's seven f populations (pass 2 of that
// step wrote it as one field instead); it is added here as each f value is
// loaded, in the adjoint-lattice precision, which is the sum pass 2 would
// have stored.
template <class R, class L, bool PULL, bool DEV = false, bool DEFER = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(adj_waves<R>())))
    k_adj_pass1(int nx, int ny, int nz, int64_t n, Coef<R> c, Coef<L> cl, int grav_on,
                const R *__restrict__ A, const R *__restrict__ phi, const L *__restrict__ lam,
                const L *__restrict__ adef, L *__restrict__ lam_out, L *__restrict__ fields,
                double *partial, int accumulate, int remap)
{
    R ap[kNumGrad];
#pragma unroll
    for (int j = 0; j < kNumGrad; ++j) ap[j] = R(0);
    ZSC_ADJ_RANGE(idx, n, remap)
    {
        int x, y, z;
        cell_xyz32((uint32_t)idx, nx, ny, x, y, z);
        // 32-bit element offsets (every (26, n) row is < 4 GiB, checked at
        // launch): scalar-base + vector-offset loads, fewer index registers
        const Nbr32 nb(x, y, z, nx, ny, nz);
        const uint32_t own = nb.at(0, 0, 0);
        R p[19], fa[7], ga[19], deta;
        L lf[7], lg[19], afa[7], aga[19], agr[3], aphi, alap;
#pragma unroll
        for (int d = 0; d < 19; ++d) p[d] = ld32z<0>(phi, nb.at(cx(d), cy(d), cz(d)));
        if constexpr (PULL) {
            // A is the post-collision lattice of the previous step: stream it
            const Lattice<R> B = {const_cast<R *>(A), const_cast<R *>(A) + n,
                                  const_cast<R *>(A) + 7 * n};
            pull_f32<R, R, 0>(B, n, nb, c, fa);
            pull_g32<R, R, 0>(B, n, nb, c, ga);
        } else {
#pragma unroll
            for (int i = 0; i < 7; ++i) fa[i] = ld32z<0>(A + (int64_t)i * n, own);
#pragma unroll
            for (int i = 0; i < 19; ++i) ga[i] = ld32z<0>(A + (int64_t)(7 + i) * n, own);
        }
        // S^T: gather the next adjoint from downwind
        R lfd[7];
        lfd[0] = R(0);
        if constexpr (DEFER) {
            const L a0 = ld32z<0>(adef, own);
            lf[0] = ld32z<0>(lam, own) + a0;
#pragma unroll
            for (int i = 1; i < 7; ++i) {
                const L *li = lam + (int64_t)i * n;
                const uint32_t od = nb.at(cx(i), cy(i), cz(i));
                const L mine = ld32z<0>(li, own) + a0;
                const L down = ld32z<0>(li, od) + ld32z<0>(adef, od);
                lf[i] = cl.eta * down + cl.eta2 * mine;
                lfd[i] = (R)down - (R)mine;
            }
        } else {
            lf[0] = ld32z<0>(lam, own);
#pragma unroll
            for (int i = 1; i < 7; ++i) {
                const L *li = lam + (int64_t)i * n;
                const L mine = ld32z<0>(li, own);
                const L down = ld32z<0>(li, nb.at(cx(i), cy(i), cz(i)));
                lf[i] = cl.eta * down + cl.eta2 * mine;
                lfd[i] = (R)down - (R)mine;  // for the eta derivative below
            }
        }
#pragma unroll
        for (int i = 0; i < 19; ++i)
            lg[i] = ld32z<0>(lam + (int64_t)(7 + i) * n, nb.at(cx(i), cy(i), cz(i)));
        adj_cell<R, L, DEV>(c, cl, p, fa, ga, lf, lfd, lg, afa, aga, aphi, alap, agr, deta, ap,
                            grav_on);
        ap[kGEta] += deta;
#pragma unroll
        for (int i = 0; i < 7; ++i) st32z<0>(lam_out + (int64_t)i * n, own, (L)(afa[i] + aphi));
#pragma unroll
        for (int i = 0; i < 19; ++i) st32z<0>(lam_out + (int64_t)(7 + i) * n, own, (L)aga[i]);
        st32z<0>(fields, own, (L)alap);
        st32z<0>(fields + n, own, (L)agr[0]);
        st32z<0>(fields + 2 * n, own, (L)agr[1]);
        st32z<0>(fields + 3 * n, own, (L)agr[2]);
    }
    double apd[kNumGrad];
#pragma unroll
    for (int j = 0; j < kNumGrad; ++j) apd[j] = (double)ap[j];
    block_reduce_grads<256>(apd, partial, accumulate != 0);
}

// Pass 2: a_phi's stencil part a = lap[a_lap] - div[a_grad] (the
// transposed stencils), added to the adjoint of every f population, or, with
// DEFER, written as one field (aout) for the next backward step's pass 1 to
// add as it loads those populations: 20 bytes per cell instead of 72.
// Tiles of 64 x TY cells march along z; the four stencil fields of planes
// k-1, k, k+1 (tile and 1-cell ring) sit in a 4-slot LDS ring, one barrier
// per plane, so each field value is loaded from memory about once instead
// of once per stencil point.  Row offsets are 32-bit (rows < 4 GiB, checked
// at launch).
template <class R, int TY, bool DEFER = false>
__global__ void __launch_bounds__(64 * TY)
    k_adj_pass2(int nx, int ny, int nz, int64_t n, int zchunk, int nbx, int nby, int nblk,
                int remap, const R *__restrict__ fields, R *__restrict__ lam_out,
                R *__restrict__ aout)
{
    __shared__ R sh[4][4][TY + 2][66];  // [slot][field][row][column]
    const int b = blockIdx.x;
    int l = b;
    if (remap && (nblk & 7) == 0) l = (b & 7) * (nblk >> 3) + (b >> 3);
    const int bx = l % nbx, by = (l / nbx) % nby, bz = l / (nbx * nby);
    const int x0 = bx * 64, y0 = by * TY, z0 = bz * zchunk;
    const int z1 = min(nz, z0 + zchunk);
    const int tw = min(64, nx - x0), th = min(TY, ny - y0);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const bool own = tx < tw && ty < th;
    const int x = x0 + min(tx, tw - 1), y = y0 + min(ty, th - 1);
    const int nh = 2 * tw + 2 * (th + 2);  // ring cells per plane
    const int tid = ty * 64 + tx;
    const uint32_t nxy = (uint32_t)nx * (uint32_t)ny;
    // Each thread loads at most one ring cell (nh <= 2 * 64 + 2 * (TY + 2)
    // <= 64 * TY): its position in the plane and in LDS are the same for
    // every plane, so they are worked out once.
    static_assert(TY >= 3, "one ring cell per thread");
    const bool ring = tid < nh;
    int rr = 0, rc = 0;
    uint32_t roff = 0;
    if (ring) {
        const int h = tid;
        int hx, hy;
        if (h < tw) {  // row below the tile
            rc = h + 1;
            rr = 0;
            hx = x0 + h;
            hy = y0 - 1;
        } else if (h < 2 * tw) {  // row above
            rc = h - tw + 1;
            rr = th + 1;
            hx = x0 + h - tw;
            hy = y0 + th;
        } else if (h < 2 * tw + th + 2) {  // column left, corners included
            rc = 0;
            rr = h - 2 * tw;
            hx = x0 - 1;
            hy = y0 - 1 + rr;
        } else {  // column right
            rc = tw + 1;
            rr = h - 2 * tw - th - 2;
            hx = x0 + tw;
            hy = y0 - 1 + rr;
        }
        hx = hx < 0 ? hx + nx : (hx >= nx ? hx - nx : hx);
        hy = hy < 0 ? hy + ny : (hy >= ny ? hy - ny : hy);
        roff = (uint32_t)hy * (uint32_t)nx + (uint32_t)hx;
    }
    const uint32_t ooff = (uint32_t)y * (uint32_t)nx + (uint32_t)x;

    // the four fields of plane zz (tile + ring) into its slot
    auto load_plane = [&](int zz) {
        const int zc = zz < 0 ? zz + nz : (zz >= nz ? zz - nz : zz);
        const int slot = (zz + 4) & 3;
        const uint32_t pz = (uint32_t)zc * nxy;
        if (own) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                sh[slot][q][ty + 1][tx + 1] = ld32z<0>(fields + q * n, pz + ooff);
        }
        if (ring) {
#pragma unroll
            for (int q = 0; q < 4; ++q) sh[slot][q][rr][rc] = ld32z<0>(fields + q * n, pz + roff);
        }
    };

    load_plane(z0 - 1);
    load_plane(z0);
    for (int k = z0; k < z1; ++k) {
        const uint32_t o = (uint32_t)k * nxy + ooff;
        R lf[7];
        if (own && !DEFER) {
#pragma unroll
            for (int i = 0; i < 7; ++i) lf[i] = ld32z<0>(lam_out + (int64_t)i * n, o);
        }
        load_plane(k + 1);
        __syncthreads();
        if (own) {
            // slots of planes k-1, k, k+1
            const int sl[3] = {(k + 3) & 3, (k + 4) & 3, (k + 5) & 3};
            auto gather = [&](int q, R *h) {
#pragma unroll
                for (int d = 0; d < 19; ++d)
                    h[d] = sh[sl[1 + cz(d)]][q][ty + 1 + cy(d)][tx + 1 + cx(d)];
            };
            R h[19];
            gather(0, h);
            R a = lap19(h);
            gather(1, h);
            R dv = R(2) * (h[1] - h[2]) + h[7] - h[8] + h[9] - h[10] + h[11] - h[12] + h[13] - h[14];
            gather(2, h);
            dv += R(2) * (h[3] - h[4]) + h[7] - h[8] + h[10] - h[9] + h[15] - h[16] + h[17] - h[18];
            gather(3, h);
            dv += R(2) * (h[5] - h[6]) + h[11] - h[12] + h[14] - h[13] + h[15] - h[16] + h[18] - h[17];
            a -= dv * W<R>::inv12;
            if constexpr (DEFER) {
                st32z<0>(aout, o, a);
            } else {
#pragma unroll
                for (int i = 0; i < 7; ++i) st32z<0>(lam_out + (int64_t)i * n, o, lf[i] + a);
            }
        }
    }
}

// Loss seed: L = sum (phi - phi_tgt)^2, lam_f,i = 2 (phi - phi_tgt), lam_g = 0.
template <class R, class L>
__global__ void __launch_bounds__(256)
    k_adj_seed(int64_t n, const R *__restrict__ A, const R *__restrict__ tgt, L *__restrict__ lam,
               double *partial)
{
    double ap[kNumGrad];
#pragma unroll
    for (int j = 0; j < kNumGrad; ++j) ap[j] = 0.0;
    ZSC_GRID_STRIDE(idx, n)
    {
        R f[7];
#pragma unroll
        for (int i = 0; i < 7; ++i) f[i] = A[(int64_t)i * n + idx];
        const R d = sum_f(f) - tgt[idx];
        ap[kGLoss] += (double)d * (double)d;
#pragma unroll
        for (int i = 0; i < 7; ++i) lam[(int64_t)i * n + idx] = (L)(R(2) * d);
#pragma unroll
        for (int i = 0; i < 19; ++i) lam[(int64_t)(7 + i) * n + idx] = L(0);
    }
    block_reduce_grads<256>(ap, partial);
}

// Adjoint of the initial equilibria (src/init_f.cu:69-78, src/init_g.cu:74-91)
// with respect to alpha4 and kappa, at fixed phi0 and rho = rho_mean.
template <class R, class L>
__global__ void __launch_bounds__(256)
    k_adj_init(int nx, int ny, int nz, int64_t n, Coef<R> c, const R *__restrict__ phi0,
               const L *__restrict__ lam, double *partial)
{
    double ap[kNumGrad];
#pragma unroll
    for (int j = 0; j < kNumGrad; ++j) ap[j] = 0.0;
    ZSC_GRID_STRIDE(idx, n)
    {
        int x, y, z;
        cell_xyz(idx, nx, ny, x, y, z);
        Nbr nb(x, y, z, nx, ny, nz);
        R p[19];
        gather19(phi0, nb, p);
        const R phi = p[0];
        auto lm = [&](int64_t i) { return (R)lam[i]; };
        R a_mu = R(-3) * c.gamma * lm(idx);
#pragma unroll
        for (int i = 1; i < 7; ++i) a_mu += R(0.5) * c.gamma * lm((int64_t)i * n + idx);
        a_mu += R(-6) * W<R>::w0 * phi * lm(7 * n + idx);
#pragma unroll
        for (int i = 1; i < 7; ++i) a_mu += R(3) * W<R>::w1 * phi * lm((int64_t)(7 + i) * n + idx);
#pragma unroll
        for (int i = 7; i < 19; ++i) a_mu += R(3) * W<R>::w2 * phi * lm((int64_t)(7 + i) * n + idx);
        ap[kGAlpha4] += (double)(phi * (phi * phi - c.phi_star_sq) * a_mu);
        ap[kGKappa] += (double)(-lap19(p) * a_mu);
    }
    block_reduce_grads<256>(ap, partial);
}

// grads[j] += sum over blocks of partial[b][j] (fixed order).
// grads[j] += sum over blocks of partial[b][j]: 256 threads, strided sums in
// a fixed order, then a fixed-shape shuffle tree (deterministic).
template <class R>
__global__ void __launch_bounds__(256) k_adj_accumulate(int nblocks, const double *partial, double *grads)
{
    double ap[kNumGrad];
#pragma unroll
    for (int j = 0; j < kNumGrad; ++j) ap[j] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256)
#pragma unroll
        for (int j = 0; j < kNumGrad; ++j) ap[j] += partial[(int64_t)b * kNumGrad + j];
    __shared__ double red[4][kNumGrad];
#pragma unroll
    for (int j = 0; j < kNumGrad; ++j) {
        double v = ap[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        ap[j] = v;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int j = 0; j < kNumGrad; ++j) red[wv][j] = ap[j];
    __syncthreads();
    if (threadIdx.x < kNumGrad) {
        const int j = threadIdx.x;
        grads[j] += (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
    }
}

// R: precision of the states, phi and the arithmetic; L: storage of the
// adjoint lattices and stencil fields (R, or float over double states).
template <class R, class L>
struct Adj {
    // scratch: 4n L stencil fields | n L deferred stencil term | partials |
    // deferred partials.  A step's pass 1 reads the term of the step after
    // it before its pass 2 writes its own there.
    static size_t field_bytes(int64_t n) { return align256((size_t)n * 5 * sizeof(L)); }
    static L *term(const Geo &g, void *scratch)
    {
        return (L *)scratch + (size_t)g.n * 4;
    }
    static size_t scratch_bytes(int64_t n)
    {
        return field_bytes(n) + 2 * (size_t)kAdjMaxBlocks * kNumGrad * sizeof(double);
    }
    static double *partials(const Geo &g, void *scratch, int which)
    {
        return (double *)((char *)scratch + field_bytes(g.n)) + (size_t)which * kAdjMaxBlocks * kNumGrad;
    }
    // A: the step's input state, or (pulled) the previous post-collision
    // lattice, streamed on the fly; dev: its g are deviations from w_i
    // rho_mean (FP32 only).
    // grads == nullptr defers the coefficient sums into per-block partials
    // kept in the scratch buffer until flush(); defer_first: no deferred
    // sums exist yet, so the partials are written rather than added to.
    // a_in (may be null): lam's deferred stencil term; a_out (may be null):
    // where this step leaves its own stencil term instead of adding it to
    // lam_out (then lam_out's f populations lack it).
    static hipError_t backward(const Geo &g, const zsc_params &p, int grav_on, hipStream_t s,
                               const R *A, bool pulled, const R *phi, const L *lam, L *lam_out,
                               void *scratch, double *grads, bool defer_first, bool dev = false,
                               const L *a_in = nullptr, L *a_out = nullptr)
    {
        if ((uint64_t)g.n * sizeof(R) > ((uint64_t)1 << 32)) return hipErrorInvalidValue;
        L *fields = (L *)scratch;
        const int acc = grads || defer_first ? 0 : 1;
        const int nb = adj_blocks(g.n);
        double *partial = partials(g, scratch, grads ? 0 : 1);
        const Coef<R> c = make_coef<R>(p, grav_on);
        const Coef<L> cl = make_coef<L>(p, grav_on);
        const int remap = 1;
        auto pass1 = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(nb), dim3(256), 0, s, g.nx, g.ny, g.nz, g.n, c, cl,
                               grav_on, A, phi, lam, a_in, lam_out, fields, partial, acc, remap);
        };
        auto pick = [&](auto dev_tag) {
            constexpr bool D = decltype(dev_tag)::value;
            if (a_in)
                pulled ? pass1(k_adj_pass1<R, L, true, D, true>) : pass1(k_adj_pass1<R, L, false, D, true>);
            else
                pulled ? pass1(k_adj_pass1<R, L, true, D>) : pass1(k_adj_pass1<R, L, false, D>);
        };
        if (!dev)
            pick(std::false_type{});
        else if constexpr (sizeof(R) == 4)  // deviation states exist at FP32 only
            pick(std::true_type{});
        else
            return hipErrorInvalidValue;
        {
            constexpr int TY = 4;
            const int nbx = (g.nx + 63) / 64, nby = (g.ny + TY - 1) / TY;
            int zchunk = 16;
            while ((int64_t)nbx * nby * ((g.nz + zchunk - 1) / zchunk) < 2048 && zchunk > 4)
                zchunk /= 2;
            const int64_t nblk = (int64_t)nbx * nby * ((g.nz + zchunk - 1) / zchunk);
            if (nblk > 0x7fffffff) return hipErrorInvalidValue;
            if (a_out)
                hipLaunchKernelGGL((k_adj_pass2<L, TY, true>), dim3((unsigned)nblk), dim3(64, TY), 0,
                                   s, g.nx, g.ny, g.nz, g.n, zchunk, nbx, nby, (int)nblk, remap,
                                   (const L *)fields, lam_out, a_out);
            else
                hipLaunchKernelGGL((k_adj_pass2<L, TY>), dim3((unsigned)nblk), dim3(64, TY), 0, s,
                                   g.nx, g.ny, g.nz, g.n, zchunk, nbx, nby, (int)nblk, remap,
                                   (const L *)fields, lam_out, (L *)nullptr);
        }
        if (grads)
            hipLaunchKernelGGL(k_adj_accumulate<R>, dim3(1), dim3(256), 0, s, nb, partial, grads);
        return hipGetLastError();
    }
    // grads += deferred partials.
    static hipError_t flush(const Geo &g, hipStream_t s, void *scratch, double *grads)
    {
        hipLaunchKernelGGL(k_adj_accumulate<R>, dim3(1), dim3(256), 0, s, adj_blocks(g.n),
                           partials(g, scratch, 1), grads);
        return hipGetLastError();
    }
    static hipError_t seed(const Geo &g, hipStream_t s, const R *A, const R *tgt, L *lam,
                           void *scratch, double *grads)
    {
        double *partial = partials(g, scratch, 0);
        const int nb = adj_blocks(g.n);
        hipLaunchKernelGGL((k_adj_seed<R, L>), dim3(nb), dim3(256), 0, s, g.n, A, tgt, lam, partial);
        hipLaunchKernelGGL(k_adj_accumulate<R>, dim3(1), dim3(256), 0, s, nb, partial, grads);
        return hipGetLastError();
    }
    static hipError_t init_backward(const Geo &g, const zsc_params &p, hipStream_t s, const R *phi0,
                                    const L *lam, void *scratch, double *grads)
    {
        double *partial = partials(g, scratch, 0);
        const int nb = adj_blocks(g.n);
        hipLaunchKernelGGL((k_adj_init<R, L>), dim3(nb), dim3(256), 0, s, g.nx, g.ny, g.nz, g.n,
                           make_coef<R>(p, 0), phi0, lam, partial);
        hipLaunchKernelGGL(k_adj_accumulate<R>, dim3(1), dim3(256), 0, s, nb, partial, grads);
        return hipGetLastError();
    }
};

}  // namespace zsc
