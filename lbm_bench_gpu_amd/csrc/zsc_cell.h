// zsc_cell.h -- per-cell ZSC arithmetic shared by every kernel of libzsc.
//
// The statements keep the operand order of the reference kernels (cited per
// function).  In the FP64 translation unit, built with -ffp-contract=off, each
// operation is then rounded exactly as an IEEE evaluation of the reference's
// expressions, which is what FP64 parity (and the rounding-sensitive gravity
// switch phi > 0, src/collision_f.cu:96) relies on.  The FP32 unit lets the
// compiler contract into FMAs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zsc {

// D3Q19 velocities (g); D3Q7 (f) uses directions 0..6.  Order follows the
// streaming targets of src/stream_g.cu:56-76 and src/stream_f.cu:52-59.
// Kept as constexpr functions so the tables fold into immediates.
__host__ __device__ constexpr int cx(int d)
{
    return d == 1 || d == 7 || d == 9 || d == 11 || d == 13 ? 1
           : d == 2 || d == 8 || d == 10 || d == 12 || d == 14 ? -1 : 0;
}
__host__ __device__ constexpr int cy(int d)
{
    return d == 3 || d == 7 || d == 10 || d == 15 || d == 17 ? 1
           : d == 4 || d == 8 || d == 9 || d == 16 || d == 18 ? -1 : 0;
}
__host__ __device__ constexpr int cz(int d)
{
    return d == 5 || d == 11 || d == 14 || d == 15 || d == 18 ? 1
           : d == 6 || d == 12 || d == 13 || d == 16 || d == 17 ? -1 : 0;
}

// Scalars of one step, at compute precision R.  The derivation lives on the
// host (src/paramread.c:141-170); grav is 0 while gravity is switched off.
template <class R>
struct Coef {
    R phi_star_sq, alpha4, kappa, gamma;
    R eta, eta2, inv_eta2;
    R inv_tau_phi, inv_tau_phi_one, inv_tau_rho_one;
    R k0, k1, k2, kc0, kc1, kc2;
    R grav;
    R rho_mean;
};

// Lattice weights of src/constants.h:43-59 at precision R.
template <class R> struct W;
template <> struct W<double> {
    static constexpr double inv6 = 1.0 / 6.0, inv12 = 1.0 / 12.0, inv_cs_sq = 3.0,
                            cs_sq = 1.0 / 3.0, w0 = 1.0 / 3.0, w1 = 1.0 / 18.0,
                            w2 = 1.0 / 36.0;
};
template <> struct W<float> {
    static constexpr float inv6 = 1.0f / 6.0f, inv12 = 1.0f / 12.0f, inv_cs_sq = 3.0f,
                           cs_sq = 1.0f / 3.0f, w0 = 1.0f / 3.0f, w1 = 1.0f / 18.0f,
                           w2 = 1.0f / 36.0f;
};

// Quotients of the step.  FP64 divides exactly, which FP64 parity needs;
// FP32 multiplies by the hardware reciprocal (v_rcp_f32, 1 ulp), as the
// reference's FP32 build approximates its divisions (-use_fast_math,
// src/Makefile:5), and folds the constant 1/6.
template <class R>
struct Q {
    __host__ __device__ static R recip(R b) { return R(1) / b; }
    __host__ __device__ static R div(R a, R b) { return a / b; }
    __host__ __device__ static R div6(R a) { return a / R(6); }
};
template <>
struct Q<float> {
    __host__ __device__ static float recip(float b)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_rcpf(b);
#else
        return 1.0f / b;
#endif
    }
    __host__ __device__ static float div(float a, float b) { return a * recip(b); }
    __host__ __device__ static float div6(float a) { return a * (1.0f / 6.0f); }
};

// Rest value of g_i used by the 16-bit deviation codecs: w_i * rho_mean.
template <class R>
__host__ __device__ inline R g_weight(int i)
{
    return i == 0 ? W<R>::w0 : (i < 7 ? W<R>::w1 : W<R>::w2);
}

// p[d] = phi(x + c_d).  Laplacian of src/collision_f.cu:61-70 (identical in
// src/init_f.cu:57-66, src/init_g.cu:62-71, src/collision_g.cu:66-75).
template <class R>
__host__ __device__ inline R lap19(const R *p)
{
    return (p[7] + p[8] + p[9] + p[10] + p[11] + p[12] + p[13] + p[14] + p[15] + p[16] +
            p[17] + p[18] + R(2) * (p[1] + p[2] + p[3] + p[4] + p[5] + p[6] - R(12) * p[0])) *
           W<R>::inv6;
}

// mu = 4 alpha phi (phi^2 - phi*^2) - kappa lap (src/collision_f.cu:76).
template <class R>
__host__ __device__ inline R chem_pot(const Coef<R> &c, R phin, R lap)
{
    return c.alpha4 * phin * (phin * phin - c.phi_star_sq) - c.kappa * lap;
}

// F = mu grad(phi) (+ grav on phi > 0 cells), src/collision_f.cu:78-96.
template <class R>
__host__ __device__ inline void force19(const Coef<R> &c, const R *p, R &mu, R &fx, R &fy,
                                        R &fz)
{
    const R phin = p[0];
    mu = chem_pot(c, phin, lap19(p));
    fx = mu * (R(2) * (p[1] - p[2]) + p[7] - p[8] + p[9] - p[10] + p[11] - p[12] + p[13] -
               p[14]) * W<R>::inv12;
    fy = mu * (R(2) * (p[3] - p[4]) + p[7] - p[8] + p[10] - p[9] + p[15] - p[16] + p[17] -
               p[18]) * W<R>::inv12;
    fz = mu * (R(2) * (p[5] - p[6]) + p[11] - p[12] + p[14] - p[13] + p[15] - p[16] + p[18] -
               p[17]) * W<R>::inv12;
    if (phin > R(0)) fz = fz + c.grav;
}

// Pressure of src/vtksave.c:75-119 (the same expression as src/logsave.c:
// 84-137): alpha (phi^2 (3 phi^2 - 2 phi*^2) - phi*^4)
// - kappa (phi lap(phi) - |grad phi|^2 / 2) + cs^2 rho.
template <class R>
__host__ __device__ inline R pressure19(R alpha, R phi_star_sq, R kappa, const R *p, R rho)
{
    const R phin = p[0], phin2 = phin * phin;
    const R lap = lap19(p);
    const R gx = (R(2) * (p[1] - p[2]) + p[7] - p[8] + p[9] - p[10] + p[11] - p[12] + p[13] -
                  p[14]) * W<R>::inv12;
    const R gy = (R(2) * (p[3] - p[4]) + p[7] - p[8] + p[10] - p[9] + p[15] - p[16] + p[17] -
                  p[18]) * W<R>::inv12;
    const R gz = (R(2) * (p[5] - p[6]) + p[11] - p[12] + p[14] - p[13] + p[15] - p[16] + p[18] -
                  p[17]) * W<R>::inv12;
    const R gsq = gx * gx + gy * gy + gz * gz;
    return alpha * (phin2 * (R(3) * phin2 - R(2) * phi_star_sq) - phi_star_sq * phi_star_sq) -
           kappa * (phin * lap - R(0.5) * gsq) + W<R>::cs_sq * rho;
}

// phi = f0 + ... + f6 (src/update_phi.cu:42-43).
template <class R>
__host__ __device__ inline R sum_f(const R *f)
{
    return f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6];
}

// rho = g0 + ... + g18 (src/update_rho.cu:46-50).
template <class R>
__host__ __device__ inline R sum_g(const R *g)
{
    R s = g[0];
#pragma unroll
    for (int i = 1; i < 19; ++i) s = s + g[i];
    return s;
}

// u = sum c g / rho with invRho = 1/rho (src/update_velocity.cu:47-59).
template <class R>
__host__ __device__ inline void velocity(const R *g, R rho, R &ux, R &uy, R &uz)
{
    const R inv_rho = Q<R>::recip(rho);
    ux = (g[1] - g[2] + g[7] - g[8] + g[9] - g[10] + g[11] - g[12] + g[13] - g[14]) * inv_rho;
    uy = (g[3] - g[4] + g[7] - g[8] - g[9] + g[10] + g[15] - g[16] + g[17] - g[18]) * inv_rho;
    uz = (g[5] - g[6] + g[11] - g[12] - g[13] + g[14] + g[15] - g[16] - g[17] + g[18]) * inv_rho;
}

// D3Q7 collision of src/collision_f.cu:100-114 (u' uses invRho = 1/rho).
template <class R>
__host__ __device__ inline void collide_f(const Coef<R> &c, R phin, R rho, R ux, R uy, R uz,
                                          R mu, R fx, R fy, R fz, const R *fa, R *fb)
{
    const R inv_rho = Q<R>::recip(rho);
    const R uxn = ux + R(0.5) * fx * inv_rho;
    const R uyn = uy + R(0.5) * fy * inv_rho;
    const R uzn = uz + R(0.5) * fz * inv_rho;
    const R af = R(0.5) * c.gamma * c.inv_tau_phi * mu;
    const R cf = c.inv_tau_phi * c.inv_eta2 * phin;
    const R t = c.inv_tau_phi_one;
    fb[0] = t * fa[0] - R(6) * af + c.inv_tau_phi * phin;
    fb[1] = t * fa[1] + af + cf * uxn;
    fb[2] = t * fa[2] + af - cf * uxn;
    fb[3] = t * fa[3] + af + cf * uyn;
    fb[4] = t * fa[4] + af - cf * uyn;
    fb[5] = t * fa[5] + af + cf * uzn;
    fb[6] = t * fa[6] + af - cf * uzn;
}

// D3Q19 collision with forcing of src/collision_g.cu:105-194 (u' divides by
// rho).  The axis pairs keep the reference's eq2 = u (K1 rho + KC1 F) and its
// two operand orders K1*rho*u*u*0.5 (x) and K1*rho*u*0.5*u (y, z).
template <class R>
__host__ __device__ inline void collide_g(const Coef<R> &c, R phin, R rhon, R ux, R uy, R uz,
                                          R mu, R fx, R fy, R fz, const R *ga, R *gb)
{
    const R K1 = c.k1, K2 = c.k2, KC1 = c.kc1, KC2 = c.kc2, t = c.inv_tau_rho_one;
    R uxn = ux + Q<R>::div(R(0.5) * fx, rhon);
    R uyn = uy + Q<R>::div(R(0.5) * fy, rhon);
    R uzn = uz + Q<R>::div(R(0.5) * fz, rhon);
    R temp = K1 * (R(3) * phin * mu + rhon * (R(1) - R(1.5) * (uxn * uxn + uyn * uyn + uzn * uzn)));
    const R uf = uxn * fx + uyn * fy + uzn * fz;
    uxn = uxn * W<R>::inv_cs_sq;
    uyn = uyn * W<R>::inv_cs_sq;
    uzn = uzn * W<R>::inv_cs_sq;
    R eq1, eq2;

    gb[0] = t * ga[0] +
            c.k0 * (rhon - R(6) * phin * mu - Q<R>::div6(rhon * (uxn * uxn + uyn * uyn + uzn * uzn))) -
            c.kc0 * uf;

    eq1 = temp + K1 * rhon * uxn * uxn * R(0.5) + KC1 * (uxn * fx - uf);
    eq2 = uxn * (K1 * rhon + KC1 * fx);
    gb[1] = t * ga[1] + eq1 + eq2;
    gb[2] = t * ga[2] + eq1 - eq2;

    eq1 = temp + K1 * rhon * uyn * R(0.5) * uyn + KC1 * (uyn * fy - uf);
    eq2 = uyn * (K1 * rhon + KC1 * fy);
    gb[3] = t * ga[3] + eq1 + eq2;
    gb[4] = t * ga[4] + eq1 - eq2;

    eq1 = temp + K1 * rhon * uzn * R(0.5) * uzn + KC1 * (uzn * fz - uf);
    eq2 = uzn * (K1 * rhon + KC1 * fz);
    gb[5] = t * ga[5] + eq1 + eq2;
    gb[6] = t * ga[6] + eq1 - eq2;

    temp = K2 * (R(3) * phin * mu + rhon * (R(1) - Q<R>::div6(uxn * uxn + uyn * uyn + uzn * uzn)));

    auto diag = [&](int ip, int im, R s, R S) {
        const R e1 = temp + K2 * rhon * s * R(0.5) * s + KC2 * (s * S - uf);
        const R e2 = K2 * rhon * s + KC2 * S;
        gb[ip] = t * ga[ip] + e1 + e2;
        gb[im] = t * ga[im] + e1 - e2;
    };
    diag(7, 8, uxn + uyn, fx + fy);
    diag(9, 10, uxn - uyn, fx - fy);
    diag(11, 12, uxn + uzn, fx + fz);
    diag(13, 14, uxn - uzn, fx - fz);
    diag(15, 16, uyn + uzn, fy + fz);
    diag(17, 18, uyn - uzn, fy - fz);
}

// collide_g on deviations dg_i = g_i - w_i rho_mean (16-bit storage).
// Because K_i = w_i / tau and 1 - t = 1 / tau, the rest state cancels
// exactly: each "rho * 1" term of the equilibria becomes drho = rho -
// rho_mean and g_B - w_i rho_mean = t dg_A + (the same equilibrium with drho).
// Terms that multiply rho by u or F keep the full rho.
template <class R>
__host__ __device__ inline void collide_g_dev(const Coef<R> &c, R phin, R rhon, R drho, R ux, R uy,
                                              R uz, R mu, R fx, R fy, R fz, const R *ga, R *gb)
{
    const R K1 = c.k1, K2 = c.k2, KC1 = c.kc1, KC2 = c.kc2, t = c.inv_tau_rho_one;
    R uxn = ux + Q<R>::div(R(0.5) * fx, rhon);
    R uyn = uy + Q<R>::div(R(0.5) * fy, rhon);
    R uzn = uz + Q<R>::div(R(0.5) * fz, rhon);
    R temp = K1 * (R(3) * phin * mu + drho - rhon * R(1.5) * (uxn * uxn + uyn * uyn + uzn * uzn));
    const R uf = uxn * fx + uyn * fy + uzn * fz;
    uxn = uxn * W<R>::inv_cs_sq;
    uyn = uyn * W<R>::inv_cs_sq;
    uzn = uzn * W<R>::inv_cs_sq;
    R eq1, eq2;
    gb[0] = t * ga[0] +
            c.k0 * (drho - R(6) * phin * mu - Q<R>::div6(rhon * (uxn * uxn + uyn * uyn + uzn * uzn))) -
            c.kc0 * uf;
    eq1 = temp + K1 * rhon * uxn * uxn * R(0.5) + KC1 * (uxn * fx - uf);
    eq2 = uxn * (K1 * rhon + KC1 * fx);
    gb[1] = t * ga[1] + eq1 + eq2;
    gb[2] = t * ga[2] + eq1 - eq2;
    eq1 = temp + K1 * rhon * uyn * R(0.5) * uyn + KC1 * (uyn * fy - uf);
    eq2 = uyn * (K1 * rhon + KC1 * fy);
    gb[3] = t * ga[3] + eq1 + eq2;
    gb[4] = t * ga[4] + eq1 - eq2;
    eq1 = temp + K1 * rhon * uzn * R(0.5) * uzn + KC1 * (uzn * fz - uf);
    eq2 = uzn * (K1 * rhon + KC1 * fz);
    gb[5] = t * ga[5] + eq1 + eq2;
    gb[6] = t * ga[6] + eq1 - eq2;
    temp = K2 * (R(3) * phin * mu + drho - Q<R>::div6(rhon * (uxn * uxn + uyn * uyn + uzn * uzn)));
    auto diag = [&](int ip, int im, R s, R S) {
        const R e1 = temp + K2 * rhon * s * R(0.5) * s + KC2 * (s * S - uf);
        const R e2 = K2 * rhon * s + KC2 * S;
        gb[ip] = t * ga[ip] + e1 + e2;
        gb[im] = t * ga[im] + e1 - e2;
    };
    diag(7, 8, uxn + uyn, fx + fy);
    diag(9, 10, uxn - uyn, fx - fy);
    diag(11, 12, uxn + uzn, fx + fz);
    diag(13, 14, uxn - uzn, fx - fz);
    diag(15, 16, uyn + uzn, fy + fz);
    diag(17, 18, uyn - uzn, fy - fz);
}

// collide_cell with g given and returned as deviations from w_i rho_mean.
template <class R>
__host__ __device__ inline void collide_cell_dev(const Coef<R> &c, const R *p, const R *fa,
                                                 const R *dga, R *fb, R *dgb)
{
    R mu, fx, fy, fz, ux, uy, uz;
    const R drho = sum_g(dga);
    const R rho = c.rho_mean + drho;
    velocity(dga, rho, ux, uy, uz);  // sum_i c_i w_i = 0: the rest state carries no momentum
    force19(c, p, mu, fx, fy, fz);
    collide_f(c, p[0], rho, ux, uy, uz, mu, fx, fy, fz, fa, fb);
    collide_g_dev(c, p[0], rho, drho, ux, uy, uz, mu, fx, fy, fz, dga, dgb);
}

// Equilibria at u = 0 of src/init_f.cu:69-78 and src/init_g.cu:74-91.
template <class R>
__host__ __device__ inline void init_f_cell(const Coef<R> &c, const R *p, R *f)
{
    const R phin = p[0];
    const R mu = chem_pot(c, phin, lap19(p));
    const R af = R(0.5) * c.gamma * mu;
    f[0] = phin - R(3) * c.gamma * mu;
#pragma unroll
    for (int i = 1; i < 7; ++i) f[i] = af;
}

template <class R>
__host__ __device__ inline void init_g_cell(const Coef<R> &c, const R *p, R rhon, R *g)
{
    const R phin = p[0];
    const R mu = chem_pot(c, phin, lap19(p));
    const R ag1 = W<R>::w1 * (R(3) * phin * mu + rhon);
    const R ag2 = W<R>::w2 * (R(3) * phin * mu + rhon);
    g[0] = W<R>::w0 * (rhon - R(6) * phin * mu);
#pragma unroll
    for (int i = 1; i < 7; ++i) g[i] = ag1;
#pragma unroll
    for (int i = 7; i < 19; ++i) g[i] = ag2;
}

// init_g_cell as deviations g_i - w_i rho_mean, from drho = rho - rho_mean:
// the "rho" of each equilibrium becomes drho (deviation storage).
template <class R>
__host__ __device__ inline void init_g_cell_dev(const Coef<R> &c, const R *p, R drho, R *g)
{
    const R phin = p[0];
    const R mu = chem_pot(c, phin, lap19(p));
    const R ag1 = W<R>::w1 * (R(3) * phin * mu + drho);
    const R ag2 = W<R>::w2 * (R(3) * phin * mu + drho);
    g[0] = W<R>::w0 * (drho - R(6) * phin * mu);
#pragma unroll
    for (int i = 1; i < 7; ++i) g[i] = ag1;
#pragma unroll
    for (int i = 7; i < 19; ++i) g[i] = ag2;
}

// Full cell update from post-stream populations and the phi stencil:
// update_rho / update_velocity, then collision_f and collision_g.
template <class R>
__host__ __device__ inline void collide_cell(const Coef<R> &c, const R *p, const R *fa,
                                             const R *ga, R *fb, R *gb)
{
    R mu, fx, fy, fz, ux, uy, uz;
    const R rho = sum_g(ga);
    velocity(ga, rho, ux, uy, uz);
    force19(c, p, mu, fx, fy, fz);
    collide_f(c, p[0], rho, ux, uy, uz, mu, fx, fy, fz, fa, fb);
    collide_g(c, p[0], rho, ux, uy, uz, mu, fx, fy, fz, ga, gb);
}

}  // namespace zsc
