// zsc_store.h -- population storage codecs.
//
// FP64 and FP32 lattices store populations as they are.  The 16-bit
// lattices store the deviation from a per-population rest value (g_i:
// w_i * rho_mean; f_1..6: 0) rounded to nearest-even, and decode to FP32 for
// compute.  f_0 ~ phi is O(1) with no useful rest value, so the 16-bit
// lattices keep f_0 in FP32 (its own array, type S0 below).  The FP32
// deviation lattice (f32d) keeps f as it is and g as the FP32 deviation
// from w_i * rho_mean: where rho_mean is large against the density
// variations (rho_mean = 500.5 for a density ratio of 1000), a stored g_i
// then keeps the bits of the variation instead of those of the rest value.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zsc {

struct bf16s {
    uint16_t v;
};
struct f16s {
    uint16_t v;
};
struct f32d {
    float v;
};

template <class S>
struct Codec;

// dec/enc take the population's rest value; dec0/enc0 are the same for a
// rest value of exactly zero (no "+ 0.0", which IEEE arithmetic cannot drop).
template <>
struct Codec<double> {
    static constexpr bool deviation = false;
    __host__ __device__ static double dec(double s, double) { return s; }
    __host__ __device__ static double enc(double v, double) { return v; }
    __host__ __device__ static double dec0(double s) { return s; }
    __host__ __device__ static double enc0(double v) { return v; }
};

template <>
struct Codec<float> {
    static constexpr bool deviation = false;
    __host__ __device__ static float dec(float s, float) { return s; }
    __host__ __device__ static float enc(float v, float) { return v; }
    __host__ __device__ static float dec0(float s) { return s; }
    __host__ __device__ static float enc0(float v) { return v; }
};

// bf16: upper half of the FP32 pattern, rounded to nearest-even; NaN kept
// quiet so divergence still shows in the reductions.  gfx950 converts in one
// instruction (v_cvt_pk_bf16_f32, round-to-nearest-even).
__host__ __device__ inline uint16_t f32_to_bf16_rne(float f)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bit_cast(uint16_t, (__bf16)f);
#endif
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

template <>
struct Codec<bf16s> {
    static constexpr bool deviation = true;
    __host__ __device__ static float dec(bf16s s, float rest) { return bf16_to_f32(s.v) + rest; }
    __host__ __device__ static bf16s enc(float v, float rest) { return {f32_to_bf16_rne(v - rest)}; }
    __host__ __device__ static float dec0(bf16s s) { return bf16_to_f32(s.v); }
    __host__ __device__ static bf16s enc0(float v) { return {f32_to_bf16_rne(v)}; }
};

template <>
struct Codec<f16s> {
    static constexpr bool deviation = true;
    __host__ __device__ static float dec(f16s s, float rest)
    {
        _Float16 h;
        __builtin_memcpy(&h, &s.v, 2);
        return (float)h + rest;
    }
    __host__ __device__ static f16s enc(float v, float rest) { return enc0(v - rest); }
    __host__ __device__ static float dec0(f16s s)
    {
        _Float16 h;
        __builtin_memcpy(&h, &s.v, 2);
        return (float)h;
    }
    __host__ __device__ static f16s enc0(float v)
    {
        _Float16 h = (_Float16)v;
        f16s s;
        __builtin_memcpy(&s.v, &h, 2);
        return s;
    }
};

template <>
struct Codec<f32d> {
    static constexpr bool deviation = true;
    __host__ __device__ static float dec(f32d s, float rest) { return s.v + rest; }
    __host__ __device__ static f32d enc(float v, float rest) { return {v - rest}; }
    __host__ __device__ static float dec0(f32d s) { return s.v; }
    __host__ __device__ static float dec0(float s) { return s; }  // an element of a float pair
    __host__ __device__ static f32d enc0(float v) { return {v}; }
};

// Compute type R and the storage types of f_0 (S0), f_1..6 (SF) and g (SG)
// for a precision tag S.  bf16 mode keeps the f populations, which carry
// phi, in fp16 (four times finer than bf16 at the same 2 bytes) and g in
// bf16; f_0 ~ phi is O(1) and stays FP32 in both 16-bit modes.
template <class S> struct Traits;
template <> struct Traits<double> { using R = double; using S0 = double; using SF = double; using SG = double; };
template <> struct Traits<float> { using R = float; using S0 = float; using SF = float; using SG = float; };
template <> struct Traits<bf16s> { using R = float; using S0 = float; using SF = f16s; using SG = bf16s; };
template <> struct Traits<f16s> { using R = float; using S0 = float; using SF = f16s; using SG = f16s; };
template <> struct Traits<f32d> { using R = float; using S0 = float; using SF = float; using SG = f32d; };

// One lattice (A or B) in the context's layout: f_0 array, f_1..6 with
// stride n, g_0..18 with stride n.  Pointers are device addresses.
template <class S>
struct Lattice {
    typename Traits<S>::S0 *f0;
    typename Traits<S>::SF *f;  // f_i at f + (i-1)*n, i = 1..6
    typename Traits<S>::SG *g;  // g_i at g + i*n
};

}  // namespace zsc
