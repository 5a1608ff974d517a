// accum.hpp -- compensated accumulation on gfx950 wave64.
//
// neumaierAdd_ / pairCombine_ follow accum/atomic.h:142-164 with the residual
// of accum/TwoSum.h:87-109 (the +0 rule: a zero residual is +0). The
// float-float add follows banded/Ff2.h:132-145 and the error-free transforms
// of banded/detail/Fp32.h:198-230.
//
// Reduction shape (fixed by the sample count alone, so a result never
// depends on the launch): a tile is 2048 consecutive samples handled by 256
// lanes; each lane walks its 16-byte vectors in order; the lanes of a wave
// are combined by an xor butterfly (1, 2, 4, 8, 16, 32 -- pairCombine_ is
// symmetric bit for bit, so every lane ends with the same pair); the four
// waves combine as ((w0, w1), (w2, w3)) through LDS; tile partials are then
// reduced in groups of 1024 by a halving tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aeth {
namespace acc {

enum Mode : int { kNeumaier = 0, kPlain = 1, kFloatFloat = 2 };

template<class T>
struct Pair {
    T h, c;
};

template<class T>
__device__ __forceinline__ bool finite(T x)
{
    return __builtin_isfinite(x);
}

template<class T>
__device__ __forceinline__ T two_sum_err(T a, T b, T s)
{
    const T bb = s - a;
    const T e = (a - (s - bb)) + (b - bb);
    return e == T(0) ? T(0) : e;
}

template<class T>
__device__ __forceinline__ void neumaier_add(T& h, T& c, T t)
{
    const T s = h + t;
    const T e = two_sum_err(h, t, s);
    h = s;
    c = finite(s) ? c + e : T(0);
}

template<class T>
__device__ __forceinline__ Pair<T> pair_combine(Pair<T> a, Pair<T> b)
{
    const T s = a.h + b.h;
    const T e = two_sum_err(a.h, b.h, s);
    return Pair<T>{s, finite(s) ? (a.c + b.c) + e : T(0)};
}

__device__ __forceinline__ void two_sum_f(float a, float b, float& s, float& e)
{
    s = a + b;
    const float bp = s - a;
    const float ap = s - bp;
    const float eb = b - bp;
    const float ea = a - ap;
    e = ea + eb;
}

__device__ __forceinline__ Pair<float> ff_add(Pair<float> a, Pair<float> b)
{
    float sh, sl, th, tl, vh, vl;
    two_sum_f(a.h, b.h, sh, sl);
    two_sum_f(a.c, b.c, th, tl);
    const float c = sl + th;
    two_sum_f(sh, c, vh, vl);
    const float w = vl + tl;
    const float hi = vh + w;
    const float bp = hi - vh;
    return Pair<float>{hi, w - bp};
}

// ff_add(a, {b, +0}) in four operations fewer, with the same bits: the second pair's low part is +0, and
// two_sum_f(a.c, +0) is exactly (a.c + 0, a.c - a.c) -- the sum (a.c, or +0 for -0) and an error that is
// +0 for a finite a.c and NaN (from a.c's own payload, or inf - inf) otherwise. Checked against ff_add
// bit for bit over signed zeros, subnormals, infinities, NaNs and random floats of every magnitude.
__device__ __forceinline__ Pair<float> ff_add_float(Pair<float> a, float b)
{
    float sh, sl, vh, vl;
    two_sum_f(a.h, b, sh, sl);
    const float th = a.c + 0.0f, tl = a.c - a.c;
    const float c = sl + th;
    two_sum_f(sh, c, vh, vl);
    const float w = vl + tl;
    const float hi = vh + w;
    const float bp = hi - vh;
    return Pair<float>{hi, w - bp};
}

// One accumulation step of `t` in mode M (square: accumulate t*t).
template<int M, bool SQUARE, class T>
__device__ __forceinline__ void accumulate(Pair<T>& p, T t)
{
    if constexpr (M == kNeumaier) {
        neumaier_add(p.h, p.c, SQUARE ? t * t : t);
    } else if constexpr (M == kPlain) {
        p.h = p.h + (SQUARE ? t * t : t);
    } else {
        static_assert(sizeof(T) == 4, "float-float accumulation is for float");
        if constexpr (SQUARE) {
            const float h = t * t;
            p = ff_add(p, Pair<float>{h, __builtin_fmaf(t, t, -h)});
        } else {
            p = ff_add_float(p, t);
        }
    }
}

template<int M, class T>
__device__ __forceinline__ Pair<T> combine(Pair<T> a, Pair<T> b)
{
    if constexpr (M == kNeumaier) {
        return pair_combine(a, b);
    } else if constexpr (M == kPlain) {
        return Pair<T>{a.h + b.h, T(0)};
    } else {
        return ff_add(a, b);
    }
}

// The 32-bit word of lane (lane ^ OFF), OFF a power of two below 64, without a trip through the LDS
// crossbar's addressed permute: data-parallel-primitive moves inside a quad (OFF 1, 2), the swizzle's
// XOR mode inside 32 lanes (4, 8, 16) and the half-wave swap (32). Every lane of the wave must be active.
template<int OFF>
__device__ __forceinline__ uint32_t xor_lane_u32(uint32_t x)
{
    static_assert(OFF > 0 && OFF < 64 && (OFF & (OFF - 1)) == 0, "a power of two below the wave size");
    if constexpr (OFF == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false); // quad_perm [1,0,3,2]
    } else if constexpr (OFF == 2) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false); // quad_perm [2,3,0,1]
    } else if constexpr (OFF < 32) {
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, (OFF << 10) | 0x1F); // and 0x1F, or 0, xor OFF
    } else {
        // after the swap the first result holds the low half in both halves, the second the high half
        const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        return (__lane_id() & 32) ? r[0] : r[1];
    }
}

template<int OFF, class T>
__device__ __forceinline__ T xor_lane(T v)
{
    if constexpr (sizeof(T) == 8) {
        const uint64_t b = __builtin_bit_cast(uint64_t, v);
        const uint64_t lo = xor_lane_u32<OFF>((uint32_t)b), hi = xor_lane_u32<OFF>((uint32_t)(b >> 32));
        return __builtin_bit_cast(T, (hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, xor_lane_u32<OFF>(__builtin_bit_cast(uint32_t, v)));
    }
}

// Butterfly over the 64 lanes of a wave (partner lane ^ 1, ^ 2, ..., ^ 32): every lane ends with the
// combined pair.
template<int M, class T, int OFF = 1>
__device__ __forceinline__ Pair<T> wave_combine(Pair<T> p)
{
    if constexpr (OFF < 64) {
        const Pair<T> q{xor_lane<OFF>(p.h), xor_lane<OFF>(p.c)};
        return wave_combine<M, T, 2 * OFF>(combine<M>(p, q));
    } else {
        return p;
    }
}

// Combine the 256 lane pairs of a block (4 waves) into one pair on thread 0.
// `lds` must hold 4 pairs. All 256 threads must call it.
template<int M, class T>
__device__ __forceinline__ Pair<T> block_combine(Pair<T> p, Pair<T>* lds)
{
    p = wave_combine<M>(p);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = p;
    __syncthreads();
    Pair<T> r{T(0), T(0)};
    if (threadIdx.x == 0) r = combine<M>(combine<M>(lds[0], lds[1]), combine<M>(lds[2], lds[3]));
    __syncthreads();
    return r;
}

} // namespace acc
} // namespace aeth
