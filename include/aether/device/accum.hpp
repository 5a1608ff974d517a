// aether/device/accum.hpp -- device-side accumulation primitives for HIP kernels on gfx950:
// compensated atomic planes, append counters and keyed sums with one atomic per run of lanes.
//
// The reference's primitives (accum/AccumPlane.h:36-470, accum/AppendCounter.h:30-120,
// accum/run.h:1-461, accum/atomic.h:120-137, accum/TwoSum.h:87-154) are written for 32-lane warps;
// these are the wave64 forms. A "run" is a set of lanes of one wavefront adding into one slot:
//   aligned<W> : the caller's fact is that the slot is equal over every aligned group of W lanes
//                (W a power of two, 2..64); one vote checks that the wave holds 64 consecutive own
//                samples starting on a multiple of 64.
//   tested     : no fact; one vote checks that every lane of the full wave has the same key.
//   sorted     : equal keys are adjacent in lane order; every maximal run is found from head flags.
// A wave the vote rejects, or one that is not full, commits lane by lane, so every call is correct for
// any key. A run is summed with a plain tree (xor butterfly for aligned/tested; a segmented
// down-shuffle tree, the lower lane on the left, for sorted) and committed with ONE atomic; the
// compensated calls make that one commit compensated: the run's sum goes into `target`, the exact
// two-sum residual of that add into `comp`, and the delivered value is target + comp.
//
// Atomics are the hardware FP add of device memory (coarse-grained hipMalloc memory).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "flag.hpp"

namespace aether {
namespace accum {

/** The "no slot" sentinel of try_append (accum/AppendCounter.h:26-29). */
inline constexpr uint32_t npos = 0xFFFFFFFFu;

namespace detail {

__device__ __forceinline__ bool finite(double x)
{
    return ((__double_as_longlong(x) >> 52) & 0x7ff) != 0x7ff;
}
__device__ __forceinline__ bool finite(float x) { return ((__float_as_uint(x) >> 23) & 0xff) != 0xff; }

__device__ __forceinline__ double positive_zero(double x)
{
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    return (b << 1) == 0 ? 0.0 : x;
}
__device__ __forceinline__ float positive_zero(float x)
{
    const uint32_t b = __float_as_uint(x);
    return (b << 1) == 0 ? 0.0f : x;
}

/** The exact residual (a + b) - s of s = fl(a + b) (Knuth's two-sum); +0 for a zero residual, and
 *  s - s (NaN) when s is not finite (accum/TwoSum.h:119-133). */
template<class E>
__device__ __forceinline__ E two_sum_err(E a, E b, E s)
{
    const E bb = s - a;
    const E e = (a - (s - bb)) + (b - bb);
    return finite(s) ? positive_zero(e) : s - s;
}

template<class E>
__device__ __forceinline__ E atomic_add(E* p, E v)
{
    return unsafeAtomicAdd(p, v);
}

__device__ __forceinline__ bool wave_full() { return __activemask() == ~0ull; }

} // namespace detail

/** Compensated atomic accumulation of `term` into the pair (value, comp) (accum/atomic.h:120-137):
 *  one atomic add of the term, one of its two-sum residual. Safe for concurrent producers. */
template<class E>
__device__ __forceinline__ void compensated_add(E* value, E* comp, E term)
{
    const E old = detail::atomic_add(value, term);
    const E s = old + term;
    detail::atomic_add(comp, detail::two_sum_err(old, term, s));
}

// ------------------------------------------------------------------ accumulation planes

enum class Policy { CompensatedAtomic, Serialized };

/** A per-sample accumulator of VD lanes of double (accum/AccumPlane.h:300-470): value planes and
 *  companion (residual) planes, `pitch` elements apart. */
template<int VD, Policy P = Policy::CompensatedAtomic>
struct PlaneView {
    double* value;
    double* comp; // unused by Serialized
    uint64_t pitch;

    struct Lane {
        const PlaneView& plane;
        uint64_t i;
        /** Fold a VD-component term into the sample, components in ascending order. */
        __device__ __forceinline__ Lane& operator+=(const double (&term)[VD])
        {
#pragma unroll
            for (int d = 0; d < VD; ++d) {
                double* v = plane.value + (uint64_t)d * plane.pitch + i;
                if constexpr (P == Policy::CompensatedAtomic)
                    compensated_add(v, plane.comp + (uint64_t)d * plane.pitch + i, term[d]);
                else
                    *v = *v + term[d];
            }
            return *this;
        }
        __device__ __forceinline__ double delivered(int d) const
        {
            const double v = plane.value[(uint64_t)d * plane.pitch + i];
            if constexpr (P == Policy::CompensatedAtomic) return v + plane.comp[(uint64_t)d * plane.pitch + i];
            else return v;
        }
        __device__ __forceinline__ void zero()
        {
#pragma unroll
            for (int d = 0; d < VD; ++d) {
                plane.value[(uint64_t)d * plane.pitch + i] = 0.0;
                if constexpr (P == Policy::CompensatedAtomic) plane.comp[(uint64_t)d * plane.pitch + i] = 0.0;
            }
        }
    };
    __device__ __forceinline__ Lane operator[](uint64_t i) const { return Lane{*this, i}; }
};

// ------------------------------------------------------------------ append counters

/** Hands out slots of a destination with `capacity` places (accum/AppendCounter.h:30-76). */
struct AppendCounterView {
    uint32_t* counter;
    uint64_t capacity;
    DeviceFlagView overflow; // set when a claim does not fit

    /** One atomic per claim: a slot < capacity, or npos (and the overflow flag) when full. */
    __device__ __forceinline__ uint32_t try_append() const
    {
        const uint32_t slot = atomicAdd(counter, 1u);
        if ((uint64_t)slot >= capacity) {
            overflow.set();
            return npos;
        }
        return slot;
    }

    /** The same contract with one atomic per wavefront: the active lanes that claim (`want`) get
     *  consecutive slots in lane order from one atomic add of their count. */
    __device__ __forceinline__ uint32_t try_append_wave(bool want) const
    {
        const uint64_t ballot = __ballot(want);
        const uint32_t lane = __lane_id();
        const int leader = __ffsll((long long)ballot) - 1;
        uint32_t base = 0;
        if (want && (int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(ballot));
        base = __shfl(base, leader < 0 ? 0 : leader);
        if (!want) return npos;
        const uint32_t slot = base + (uint32_t)__popcll(ballot & ((1ull << lane) - 1ull));
        if ((uint64_t)slot >= capacity || slot < base) {
            overflow.set();
            return npos;
        }
        return slot;
    }
};

// ------------------------------------------------------------------ keyed sums

namespace run {

/** Counters of (run commits, lane commits) for diagnostics; null in production. */
struct Commits {
    unsigned long long* c = nullptr;
    __device__ __forceinline__ void count(bool run) const
    {
        if (c) atomicAdd(&c[run ? 0 : 1], 1ull);
    }
};

namespace detail {

template<class E>
__device__ __forceinline__ void commit(E* slot, E s, bool run, Commits k)
{
    accum::detail::atomic_add(slot, s);
    k.count(run);
}

template<class E>
__device__ __forceinline__ void commit_comp(E* slot, E* comp, E s, bool run, Commits k)
{
    const E old = accum::detail::atomic_add(slot, s);
    accum::detail::atomic_add(comp, accum::detail::two_sum_err(old, s, old + s));
    k.count(run);
}

/** The full wave holds 64 consecutive own samples starting on a multiple of 64 (one vote). */
__device__ __forceinline__ bool wave_aligned(uint64_t own)
{
    return accum::detail::wave_full() && __all((uint32_t)(own & 63u) == __lane_id());
}

/** The 32-bit word of lane (lane ^ OFF) of a full wave: data-parallel-primitive moves inside a quad
 *  (OFF 1, 2), the swizzle's XOR mode inside 32 lanes (4, 8, 16), the half-wave swap (32). */
template<int OFF>
__device__ __forceinline__ uint32_t xor_lane_u32(uint32_t x)
{
    if constexpr (OFF == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);
    } else if constexpr (OFF == 2) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);
    } else if constexpr (OFF < 32) {
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, (OFF << 10) | 0x1F);
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        return (__lane_id() & 32) ? r[0] : r[1];
    }
}

template<int OFF, class E>
__device__ __forceinline__ E xor_lane(E v)
{
    if constexpr (sizeof(E) == 8) {
        const uint64_t b = __builtin_bit_cast(uint64_t, v);
        const uint64_t lo = xor_lane_u32<OFF>((uint32_t)b), hi = xor_lane_u32<OFF>((uint32_t)(b >> 32));
        return __builtin_bit_cast(E, (hi << 32) | lo);
    } else {
        return __builtin_bit_cast(E, xor_lane_u32<OFF>(__builtin_bit_cast(uint32_t, v)));
    }
}

/** Plain xor butterfly over aligned groups of W lanes (offsets W/2, ..., 1) of a full wave. */
template<int W, class E>
__device__ __forceinline__ E group_sum(E t)
{
    if constexpr (W > 1) {
        t += xor_lane<W / 2>(t);
        return group_sum<W / 2>(t);
    } else {
        return t;
    }
}

/** Segmented sum of a full wave's sorted runs (the run's head lane ends with its sum): lane l adds the
 *  lanes below the next head, offsets 1, 2, 4, ..., 32, the lower lane on the left. */
template<class E>
__device__ __forceinline__ E sorted_sum(E t, uint64_t heads, uint32_t lane)
{
    if (heads == ~0ull) return t;
    const uint64_t later = lane == 63u ? 0ull : heads >> (lane + 1u);
    const uint32_t end = later ? lane + 1u + (uint32_t)(__ffsll((long long)later) - 1) : 64u;
#pragma unroll
    for (uint32_t o = 1; o < 64u; o <<= 1) {
        const E v = __shfl_down(t, o);
        if (lane + o < end) t = t + v;
    }
    return t;
}

template<class K>
__device__ __forceinline__ uint64_t sorted_heads(K key, uint32_t lane)
{
    const K below = __shfl_up(key, 1);
    return __ballot(lane == 0u || key != below);
}

template<bool COMP, class E>
__device__ __forceinline__ void put(E* slot, E* comp, E s, bool run, Commits k)
{
    if constexpr (COMP) commit_comp(slot, comp, s, run, k);
    else commit(slot, s, run, k);
}

} // namespace detail

/** add<W> (accum/run.h:272-292): one atomic per aligned group of W lanes when the vote on `own`
 *  (the lane's own sample index) passes, lane by lane otherwise. */
template<int W, bool COMP = false, class E>
__device__ __forceinline__ void add(E* target, E* comp, uint64_t at, E term, uint64_t own, Commits k = {})
{
    static_assert(W >= 2 && W <= 64 && (W & (W - 1)) == 0, "a run is 2..64 lanes");
    if (detail::wave_aligned(own)) {
        const E s = detail::group_sum<W>(term);
        if ((__lane_id() & (W - 1)) == 0u) detail::put<COMP>(target + at, comp + at, s, true, k);
        return;
    }
    detail::put<COMP>(target + at, comp + at, term, false, k);
}

/** add_tested (accum/run.h:294-316): one atomic when every lane of the full wave has the same key. */
template<bool COMP = false, class E, class K>
__device__ __forceinline__ void add_tested(E* target, E* comp, K key, E term, Commits k = {})
{
    if (accum::detail::wave_full() && __all(key == __shfl(key, 0))) {
        const E s = detail::group_sum<64>(term);
        if (__lane_id() == 0u) detail::put<COMP>(target + key, comp + key, s, true, k);
        return;
    }
    detail::put<COMP>(target + key, comp + key, term, false, k);
}

/** add_sorted (accum/run.h:318-344): one atomic per maximal run of equal keys in a full wave. */
template<bool COMP = false, class E, class K>
__device__ __forceinline__ void add_sorted(E* target, E* comp, K key, E term, Commits k = {})
{
    if (!accum::detail::wave_full()) {
        detail::put<COMP>(target + key, comp + key, term, false, k);
        return;
    }
    const uint32_t lane = __lane_id();
    const uint64_t heads = detail::sorted_heads(key, lane);
    const E s = detail::sorted_sum(term, heads, lane);
    if ((heads >> lane) & 1ull) detail::put<COMP>(target + key, comp + key, s, true, k);
}

} // namespace run
} // namespace accum
} // namespace aether
