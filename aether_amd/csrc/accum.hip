// accum.hip -- bulk entries over the device accumulation primitives (include/aether/device/accum.hpp):
// keyed sums with one atomic per run of lanes (accum/run.h), compensated accumulation planes
// (accum/AccumPlane.h) and append counters (accum/AppendCounter.h). One lane per input element,
// 256-lane blocks (four wavefronts); every kernel is a single pass over its inputs.
#include <hip/hip_runtime.h>

#include "../../include/aether/device/accum.hpp"
#include "../../include/aether_hip.h"
#include "common.hpp"

namespace aeth {
namespace {

constexpr int kLanes = 256;

inline dim3 lanes_grid(uint64_t n)
{
    return dim3((unsigned)((n + kLanes - 1) / kLanes));
}

template<class E, bool COMP, int FORM, int W>
__global__ __launch_bounds__(kLanes) void keyed_sum_kernel(const int64_t* __restrict__ keys, const E* __restrict__ terms,
                                                           uint64_t n, int64_t own0, uint64_t slots, E* target, E* comp,
                                                           unsigned long long* commits)
{
    namespace run = aether::accum::run;
    const uint64_t l = (uint64_t)blockIdx.x * kLanes + threadIdx.x;
    if (l >= n) return;
    const int64_t key = keys[l];
    if (key < 0 || (uint64_t)key >= slots) return; // a lane with no slot takes no part
    const E t = terms[l];
    const run::Commits k{commits};
    if constexpr (FORM == AETH_RUN_ALIGNED)
        run::add<W, COMP>(target, comp, (uint64_t)key, t, (uint64_t)((int64_t)l + own0), k);
    else if constexpr (FORM == AETH_RUN_TESTED)
        run::add_tested<COMP>(target, comp, key, t, k);
    else
        run::add_sorted<COMP>(target, comp, key, t, k);
}

template<class E, bool COMP, int FORM>
void keyed_launch(int width, const int64_t* keys, const E* terms, uint64_t n, int64_t own, uint64_t slots, E* target,
                  E* comp, unsigned long long* commits, hipStream_t st)
{
#define AETH_KS(W)                                                                                                  \
    hipLaunchKernelGGL((keyed_sum_kernel<E, COMP, FORM, W>), lanes_grid(n), dim3(kLanes), 0, st, keys, terms, n, own, \
                       slots, target, comp, commits)
    if constexpr (FORM != AETH_RUN_ALIGNED) {
        (void)width;
        AETH_KS(64);
    } else {
        switch (width) {
        case 2: AETH_KS(2); break;
        case 4: AETH_KS(4); break;
        case 8: AETH_KS(8); break;
        case 16: AETH_KS(16); break;
        case 32: AETH_KS(32); break;
        case 64: AETH_KS(64); break;
        default: fail("aeth_keyed_sum", "aligned runs are 2, 4, 8, 16, 32 or 64 lanes");
        }
    }
#undef AETH_KS
}

template<class E>
void keyed_dispatch(int form, int width, bool comp_on, const int64_t* keys, const E* terms, uint64_t n, int64_t own,
                    uint64_t slots, E* target, E* comp, unsigned long long* commits, hipStream_t st)
{
    if (comp_on) {
        if (form == AETH_RUN_ALIGNED) keyed_launch<E, true, AETH_RUN_ALIGNED>(width, keys, terms, n, own, slots, target, comp, commits, st);
        else if (form == AETH_RUN_TESTED) keyed_launch<E, true, AETH_RUN_TESTED>(width, keys, terms, n, own, slots, target, comp, commits, st);
        else keyed_launch<E, true, AETH_RUN_SORTED>(width, keys, terms, n, own, slots, target, comp, commits, st);
    } else {
        if (form == AETH_RUN_ALIGNED) keyed_launch<E, false, AETH_RUN_ALIGNED>(width, keys, terms, n, own, slots, target, comp, commits, st);
        else if (form == AETH_RUN_TESTED) keyed_launch<E, false, AETH_RUN_TESTED>(width, keys, terms, n, own, slots, target, comp, commits, st);
        else keyed_launch<E, false, AETH_RUN_SORTED>(width, keys, terms, n, own, slots, target, comp, commits, st);
    }
    check(hipGetLastError(), "keyed_sum_kernel launch");
}

template<int VD, aether::accum::Policy P>
__global__ __launch_bounds__(kLanes) void plane_add_kernel(const double* __restrict__ terms, uint64_t term_pitch,
                                                           uint64_t n, aether::accum::PlaneView<VD, P> plane)
{
    const uint64_t i = (uint64_t)blockIdx.x * kLanes + threadIdx.x;
    if (i >= n) return;
    double t[VD];
#pragma unroll
    for (int d = 0; d < VD; ++d) t[d] = terms[(uint64_t)d * term_pitch + i];
    plane[i] += t;
}

template<aether::accum::Policy P>
void plane_launch(int vd, const double* terms, uint64_t term_pitch, uint64_t n, double* value, double* comp,
                  uint64_t pitch, hipStream_t st)
{
#define AETH_PA(V)                                                                                                  \
    hipLaunchKernelGGL((plane_add_kernel<V, P>), lanes_grid(n), dim3(kLanes), 0, st, terms, term_pitch, n,          \
                       aether::accum::PlaneView<V, P>{value, comp, pitch})
    switch (vd) {
    case 1: AETH_PA(1); break;
    case 2: AETH_PA(2); break;
    case 3: AETH_PA(3); break;
    case 4: AETH_PA(4); break;
    case 6: AETH_PA(6); break;
    case 8: AETH_PA(8); break;
    default: fail("aeth_plane_add", "planes have 1, 2, 3, 4, 6 or 8 lanes");
    }
#undef AETH_PA
    check(hipGetLastError(), "plane_add_kernel launch");
}

template<class T, bool WAVE>
__global__ __launch_bounds__(kLanes) void append_where_kernel(const T* __restrict__ x, uint64_t n, double threshold,
                                                              uint32_t* __restrict__ out,
                                                              aether::accum::AppendCounterView v)
{
    // every lane of a wave takes part in the wave's claim, so none returns early
    const uint64_t i = (uint64_t)blockIdx.x * kLanes + threadIdx.x;
    const bool want = i < n && (double)x[i] > threshold;
    uint32_t slot;
    if constexpr (WAVE) slot = v.try_append_wave(want);
    else slot = want ? v.try_append() : aether::accum::npos;
    if (slot != aether::accum::npos) out[slot] = (uint32_t)i;
}

// One atomic per tile of 16384 elements: every wave ballots its 64 rounds into LDS, wave 0 scans the
// 256 (round, wave) counts and claims the tile's total with one atomic, and each claimant writes at
// base + (claims before it in round-major, wave, lane order): a tile's elements keep their order.
constexpr int kTileRounds = 64;
constexpr int kWaves = kLanes / 64;

template<class T>
__global__ __launch_bounds__(kLanes) void append_tile_kernel(const T* __restrict__ x, uint64_t n, double threshold,
                                                             uint32_t* __restrict__ out,
                                                             aether::accum::AppendCounterView v)
{
    __shared__ uint64_t ball[kTileRounds * kWaves];
    __shared__ uint32_t before[kTileRounds * kWaves];
    __shared__ uint32_t base_s;
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
    const uint64_t tile0 = (uint64_t)blockIdx.x * (kLanes * kTileRounds);
#pragma unroll 8
    for (int r = 0; r < kTileRounds; ++r) {
        const uint64_t i = tile0 + (uint64_t)r * kLanes + threadIdx.x;
        const bool want = i < n && (double)x[i] > threshold;
        const uint64_t b = __ballot(want);
        if (lane == 0) ball[r * kWaves + wave] = b;
    }
    __syncthreads();
    if (wave == 0) {
        // lane l owns entries 4l .. 4l+3 (entry = round * kWaves + wave)
        uint32_t c[kWaves], own = 0;
#pragma unroll
        for (int k = 0; k < kWaves; ++k) {
            c[k] = (uint32_t)__popcll(ball[lane * kWaves + k]);
            own += c[k];
        }
        uint32_t incl = own; // inclusive scan of the lanes' totals
#pragma unroll
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        uint32_t run = incl - own;
#pragma unroll
        for (int k = 0; k < kWaves; ++k) {
            before[lane * kWaves + k] = run;
            run += c[k];
        }
        if (lane == 63) base_s = incl ? atomicAdd(v.counter, incl) : 0u;
    }
    __syncthreads();
    const uint32_t base = base_s;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    bool over = false;
    for (int r = 0; r < kTileRounds; ++r) {
        const uint64_t b = ball[r * kWaves + wave];
        if ((b >> lane) & 1ull) {
            const uint32_t slot = base + before[r * kWaves + wave] + (uint32_t)__popcll(b & below);
            if ((uint64_t)slot < v.capacity && slot >= base) out[slot] = (uint32_t)(tile0 + (uint64_t)r * kLanes + threadIdx.x);
            else over = true;
        }
    }
    if (over) v.overflow.set();
}

} // namespace
} // namespace aeth

using namespace aeth;

extern "C" int aeth_keyed_sum(int dtype, int form, int width, int compensated, const int64_t* keys,
                              const void* terms, uint64_t n, int64_t own, uint64_t slots, void* target, void* comp,
                              unsigned long long* commits, void* stream)
{
    return guarded([&] {
        if (form < AETH_RUN_ALIGNED || form > AETH_RUN_SORTED) fail("aeth_keyed_sum", "unknown run form");
        if (compensated && !comp) fail("aeth_keyed_sum", "a compensated sum needs the comp array");
        if (n == 0) return;
        if (!keys || !terms || !target) fail("aeth_keyed_sum", "keys, terms and target are required");
        hipStream_t st = as_stream(stream);
        if (dtype == AETH_F64)
            keyed_dispatch<double>(form, width, compensated != 0, keys, static_cast<const double*>(terms), n, own,
                                   slots, static_cast<double*>(target), static_cast<double*>(comp), commits, st);
        else if (dtype == AETH_F32)
            keyed_dispatch<float>(form, width, compensated != 0, keys, static_cast<const float*>(terms), n, own,
                                  slots, static_cast<float*>(target), static_cast<float*>(comp), commits, st);
        else
            fail("aeth_keyed_sum", "unknown dtype");
    });
}

extern "C" int aeth_plane_add(int policy, int vd, const double* terms, uint64_t term_pitch, uint64_t n,
                              double* value, double* comp, uint64_t pitch, void* stream)
{
    return guarded([&] {
        if (policy != AETH_PLANE_COMPENSATED && policy != AETH_PLANE_SERIALIZED) fail("aeth_plane_add", "unknown policy");
        if (policy == AETH_PLANE_COMPENSATED && !comp) fail("aeth_plane_add", "a compensated plane needs its comp planes");
        if (n == 0) return;
        if (term_pitch < n || pitch < n) fail("aeth_plane_add", "pitch < n");
        hipStream_t st = as_stream(stream);
        if (policy == AETH_PLANE_COMPENSATED)
            plane_launch<aether::accum::Policy::CompensatedAtomic>(vd, terms, term_pitch, n, value, comp, pitch, st);
        else
            plane_launch<aether::accum::Policy::Serialized>(vd, terms, term_pitch, n, value, comp, pitch, st);
    });
}

extern "C" int aeth_append_where(int dtype, const void* x, uint64_t n, double threshold, uint32_t* out,
                                 uint64_t capacity, uint32_t* counter, uint32_t* overflow, int claim, void* stream)
{
    return guarded([&] {
        if (!counter || !overflow) fail("aeth_append_where", "counter and overflow flag are required");
        if (n == 0) return;
        if (n > 0xFFFFFFFFull) fail("aeth_append_where", "indices are 32-bit");
        const aether::accum::AppendCounterView v{counter, capacity, aether::DeviceFlagView(overflow)};
        if (claim < AETH_CLAIM_LANE || claim > AETH_CLAIM_TILE) fail("aeth_append_where", "unknown claim granularity");
        hipStream_t st = as_stream(stream);
        const dim3 g = lanes_grid(n);
        const dim3 gt((unsigned)((n + kLanes * kTileRounds - 1) / (kLanes * kTileRounds)));
        auto go = [&](auto tag) {
            using T = decltype(tag);
            auto X = static_cast<const T*>(x);
            if (claim == AETH_CLAIM_TILE)
                hipLaunchKernelGGL((append_tile_kernel<T>), gt, dim3(kLanes), 0, st, X, n, threshold, out, v);
            else if (claim == AETH_CLAIM_WAVE)
                hipLaunchKernelGGL((append_where_kernel<T, true>), g, dim3(kLanes), 0, st, X, n, threshold, out, v);
            else
                hipLaunchKernelGGL((append_where_kernel<T, false>), g, dim3(kLanes), 0, st, X, n, threshold, out, v);
        };
        if (dtype == AETH_F64) go(double{});
        else if (dtype == AETH_F32) go(float{});
        else fail("aeth_append_where", "unknown dtype");
        check(hipGetLastError(), "append_where_kernel launch");
    });
}
