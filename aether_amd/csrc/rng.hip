// rng.hip -- Philox4x32-10 stream fills (gfx950).
//
// Pure integer VALU: one Philox block per output element (per two words for
// BITS32), several elements per lane so each lane issues one 16-byte store
// (global_store_dwordx4) per step. Streams and conversions are aether's
// (random/detail/Backend.h:123-236), element j of a fill is the draw at
// global sample j, counter 0 ("key = seed, counter = sample index"
// streams).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/aether_hip.h"
#include "common.hpp"
#include "rng.hpp"

namespace aeth {
namespace {

constexpr int kThreads = 256;

// 4 words per lane-step: blocks (j0>>1) and (j0>>1)+1 give words j0..j0+3.
__global__ __launch_bounds__(kThreads) void bits32_kernel(uint64_t seed, uint64_t first, uint64_t n,
                                                          uint32_t* __restrict__ out, bool vec)
{
    const uint64_t quads = (n + 3) / 4;
    for (uint64_t q = (uint64_t)blockIdx.x * kThreads + threadIdx.x; q < quads; q += (uint64_t)gridDim.x * kThreads) {
        const uint64_t k0 = 4 * q;
        if (vec && (first & 1) == 0 && k0 + 4 <= n) {
            const uint64_t j = first + k0;
            const rng::U4 b0 = rng::block(seed, (uint32_t)(j >> 1), 0);
            const rng::U4 b1 = rng::block(seed, (uint32_t)((j >> 1) + 1), 0);
            *reinterpret_cast<uint4*>(out + k0) = make_uint4(b0.x, b0.y, b1.x, b1.y);
        } else {
            for (int t = 0; t < 4; ++t) {
                const uint64_t k = k0 + t;
                if (k < n) {
                    const uint64_t j = first + k;
                    const rng::U4 b = rng::block(seed, (uint32_t)(j >> 1), 0);
                    out[k] = (j & 1) ? b.y : b.x;
                }
            }
        }
    }
}

template<int KIND>
__device__ __forceinline__ double draw_d(uint64_t seed, uint32_t j)
{
    const rng::U4 b = rng::block(seed, j, 0);
    if constexpr (KIND == AETH_RNG_UNIFORM_F64) return rng::u01_from_word_d(b.x);
    else return rng::normal_from_block_d(b);
}

template<int KIND>
__device__ __forceinline__ float draw_f(uint64_t seed, uint32_t j)
{
    const rng::U4 b = rng::block(seed, j, 0);
    if constexpr (KIND == AETH_RNG_UNIFORM_F32) return rng::u01_from_word_f(b.x);
    else return rng::normal_from_block_f(b);
}

template<int KIND>
__global__ __launch_bounds__(kThreads) void f64_kernel(uint64_t seed, uint64_t first, uint64_t n,
                                                       double* __restrict__ out, bool vec)
{
    const uint64_t pairs = (n + 1) / 2;
    for (uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x; p < pairs; p += (uint64_t)gridDim.x * kThreads) {
        const uint64_t k0 = 2 * p;
        if (vec && k0 + 2 <= n) {
            const double x0 = draw_d<KIND>(seed, (uint32_t)(first + k0));
            const double x1 = draw_d<KIND>(seed, (uint32_t)(first + k0 + 1));
            *reinterpret_cast<double2*>(out + k0) = make_double2(x0, x1);
        } else {
            for (int t = 0; t < 2; ++t)
                if (k0 + t < n) out[k0 + t] = draw_d<KIND>(seed, (uint32_t)(first + k0 + t));
        }
    }
}

template<int KIND>
__global__ __launch_bounds__(kThreads) void f32_kernel(uint64_t seed, uint64_t first, uint64_t n,
                                                       float* __restrict__ out, bool vec)
{
    const uint64_t quads = (n + 3) / 4;
    for (uint64_t q = (uint64_t)blockIdx.x * kThreads + threadIdx.x; q < quads; q += (uint64_t)gridDim.x * kThreads) {
        const uint64_t k0 = 4 * q;
        if (vec && k0 + 4 <= n) {
            float4 v;
            v.x = draw_f<KIND>(seed, (uint32_t)(first + k0));
            v.y = draw_f<KIND>(seed, (uint32_t)(first + k0 + 1));
            v.z = draw_f<KIND>(seed, (uint32_t)(first + k0 + 2));
            v.w = draw_f<KIND>(seed, (uint32_t)(first + k0 + 3));
            *reinterpret_cast<float4*>(out + k0) = v;
        } else {
            for (int t = 0; t < 4; ++t)
                if (k0 + t < n) out[k0 + t] = draw_f<KIND>(seed, (uint32_t)(first + k0 + t));
        }
    }
}

// Array fill: plane c, sample i = p0 + (p1 - p0) u01(seed, i, c) or p0 + p1 N(seed, i, c). Blocks walk
// the components in y and the samples of a plane in x, so no per-element division is needed.
template<class T, int DIST>
__global__ __launch_bounds__(kThreads) void array_fill_kernel(uint64_t seed, T p0, T p1, int ncomp, uint32_t n,
                                                              uint64_t pitch, T* __restrict__ out)
{
    for (uint32_t c = blockIdx.y; c < (uint32_t)ncomp; c += gridDim.y) {
        T* plane = out + (uint64_t)c * pitch;
        for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
            const rng::U4 b = rng::block(seed, (uint32_t)i, c);
            T v;
            if constexpr (sizeof(T) == 8) {
                if constexpr (DIST == AETH_DIST_UNIFORM) v = p0 + (p1 - p0) * rng::u01_from_word_d(b.x);
                else v = p0 + p1 * rng::normal_from_block_d(b);
            } else {
                if constexpr (DIST == AETH_DIST_UNIFORM) v = p0 + (p1 - p0) * rng::u01_from_word_f(b.x);
                else v = p0 + p1 * rng::normal_from_block_f(b);
            }
            plane[i] = v;
        }
    }
}

int grid_for(uint64_t work)
{
    const uint64_t blocks = (work + kThreads - 1) / kThreads;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(blocks, (uint64_t)device_cus() * 32u));
}

} // namespace
} // namespace aeth

using namespace aeth;

extern "C" int aeth_rng_fill(int kind, uint64_t seed, uint64_t first, uint64_t n, void* out, void* stream)
{
    return guarded([&] {
        if (n == 0) return;
        if (!out) fail("aeth_rng_fill", "out is NULL");
        const uint64_t limit = kind == AETH_RNG_BITS32 ? (1ull << 33) : (1ull << 32);
        if (first + n > limit || first + n < first) fail("aeth_rng_fill", "stream index beyond the 32-bit global span");
        hipStream_t st = as_stream(stream);
        const bool vec = aligned16(out);
        switch (kind) {
        case AETH_RNG_BITS32:
            hipLaunchKernelGGL(bits32_kernel, dim3(grid_for((n + 3) / 4)), dim3(kThreads), 0, st, seed, first, n,
                               static_cast<uint32_t*>(out), vec);
            break;
        case AETH_RNG_UNIFORM_F64:
            hipLaunchKernelGGL((f64_kernel<AETH_RNG_UNIFORM_F64>), dim3(grid_for((n + 1) / 2)), dim3(kThreads), 0, st,
                               seed, first, n, static_cast<double*>(out), vec);
            break;
        case AETH_RNG_NORMAL_F64:
            hipLaunchKernelGGL((f64_kernel<AETH_RNG_NORMAL_F64>), dim3(grid_for((n + 1) / 2)), dim3(kThreads), 0, st,
                               seed, first, n, static_cast<double*>(out), vec);
            break;
        case AETH_RNG_UNIFORM_F32:
            hipLaunchKernelGGL((f32_kernel<AETH_RNG_UNIFORM_F32>), dim3(grid_for((n + 3) / 4)), dim3(kThreads), 0, st,
                               seed, first, n, static_cast<float*>(out), vec);
            break;
        case AETH_RNG_NORMAL_F32:
            hipLaunchKernelGGL((f32_kernel<AETH_RNG_NORMAL_F32>), dim3(grid_for((n + 3) / 4)), dim3(kThreads), 0, st,
                               seed, first, n, static_cast<float*>(out), vec);
            break;
        default: fail("aeth_rng_fill", "unknown stream kind");
        }
        check(hipGetLastError(), "aeth_rng_fill launch");
    });
}

extern "C" int aeth_array_fill(int dist, int dtype, uint64_t seed, double p0, double p1, int ncomp, uint64_t n,
                               uint64_t pitch, void* out, void* stream)
{
    return guarded([&] {
        if (ncomp < 1 || pitch < n) fail("aeth_array_fill", "bad shape");
        if ((uint64_t)ncomp * pitch > 0xFFFFFFFFull) fail("aeth_array_fill", "span exceeds the 32-bit offset_t");
        if (n == 0) return;
        hipStream_t st = as_stream(stream);
        const uint32_t planes = (uint32_t)std::min<int>(ncomp, 65535);
        const uint64_t per_plane = std::max<uint64_t>(1, (uint64_t)grid_for((uint64_t)ncomp * n) / planes);
        const dim3 grid((uint32_t)std::min<uint64_t>(per_plane, (n + kThreads - 1) / kThreads), planes);
        const uint32_t nn = (uint32_t)n;
        if (dtype == AETH_F64) {
            auto* o = static_cast<double*>(out);
            if (dist == AETH_DIST_UNIFORM)
                hipLaunchKernelGGL((array_fill_kernel<double, AETH_DIST_UNIFORM>), grid, dim3(kThreads), 0, st,
                                   seed, p0, p1, ncomp, nn, pitch, o);
            else if (dist == AETH_DIST_NORMAL)
                hipLaunchKernelGGL((array_fill_kernel<double, AETH_DIST_NORMAL>), grid, dim3(kThreads), 0, st,
                                   seed, p0, p1, ncomp, nn, pitch, o);
            else
                fail("aeth_array_fill", "unknown distribution");
        } else if (dtype == AETH_F32) {
            auto* o = static_cast<float*>(out);
            if (dist == AETH_DIST_UNIFORM)
                hipLaunchKernelGGL((array_fill_kernel<float, AETH_DIST_UNIFORM>), grid, dim3(kThreads), 0, st,
                                   seed, (float)p0, (float)p1, ncomp, nn, pitch, o);
            else if (dist == AETH_DIST_NORMAL)
                hipLaunchKernelGGL((array_fill_kernel<float, AETH_DIST_NORMAL>), grid, dim3(kThreads), 0, st,
                                   seed, (float)p0, (float)p1, ncomp, nn, pitch, o);
            else
                fail("aeth_array_fill", "unknown distribution");
        } else {
            fail("aeth_array_fill", "unknown dtype");
        }
        check(hipGetLastError(), "aeth_array_fill launch");
    });
}
