// eval.hip -- fused SoA evaluation and compensated reductions (gfx950).
//
// The reference evaluates `dv[i] = expr` one sample per thread with the
// component loop unrolled (expr/Assign.h:124-160), and its 128-bit bundle
// path only applies to layout_right views (backend/cuda/bundle/LoadStore.h:
// 150-208). Here each whole-array assignment is one kernel whose unit of
// work is a (component, tile) pair: a tile is 2048 consecutive samples of
// one component plane, 256 lanes each issuing 16-byte loads (double2 /
// float4) along the sample axis of the plane, so every load instruction of a
// wave covers 1 KiB of contiguous memory. The optional per-component sum is
// folded into the same pass (0 extra HBM bytes), in the fixed tile shape of
// accum.hpp, and finished by a small group tree.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/aether_hip.h"
#include "accum.hpp"
#include "common.hpp"
#include "dmath.hpp"
#include "units.hpp"

namespace aeth {
namespace {

using namespace units;

// One block per (component, tile) unit, grid-stride over units.
template<class T, int M, bool SUM, bool VECTOR, bool NT>
__global__ __launch_bounds__(kThreads) void cfg0_kernel(const T* __restrict__ a, const T* __restrict__ b,
                                                        const T* __restrict__ c, T* __restrict__ d, uint32_t n,
                                                        uint64_t pitch, uint32_t tiles, uint32_t units,
                                                        acc::Pair<T>* __restrict__ part)
{
    __shared__ acc::Pair<T> lds[4];
    constexpr uint32_t K = xcd_chunk(kTile * sizeof(T));
    for (uint32_t v = blockIdx.x; v < units; v += gridDim.x) {
        const uint32_t u = xcd_unit(v, units, K);
        const uint32_t comp = u / tiles;
        const uint32_t tile = u - comp * tiles;
        const acc::Pair<T> p = cfg0_tile<T, M, SUM, VECTOR, NT>(a, b, c, d, n, pitch, comp, tile);
        if constexpr (SUM) {
            const acc::Pair<T> r = acc::block_combine<M>(p, lds);
            if (threadIdx.x == 0) part[u] = r;
        }
    }
}

// Sum (or sum of squares) of an SoA array in the tile shape.
template<class T, int M, bool SQUARE, bool VECTOR, bool NT>
__global__ __launch_bounds__(kThreads) void tile_reduce_kernel(const T* __restrict__ x, uint32_t n, uint64_t pitch,
                                                               uint32_t tiles, uint32_t units,
                                                               acc::Pair<T>* __restrict__ part)
{
    __shared__ acc::Pair<T> lds[4];
    constexpr uint32_t K = xcd_chunk(kTile * sizeof(T));
    for (uint32_t v = blockIdx.x; v < units; v += gridDim.x) {
        const uint32_t u = xcd_unit(v, units, K);
        const uint32_t comp = u / tiles;
        const uint32_t tile = u - comp * tiles;
        const acc::Pair<T> p = reduce_tile<T, M, SQUARE, VECTOR, NT>(x, n, pitch, comp, tile);
        const acc::Pair<T> r = acc::block_combine<M>(p, lds);
        if (threadIdx.x == 0) part[u] = r;
    }
}

// One level of the group tree: block (g, comp) reduces partials
// [g*1024, min(T, (g+1)*1024)) of component `comp` by the halving tree.
template<class T, int M>
__global__ __launch_bounds__(512) void group_tree_kernel(const acc::Pair<T>* __restrict__ in,
                                                         acc::Pair<T>* __restrict__ out, uint32_t count,
                                                         uint32_t in_stride, uint32_t out_stride)
{
    __shared__ acc::Pair<T> s[kGroup];
    const uint32_t g = blockIdx.x, comp = blockIdx.y, t = threadIdx.x;
    const uint32_t first = g * kGroup;
    const uint32_t m = min((uint32_t)kGroup, count - first);
    const acc::Pair<T>* src = in + (uint64_t)comp * in_stride + first;
    if (t < m) s[t] = src[t];
    if (t + 512 < m) s[t + 512] = src[t + 512];
    group_halving<T, M>(s, m);
    if (t == 0) out[(uint64_t)comp * out_stride + g] = s[0];
}

template<class T>
__global__ void store_pairs_kernel(const acc::Pair<T>* __restrict__ in, uint32_t in_stride, T* __restrict__ out,
                                   int ncomp)
{
    const int c = threadIdx.x;
    if (c < ncomp) {
        out[2 * c] = in[(uint64_t)c * in_stride].h;
        out[2 * c + 1] = in[(uint64_t)c * in_stride].c;
    }
}

// One block per (component, tile) unit: measured on MI355X this beats a
// persistent grid-stride launch by ~12% for the 3-read/1-write stream (the
// hardware dispatcher keeps every CU fed; a block's tail never idles a CU).
// The kernels still walk units grid-stride, so any grid is correct.
int grid_for(uint32_t units)
{
    return (int)std::max<uint32_t>(1u, std::min<uint32_t>(units, 0x7FFFFFFFu));
}

// Streaming cache policy once the operands dwarf the 256 MB last-level cache.
bool streaming(uint64_t bytes) { return bytes >= (512ull << 20); }

size_t tree_scratch_pairs(uint32_t tiles, int ncomp)
{
    size_t total = 0;
    uint32_t t = tiles;
    total += (size_t)t * ncomp;
    while (t > 1) {
        t = (t + kGroup - 1) / kGroup;
        total += (size_t)t * ncomp;
    }
    return total;
}

// Reduce `ncomp` rows of `tiles` partials (in scratch) to (hi, comp) pairs in `out`.
template<class T, int M>
void finish_tree(acc::Pair<T>* scratch, uint32_t tiles, int ncomp, T* out, hipStream_t st)
{
    acc::Pair<T>* cur = scratch;
    uint32_t count = tiles, stride = tiles;
    acc::Pair<T>* next = scratch + (size_t)tiles * ncomp;
    while (count > 1) {
        const uint32_t groups = (count + kGroup - 1) / kGroup;
        hipLaunchKernelGGL((group_tree_kernel<T, M>), dim3(groups, ncomp), dim3(512), 0, st, cur, next, count, stride,
                           groups);
        cur = next;
        next = next + (size_t)groups * ncomp;
        stride = groups;
        count = groups;
    }
    hipLaunchKernelGGL((store_pairs_kernel<T>), dim3(1), dim3(64), 0, st, cur, stride, out, ncomp);
    check(hipGetLastError(), "group tree launch");
}

template<class T>
void check_sizes(int ncomp, uint64_t n, uint64_t pitch, const char* op)
{
    if (ncomp < 1 || ncomp > 64) fail(op, "ncomp must be in [1, 64]");
    if (pitch < n) fail(op, "pitch < n");
    if ((uint64_t)ncomp * pitch > 0xFFFFFFFFull) fail(op, "ncomp*pitch exceeds the 32-bit offset_t span");
}

// Scratch of the allocation-free entry point (aeth_eval_cfg0_noalloc): the tile partials and the group
// tree's levels of its sums, in a file-scope device array that each GPU holds from the library's first
// use on it until the process ends. Every call on a GPU shares that GPU's copy; nothing in it has to
// start at zero (each partial is written before it is read), and calls that overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = 4u << 20;
} // namespace
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current device's copy of `scratch`, looked up once per device with that
// device current. In host code the array's name is a host-side stand-in and never a device address.
void* scratch_on_current_device()
{
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) fail("aeth_eval_cfg0_noalloc", "device ordinal beyond 64");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) fail("aeth_eval_cfg0_noalloc", "hipGetSymbolAddress returned no address");
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

// `fixed` (may be NULL): caller-owned scratch of kNoAllocScratchBytes used instead of a stream-ordered
// allocation; sizes whose partials do not fit are refused before anything is launched.
template<class T, int M, bool SUM>
void launch_cfg0(const T* a, const T* b, const T* c, T* d, int ncomp, uint64_t n, uint64_t pitch, T* sums,
                 hipStream_t st, void* fixed = nullptr)
{
    const uint32_t tiles = n ? (uint32_t)((n + kTile - 1) / kTile) : 1u;
    const uint32_t units = tiles * (uint32_t)ncomp;
    const bool vec = aligned16(a) && aligned16(b) && aligned16(c) && aligned16(d) && (pitch * sizeof(T)) % 16 == 0;
    const size_t need = SUM ? tree_scratch_pairs(tiles, ncomp) * sizeof(acc::Pair<T>) : 0;
    if (fixed && need > kNoAllocScratchBytes)
        fail("aeth_eval_cfg0_noalloc",
             "the sums' partials need " + std::to_string(need) + " bytes, beyond the fixed scratch of " +
                 std::to_string(kNoAllocScratchBytes) + " bytes",
             need);
    StreamLease scratch(fixed ? 0 : need, st);
    auto* part = static_cast<acc::Pair<T>*>(fixed ? fixed : scratch.get());
    const int grid = grid_for(units);
    const bool nt = streaming(4ull * ncomp * n * sizeof(T));
    if (vec && nt)
        hipLaunchKernelGGL((cfg0_kernel<T, M, SUM, true, true>), dim3(grid), dim3(kThreads), 0, st, a, b, c, d,
                           (uint32_t)n, pitch, tiles, units, part);
    else if (vec)
        hipLaunchKernelGGL((cfg0_kernel<T, M, SUM, true, false>), dim3(grid), dim3(kThreads), 0, st, a, b, c, d,
                           (uint32_t)n, pitch, tiles, units, part);
    else
        hipLaunchKernelGGL((cfg0_kernel<T, M, SUM, false, false>), dim3(grid), dim3(kThreads), 0, st, a, b, c, d,
                           (uint32_t)n, pitch, tiles, units, part);
    check(hipGetLastError(), "cfg0_kernel launch");
    if constexpr (SUM) finish_tree<T, M>(part, tiles, ncomp, sums, st);
}

template<class T>
void dispatch_cfg0(const void* a, const void* b, const void* c, void* d, int ncomp, uint64_t n, uint64_t pitch,
                   int mode, void* sums, hipStream_t st, void* fixed = nullptr)
{
    auto A = static_cast<const T*>(a), B = static_cast<const T*>(b), Cc = static_cast<const T*>(c);
    auto D = static_cast<T*>(d);
    auto S = static_cast<T*>(sums);
    if (mode != AETH_SUM_NONE && !sums) fail("aeth_eval_cfg0", "sums is NULL");
    switch (mode) {
    case AETH_SUM_NONE: launch_cfg0<T, acc::kNeumaier, false>(A, B, Cc, D, ncomp, n, pitch, S, st, fixed); break;
    case AETH_SUM_NEUMAIER: launch_cfg0<T, acc::kNeumaier, true>(A, B, Cc, D, ncomp, n, pitch, S, st, fixed); break;
    case AETH_SUM_PLAIN: launch_cfg0<T, acc::kPlain, true>(A, B, Cc, D, ncomp, n, pitch, S, st, fixed); break;
    case AETH_SUM_FLOATFLOAT:
        if constexpr (sizeof(T) == 4) {
            launch_cfg0<T, acc::kFloatFloat, true>(A, B, Cc, D, ncomp, n, pitch, S, st, fixed);
            break;
        } else {
            fail("aeth_eval_cfg0", "float-float accumulation needs dtype F32", 0, 2);
        }
    default: fail("aeth_eval_cfg0", "unknown sum mode");
    }
}

template<class T, int M, bool SQ>
void launch_reduce(const T* x, int ncomp, uint64_t n, uint64_t pitch, T* out, hipStream_t st)
{
    const uint32_t tiles = n ? (uint32_t)((n + kTile - 1) / kTile) : 1u;
    const uint32_t units = tiles * (uint32_t)ncomp;
    const bool vec = aligned16(x) && (pitch * sizeof(T)) % 16 == 0;
    StreamLease scratch(tree_scratch_pairs(tiles, ncomp) * sizeof(acc::Pair<T>), st);
    auto* part = static_cast<acc::Pair<T>*>(scratch.get());
    const int grid = grid_for(units);
    const bool nt = streaming((uint64_t)ncomp * n * sizeof(T));
    if (vec && nt)
        hipLaunchKernelGGL((tile_reduce_kernel<T, M, SQ, true, true>), dim3(grid), dim3(kThreads), 0, st, x,
                           (uint32_t)n, pitch, tiles, units, part);
    else if (vec)
        hipLaunchKernelGGL((tile_reduce_kernel<T, M, SQ, true, false>), dim3(grid), dim3(kThreads), 0, st, x,
                           (uint32_t)n, pitch, tiles, units, part);
    else
        hipLaunchKernelGGL((tile_reduce_kernel<T, M, SQ, false, false>), dim3(grid), dim3(kThreads), 0, st, x,
                           (uint32_t)n, pitch, tiles, units, part);
    check(hipGetLastError(), "tile_reduce_kernel launch");
    finish_tree<T, M>(part, tiles, ncomp, out, st);
}

template<class T, int M>
void dispatch_reduce_sq(const T* x, int ncomp, uint64_t n, uint64_t pitch, int square, T* out, hipStream_t st)
{
    if (square)
        launch_reduce<T, M, true>(x, ncomp, n, pitch, out, st);
    else
        launch_reduce<T, M, false>(x, ncomp, n, pitch, out, st);
}

template<class T>
void dispatch_reduce(int mode, int square, const void* x, int ncomp, uint64_t n, uint64_t pitch, void* out,
                     hipStream_t st)
{
    auto X = static_cast<const T*>(x);
    auto O = static_cast<T*>(out);
    switch (mode) {
    case AETH_SUM_NEUMAIER: dispatch_reduce_sq<T, acc::kNeumaier>(X, ncomp, n, pitch, square, O, st); break;
    case AETH_SUM_PLAIN: dispatch_reduce_sq<T, acc::kPlain>(X, ncomp, n, pitch, square, O, st); break;
    case AETH_SUM_FLOATFLOAT:
        if constexpr (sizeof(T) == 4) {
            dispatch_reduce_sq<T, acc::kFloatFloat>(X, ncomp, n, pitch, square, O, st);
            break;
        } else {
            fail("aeth_reduce", "float-float accumulation needs dtype F32", 0, 2);
        }
    default: fail("aeth_reduce", "unknown mode");
    }
}

// e = cross(a,b).dot(c) + sqrt(a.dot(a) + 1) (expr/nodes/Geometric.h:55-71, expr/Reduce.h:172-188:
// right-nested folds, nothing contracted).
__device__ __forceinline__ double vec3geo_expr(const double (&A)[3], const double (&B)[3], const double (&C)[3])
{
    const double x0 = A[1] * B[2] - A[2] * B[1];
    const double x1 = A[2] * B[0] - A[0] * B[2];
    const double x2 = A[0] * B[1] - A[1] * B[0];
    const double dxc = (x2 * C[2]) + ((x1 * C[1]) + (x0 * C[0]));
    const double daa = (A[2] * A[2]) + ((A[1] * A[1]) + (A[0] * A[0]));
    return dxc + __builtin_sqrt(daa + 1.0);
}

// One block per tile of kV3Tile samples: a full tile issues all 9 x kV3Rounds
// 16-byte loads of a lane before any arithmetic (each lane holds 2 samples per
// round; consecutive lanes take consecutive vectors of a plane), then writes
// its kV3Rounds 16-byte results. The last, partial tile goes element-wise.
constexpr int kV3Rounds = 2;
constexpr uint32_t kV3Tile = kThreads * 2 * kV3Rounds;

template<bool VECTOR, bool NT>
__global__ __launch_bounds__(kThreads) void vec3geo_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                           const double* __restrict__ c, double* __restrict__ e,
                                                           uint32_t n, uint64_t pitch)
{
    const uint64_t base = (uint64_t)xcd_unit(blockIdx.x, gridDim.x, xcd_chunk(kV3Tile * sizeof(double))) * kV3Tile;
    if (VECTOR && base + kV3Tile <= n) {
        double A[kV3Rounds][3][2], B[kV3Rounds][3][2], C[kV3Rounds][3][2];
#pragma unroll
        for (int r = 0; r < kV3Rounds; ++r) {
            const uint64_t i = base + 2 * (threadIdx.x + (uint64_t)r * kThreads);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                vload<NT>(a + (uint64_t)k * pitch + i, A[r][k]);
                vload<NT>(b + (uint64_t)k * pitch + i, B[r][k]);
                vload<NT>(c + (uint64_t)k * pitch + i, C[r][k]);
            }
        }
#pragma unroll
        for (int r = 0; r < kV3Rounds; ++r) {
            double out[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const double As[3] = {A[r][0][j], A[r][1][j], A[r][2][j]};
                const double Bs[3] = {B[r][0][j], B[r][1][j], B[r][2][j]};
                const double Cs[3] = {C[r][0][j], C[r][1][j], C[r][2][j]};
                out[j] = vec3geo_expr(As, Bs, Cs);
            }
            vstore<NT>(e + base + 2 * (threadIdx.x + (uint64_t)r * kThreads), out);
        }
        return;
    }
    const uint64_t end = base + kV3Tile < n ? base + kV3Tile : n;
    for (uint64_t i = base + threadIdx.x; i < end; i += kThreads) {
        double As[3], Bs[3], Cs[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            As[k] = a[(uint64_t)k * pitch + i];
            Bs[k] = b[(uint64_t)k * pitch + i];
            Cs[k] = c[(uint64_t)k * pitch + i];
        }
        e[i] = vec3geo_expr(As, Bs, Cs);
    }
}

} // namespace
} // namespace aeth

using namespace aeth;

extern "C" int aeth_eval_cfg0(int dtype, const void* a, const void* b, const void* c, void* d, int ncomp, uint64_t n,
                              uint64_t pitch, int sum_mode, void* sums, void* stream)
{
    return guarded([&] {
        if (dtype == AETH_F64) {
            check_sizes<double>(ncomp, n, pitch, "aeth_eval_cfg0");
            dispatch_cfg0<double>(a, b, c, d, ncomp, n, pitch, sum_mode, sums, as_stream(stream));
        } else if (dtype == AETH_F32) {
            check_sizes<float>(ncomp, n, pitch, "aeth_eval_cfg0");
            dispatch_cfg0<float>(a, b, c, d, ncomp, n, pitch, sum_mode, sums, as_stream(stream));
        } else {
            fail("aeth_eval_cfg0", "unknown dtype");
        }
    });
}

// aeth_eval_cfg0 for callers that cannot allocate (code under graph capture, tight serving loops):
// the same kernels and bits, with the sums' partials in the fixed per-device `scratch` instead of a
// stream-ordered allocation. Calls no hipMalloc/hipFree; refuses sizes whose partials exceed the scratch.
extern "C" int aeth_eval_cfg0_noalloc(int dtype, const void* a, const void* b, const void* c, void* d, int ncomp,
                                      uint64_t n, uint64_t pitch, int sum_mode, void* sums, void* stream)
{
    return guarded([&] {
        void* fixed = scratch_on_current_device();
        if (dtype == AETH_F64) {
            check_sizes<double>(ncomp, n, pitch, "aeth_eval_cfg0_noalloc");
            dispatch_cfg0<double>(a, b, c, d, ncomp, n, pitch, sum_mode, sums, as_stream(stream), fixed);
        } else if (dtype == AETH_F32) {
            check_sizes<float>(ncomp, n, pitch, "aeth_eval_cfg0_noalloc");
            dispatch_cfg0<float>(a, b, c, d, ncomp, n, pitch, sum_mode, sums, as_stream(stream), fixed);
        } else {
            fail("aeth_eval_cfg0_noalloc", "unknown dtype");
        }
    });
}

extern "C" int aeth_reduce(int dtype, int mode, int square, const void* x, int ncomp, uint64_t n, uint64_t pitch,
                           void* out, void* stream)
{
    return guarded([&] {
        if (!out) fail("aeth_reduce", "out is NULL");
        if (dtype == AETH_F64) {
            check_sizes<double>(ncomp, n, pitch, "aeth_reduce");
            dispatch_reduce<double>(mode, square, x, ncomp, n, pitch, out, as_stream(stream));
        } else if (dtype == AETH_F32) {
            check_sizes<float>(ncomp, n, pitch, "aeth_reduce");
            dispatch_reduce<float>(mode, square, x, ncomp, n, pitch, out, as_stream(stream));
        } else {
            fail("aeth_reduce", "unknown dtype");
        }
    });
}

extern "C" int aeth_eval_vec3geo(const void* a, const void* b, const void* c, void* e, uint64_t n, uint64_t pitch,
                                 void* stream)
{
    return guarded([&] {
        check_sizes<double>(3, n, pitch, "aeth_eval_vec3geo");
        if (n == 0) return;
        auto A = static_cast<const double*>(a), B = static_cast<const double*>(b), C = static_cast<const double*>(c);
        auto E = static_cast<double*>(e);
        const bool vec = aligned16(a) && aligned16(b) && aligned16(c) && aligned16(e) && pitch % 2 == 0;
        const bool nt = streaming(80ull * n);
        const dim3 grid((unsigned)((n + kV3Tile - 1) / kV3Tile));
        hipStream_t st = as_stream(stream);
        if (vec && nt)
            hipLaunchKernelGGL((vec3geo_kernel<true, true>), grid, dim3(kThreads), 0, st, A, B, C, E, (uint32_t)n, pitch);
        else if (vec)
            hipLaunchKernelGGL((vec3geo_kernel<true, false>), grid, dim3(kThreads), 0, st, A, B, C, E, (uint32_t)n, pitch);
        else
            hipLaunchKernelGGL((vec3geo_kernel<false, false>), grid, dim3(kThreads), 0, st, A, B, C, E, (uint32_t)n,
                               pitch);
        check(hipGetLastError(), "vec3geo_kernel launch");
    });
}
