// units.hpp -- the per-tile bodies shared by the single-call kernels
// (eval.hip) and the batched serving kernels (serve.hip). Both paths run the
// exact same code for a (component, tile) unit, which is what makes a job's
// result identical alone and inside a served batch.
#pragma once

#include <hip/hip_runtime.h>

#include "accum.hpp"
#include "dmath.hpp"

namespace aeth {
namespace units {

constexpr int kTile = 2048;
constexpr int kThreads = 256;
constexpr int kGroup = 1024;

// Whether the tile of `tile_len` samples starting at s0 lies wholly below n. Sample counts reach 2^32-1
// (the 32-bit offset_t span), so the end of the last tile is formed in 64 bits: s0 + tile_len can
// exceed 2^32 and must not wrap.
__host__ __device__ constexpr bool full_tile(uint32_t s0, uint32_t tile_len, uint64_t n)
{
    return (uint64_t)s0 + tile_len <= n;
}

// XCD-aware unit order. MI355X's dispatcher hands workgroups to its eight XCDs (each with its own L2)
// round robin, workgroup b to XCD b mod 8. In launch order every XCD would take every eighth unit;
// remapped, XCD x takes the K consecutive units x*K .. x*K+K-1 of each span of 8K units, so each XCD
// streams through its own contiguous stretch of every plane. Measured on MI355X over 3 x 2^27 doubles:
// the 3-read/1-write fused pass with Neumaier sums 2.15 -> 2.10 ms (K = 32 units of 2048 samples),
// 9-read/1-write 1.95 -> 1.86 ms, write-only 0.55 -> 0.49 ms; a stretch of about 512 KiB of one plane
// was the best chunk at every tile size tried. The map is a permutation of [0, units) (the last, partial
// span keeps launch order) and a unit's work does not depend on the block that runs it, so results are
// the same bits under any K and any grid.
__host__ __device__ constexpr uint32_t xcd_unit(uint32_t v, uint32_t units, uint32_t K)
{
    const uint32_t span = 8u * K;
    if (K <= 1u || v >= units - units % span) return v;
    const uint32_t w = v % span;
    return v - w + (w % 8u) * K + w / 8u;
}

// K for units whose plane stretch is `unit_bytes` (samples per tile x element size).
__host__ __device__ constexpr uint32_t xcd_chunk(uint32_t unit_bytes)
{
    return unit_bytes >= (512u << 10) ? 1u : (512u << 10) / unit_bytes;
}

template<class T>
struct V16 {
    static constexpr int n = 16 / sizeof(T);
};

// 16-byte lane vectors. NT selects the nontemporal (streaming) cache policy,
// used when the operands are far larger than the 256 MB last-level cache:
// measured on MI355X, a 3-read/1-write stream gains ~8% with it.
template<class T>
using ext16 = T __attribute__((ext_vector_type(16 / sizeof(T))));

template<bool NT = false, class T>
__device__ __forceinline__ void vload(const T* p, T (&v)[16 / sizeof(T)])
{
    using VT = ext16<T>;
    VT x;
    if constexpr (NT) x = __builtin_nontemporal_load(reinterpret_cast<const VT*>(p));
    else x = *reinterpret_cast<const VT*>(p);
#pragma unroll
    for (int k = 0; k < (int)(16 / sizeof(T)); ++k) v[k] = x[k];
}

template<bool NT = false, class T>
__device__ __forceinline__ void vstore(T* p, const T (&v)[16 / sizeof(T)])
{
    using VT = ext16<T>;
    VT x;
#pragma unroll
    for (int k = 0; k < (int)(16 / sizeof(T)); ++k) x[k] = v[k];
    if constexpr (NT) __builtin_nontemporal_store(x, reinterpret_cast<VT*>(p));
    else *reinterpret_cast<VT*>(p) = x;
}

// The benchmark expression, spelled as aether evaluates it:
// Sum(Sum(a, Product(b, c)), Product(CWiseSin(a), CWiseExp(CWiseScale(-1, Product(b, b))))).
template<class T>
__device__ __forceinline__ T cfg0_expr(T a, T b, T c)
{
    const T lhs = a + b * c;
    const T nb2 = T(-1) * (b * b);
    if constexpr (sizeof(T) == 8) {
        return lhs + dm::sin(a) * dm::exp(nb2);
    } else {
        return lhs + dm::sinf(a) * dm::expf(nb2);
    }
}

// d = cfg0(a, b, c) over one (component, tile) unit; returns this lane's
// accumulation pair (combine with acc::block_combine). VECTOR: planes 16-byte
// aligned; the tail tile always takes the predicated path.
template<class T, int M, bool SUM, bool VECTOR, bool NT>
__device__ __forceinline__ acc::Pair<T> cfg0_tile(const T* __restrict__ a, const T* __restrict__ b,
                                                  const T* __restrict__ c, T* __restrict__ d, uint32_t n,
                                                  uint64_t pitch, uint32_t comp, uint32_t tile)
{
    constexpr int VN = V16<T>::n;
    constexpr int ROUNDS = kTile / (kThreads * VN);
    const uint32_t s0 = tile * kTile;
    const uint64_t base = (uint64_t)comp * pitch + s0;
    acc::Pair<T> p{T(0), T(0)};
    if (VECTOR && full_tile(s0, kTile, n)) {
        T va[ROUNDS][VN], vb[ROUNDS][VN], vc[ROUNDS][VN];
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            const uint64_t o = base + (uint64_t)r * (kThreads * VN) + (uint64_t)threadIdx.x * VN;
            vload<NT>(a + o, va[r]);
            vload<NT>(b + o, vb[r]);
            vload<NT>(c + o, vc[r]);
        }
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            T vd[VN];
#pragma unroll
            for (int k = 0; k < VN; ++k) {
                vd[k] = cfg0_expr(va[r][k], vb[r][k], vc[r][k]);
                if constexpr (SUM) acc::accumulate<M, false>(p, vd[k]);
            }
            const uint64_t o = base + (uint64_t)r * (kThreads * VN) + (uint64_t)threadIdx.x * VN;
            vstore<NT>(d + o, vd);
        }
    } else {
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
#pragma unroll
            for (int k = 0; k < VN; ++k) {
                const uint32_t i = s0 + r * (kThreads * VN) + threadIdx.x * VN + k;
                if (i < n) {
                    const uint64_t o = (uint64_t)comp * pitch + i;
                    const T v = cfg0_expr(a[o], b[o], c[o]);
                    d[o] = v;
                    if constexpr (SUM) acc::accumulate<M, false>(p, v);
                }
            }
        }
    }
    return p;
}

// Sum (or sum of squares) of one (component, tile) unit; lane pair.
template<class T, int M, bool SQUARE, bool VECTOR, bool NT>
__device__ __forceinline__ acc::Pair<T> reduce_tile(const T* __restrict__ x, uint32_t n, uint64_t pitch,
                                                    uint32_t comp, uint32_t tile)
{
    constexpr int VN = V16<T>::n;
    constexpr int ROUNDS = kTile / (kThreads * VN);
    const uint32_t s0 = tile * kTile;
    const uint64_t base = (uint64_t)comp * pitch + s0;
    acc::Pair<T> p{T(0), T(0)};
    if (VECTOR && full_tile(s0, kTile, n)) {
        T v[ROUNDS][VN];
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r)
            vload<NT>(x + base + (uint64_t)r * (kThreads * VN) + (uint64_t)threadIdx.x * VN, v[r]);
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r)
#pragma unroll
            for (int k = 0; k < VN; ++k) acc::accumulate<M, SQUARE>(p, v[r][k]);
    } else {
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r)
#pragma unroll
            for (int k = 0; k < VN; ++k) {
                const uint32_t i = s0 + r * (kThreads * VN) + threadIdx.x * VN + k;
                if (i < n) acc::accumulate<M, SQUARE>(p, x[(uint64_t)comp * pitch + i]);
            }
    }
    return p;
}

// Halving tree over m <= 1024 pairs already in LDS `s` (512 threads):
// (t, t + stride), stride = 512 .. 1, absent partner = nothing. Result in s[0].
template<class T, int M>
__device__ __forceinline__ void group_halving(acc::Pair<T>* s, uint32_t m)
{
    const uint32_t t = threadIdx.x;
    for (uint32_t stride = kGroup / 2; stride >= 1; stride >>= 1) {
        __syncthreads();
        if (t < stride && t + stride < m) s[t] = acc::combine<M>(s[t], s[t + stride]);
    }
    __syncthreads();
}

} // namespace units
} // namespace aeth
