// runtime.hip -- the runtime elementwise evaluator over strided N-d views (eval/Runtime.h:1-342,
// view/RuntimeView.h:71-120): out = a (with a double <-> float cast), a + b, a - b, s * a, a + s * b,
// a - s * b, where every operand has its own element strides over one logical row-major shape.
//
// Dense row-major operands of one dtype take a flat kernel of 16 KiB units of 16-byte lanes; any other layout takes
// the general kernel, which splits each flat index into a multi-index over the output's extents and
// addresses every operand through its own strides (the reference's addressing). Nothing is contracted:
// a + s * b rounds the product, then the sum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <type_traits>

#include "../../include/aether_hip.h"
#include "common.hpp"
#include "units.hpp"

namespace aeth {
namespace {

constexpr int kThreads = 256;
constexpr int kRank = AETH_RUNTIME_RANK_CAP;

struct Strided {
    const void* data;
    int64_t strides[kRank];
};

struct Shape {
    int32_t rank;
    uint32_t extents[kRank];
};

template<class TIn, class TOut>
__device__ __forceinline__ TOut apply(int op, TIn a, TIn b, TIn s)
{
    switch (op) {
    case AETH_RT_ASSIGN: return (TOut)a;
    case AETH_RT_ADD: return (TOut)(a + b);
    case AETH_RT_SUB: return (TOut)(a - b);
    case AETH_RT_SCALE: return (TOut)(s * a);
    case AETH_RT_ADD_SCALED: return (TOut)(a + s * b);
    default: return (TOut)(a - s * b);
    }
}

template<class TIn, class TOut>
__global__ __launch_bounds__(kThreads) void strided_kernel(int op, Strided A, Strided B, Strided O, Shape sh, TIn s,
                                                           uint64_t total, bool uses_b)
{
    for (uint64_t linear = (uint64_t)blockIdx.x * kThreads + threadIdx.x; linear < total;
         linear += (uint64_t)gridDim.x * kThreads) {
        uint64_t rem = linear;
        int64_t oa = 0, ob = 0, oo = 0;
        for (int k = sh.rank - 1; k >= 0; --k) {
            const uint64_t e = sh.extents[k];
            const int64_t idx = (int64_t)(rem % e);
            rem /= e;
            oa += idx * A.strides[k];
            ob += idx * B.strides[k];
            oo += idx * O.strides[k];
        }
        const TIn av = static_cast<const TIn*>(A.data)[oa];
        const TIn bv = uses_b ? static_cast<const TIn*>(B.data)[ob] : TIn(0);
        static_cast<TOut*>(const_cast<void*>(O.data))[oo] = apply<TIn, TOut>(op, av, bv, s);
    }
}

// Dense, one dtype: 16 KiB units (four 16-byte vectors a lane, every load of a unit issued before its
// stores), one workgroup a unit in the XCD-aware unit order of the evaluation kernels; the vectors past the
// last whole unit grid-stride, then the scalar tail. The output may be one of the inputs (an in-place
// update): every element is read and written by the same lane, its loads before its store, so the
// pointers carry no __restrict__. NT: the streaming cache policy, for operands far beyond the last-level
// cache.
constexpr uint32_t kUnitVecs = kThreads * 4;

template<class T, bool NT>
__global__ __launch_bounds__(kThreads) void dense_kernel(int op, const T* a, const T* b, T* o, T s, uint64_t total,
                                                         uint32_t units, bool uses_b)
{
    using V = T __attribute__((ext_vector_type(16 / sizeof(T))));
    constexpr int VN = 16 / sizeof(T);
    constexpr uint32_t K = units::xcd_chunk(kUnitVecs * 16);
    const V* av = reinterpret_cast<const V*>(a);
    const V* bv = reinterpret_cast<const V*>(b);
    V* ov = reinterpret_cast<V*>(o);
    const uint64_t vecs = total / VN;
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const uint64_t base = (uint64_t)units::xcd_unit(u, units, K) * kUnitVecs + threadIdx.x;
        V x[4], y[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if constexpr (NT) x[r] = __builtin_nontemporal_load(av + base + r * kThreads);
            else x[r] = av[base + r * kThreads];
            if (uses_b) {
                if constexpr (NT) y[r] = __builtin_nontemporal_load(bv + base + r * kThreads);
                else y[r] = bv[base + r * kThreads];
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            V res;
#pragma unroll
            for (int k = 0; k < VN; ++k) res[k] = apply<T, T>(op, x[r][k], uses_b ? y[r][k] : T(0), s);
            if constexpr (NT) __builtin_nontemporal_store(res, ov + base + r * kThreads);
            else ov[base + r * kThreads] = res;
        }
    }
    for (uint64_t v = (uint64_t)units * kUnitVecs + (uint64_t)blockIdx.x * kThreads + threadIdx.x; v < vecs;
         v += (uint64_t)gridDim.x * kThreads) {
        const V x = av[v];
        V y;
        if (uses_b) y = bv[v];
        V r;
#pragma unroll
        for (int k = 0; k < VN; ++k) r[k] = apply<T, T>(op, x[k], uses_b ? y[k] : T(0), s);
        ov[v] = r;
    }
    const uint64_t t = vecs * VN + (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (blockIdx.x == 0 && t < total) o[t] = apply<T, T>(op, a[t], uses_b ? b[t] : T(0), s);
}

uint64_t flat(const aeth_runtime_view& v)
{
    uint64_t t = 1;
    for (int k = 0; k < v.rank; ++k) t *= v.extents[k];
    return t;
}

bool same_shape(const aeth_runtime_view& x, const aeth_runtime_view& y)
{
    if (x.rank != y.rank) return false;
    for (int k = 0; k < x.rank; ++k)
        if (x.extents[k] != y.extents[k]) return false;
    return true;
}

bool dense(const aeth_runtime_view& v)
{
    int64_t want = 1;
    for (int k = v.rank - 1; k >= 0; --k) {
        if (v.extents[k] != 1 && v.strides[k] != want) return false;
        want *= v.extents[k];
    }
    return true;
}

void check_view(const aeth_runtime_view& v, const char* name)
{
    if (v.rank < 0 || v.rank > kRank)
        fail("aeth_runtime_eval", std::string(name) + ": rank must be 0.." + std::to_string(kRank));
    if (v.dtype != AETH_F64 && v.dtype != AETH_F32)
        fail("aeth_runtime_eval", std::string(name) + ": unsupported dtype (double/float only)");
    for (int k = 0; k < v.rank; ++k)
        if (v.strides[k] < 0) fail("aeth_runtime_eval", std::string(name) + ": negative strides are not supported");
    if (flat(v) && !v.data) fail("aeth_runtime_eval", std::string(name) + ": NULL data");
    if (flat(v)) {
        hipPointerAttribute_t at;
        if (hipPointerGetAttributes(&at, v.data) != hipSuccess ||
            (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged)) {
            (void)hipGetLastError();
            fail("aeth_runtime_eval", std::string(name) + ": data is not device memory");
        }
    }
}

Strided strided(const aeth_runtime_view& v)
{
    Strided s{v.data, {}};
    for (int k = 0; k < kRank; ++k) s.strides[k] = k < v.rank ? v.strides[k] : 0;
    return s;
}

template<class TIn, class TOut>
void launch(int op, const aeth_runtime_view& a, const aeth_runtime_view& b, const aeth_runtime_view& o, double scalar,
            bool uses_b, hipStream_t st)
{
    const uint64_t total = flat(o);
    if (total == 0) return;
    const TIn s = (TIn)scalar;
    const unsigned grid = (unsigned)std::min<uint64_t>((total + kThreads - 1) / kThreads, (uint64_t)device_cus() * 64u);
    constexpr bool same = std::is_same_v<TIn, TOut>;
    if constexpr (same) {
        const bool vec = dense(a) && dense(o) && (!uses_b || dense(b)) && aligned16(a.data) && aligned16(o.data) &&
                         (!uses_b || aligned16(b.data));
        if (vec) {
            const uint64_t vecs = total / (16 / sizeof(TIn));
            const uint32_t units = (uint32_t)std::min<uint64_t>(vecs / kUnitVecs, 0x7FFFFFFFull);
            const unsigned dgrid = units ? units : grid;
            const bool nt = (uses_b ? 3 : 2) * total * sizeof(TIn) >= (512ull << 20);
            const TIn* A = static_cast<const TIn*>(a.data);
            const TIn* B = static_cast<const TIn*>(uses_b ? b.data : a.data);
            TIn* O = static_cast<TIn*>(o.data);
            if (nt)
                hipLaunchKernelGGL((dense_kernel<TIn, true>), dim3(dgrid), dim3(kThreads), 0, st, op, A, B, O, s, total,
                                   units, uses_b);
            else
                hipLaunchKernelGGL((dense_kernel<TIn, false>), dim3(dgrid), dim3(kThreads), 0, st, op, A, B, O, s,
                                   total, units, uses_b);
            check(hipGetLastError(), "dense_kernel launch");
            return;
        }
    }
    Shape sh{o.rank, {}};
    for (int k = 0; k < kRank; ++k) sh.extents[k] = k < o.rank ? o.extents[k] : 1;
    hipLaunchKernelGGL((strided_kernel<TIn, TOut>), dim3(grid), dim3(kThreads), 0, st, op, strided(a),
                       uses_b ? strided(b) : strided(a), strided(o), sh, s, total, uses_b);
    check(hipGetLastError(), "strided_kernel launch");
}

} // namespace
} // namespace aeth

using namespace aeth;

extern "C" int aeth_runtime_eval(int op, const aeth_runtime_view* a, const aeth_runtime_view* b,
                                 const aeth_runtime_view* out, double scalar, void* stream)
{
    return guarded([&] {
        if (op < AETH_RT_ASSIGN || op > AETH_RT_SUB_SCALED) fail("aeth_runtime_eval", "unknown op");
        if (!a || !out) fail("aeth_runtime_eval", "a and out are required");
        const bool uses_b = op == AETH_RT_ADD || op == AETH_RT_SUB || op == AETH_RT_ADD_SCALED || op == AETH_RT_SUB_SCALED;
        if (uses_b && !b) fail("aeth_runtime_eval", "this op reads b");
        check_view(*out, "out");
        check_view(*a, "a");
        if (uses_b) check_view(*b, "b");
        if (!same_shape(*a, *out)) fail("aeth_runtime_eval", "shape mismatch: a vs out");
        if (uses_b && !same_shape(*b, *out)) fail("aeth_runtime_eval", "shape mismatch: b vs out");
        if (flat(*out) > 0xFFFFFFFFull) fail("aeth_runtime_eval", "flat element count exceeds the 32-bit offset_t");
        hipStream_t st = as_stream(stream);
        const aeth_runtime_view& B = uses_b ? *b : *a;
        if (op == AETH_RT_ASSIGN) {
            if (a->dtype == AETH_F64 && out->dtype == AETH_F64) launch<double, double>(op, *a, B, *out, scalar, false, st);
            else if (a->dtype == AETH_F32 && out->dtype == AETH_F32) launch<float, float>(op, *a, B, *out, scalar, false, st);
            else if (a->dtype == AETH_F64) launch<double, float>(op, *a, B, *out, scalar, false, st);
            else launch<float, double>(op, *a, B, *out, scalar, false, st);
            return;
        }
        if (a->dtype != out->dtype) fail("aeth_runtime_eval", "dtype mismatch: a vs out (only ASSIGN casts)");
        if (uses_b && b->dtype != out->dtype) fail("aeth_runtime_eval", "dtype mismatch: b vs out");
        if (out->dtype == AETH_F64) launch<double, double>(op, *a, B, *out, scalar, uses_b, st);
        else launch<float, float>(op, *a, B, *out, scalar, uses_b, st);
    });
}
