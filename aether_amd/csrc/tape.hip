// tape.hip -- generic lowering of `dest = expr` to one gfx950 kernel.
//
// aether evaluates an assignment per sample with the component loop unrolled
// at compile time (expr/Assign.h:124-160) and its runtime evaluator applies
// one of six fixed ops (eval/Runtime.h:73-167). Here any expression tree the
// host lowers becomes one PROGRAM per output component: the planes it reads
// and a postfix tape. A block owns one (program, 2048-sample tile) unit; all
// planes of the program are loaded first with 16-byte lanes (issued together,
// so their latencies overlap), then the tape runs on a register stack. The
// tape is wave-uniform (it lives in the kernel arguments), so dispatch is
// scalar branching only; each transcendental is instantiated once and fed
// through a register. The tile and sum shape are those of eval.hip, so a tape
// and the specialised kernel for the same tree agree bit for bit.
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "jit_prelude.inc"

#include "../../include/aether_hip.h"
#include "accum.hpp"
#include "common.hpp"
#include "dmath.hpp"
#include "rng.hpp"
#include "units.hpp"

namespace aeth {
namespace {

constexpr int kTile = 2048;
constexpr int kThreads = 256;
constexpr int kGroup = 1024;
constexpr int kDepth = AETH_TAPE_MAX_DEPTH;

struct TapeArgs {
    aeth_program prog[AETH_TAPE_MAX_PROGRAMS];
    const void* leaf[AETH_TAPE_MAX_LEAVES];
    uint64_t pitch[AETH_TAPE_MAX_LEAVES];
    double consts[AETH_TAPE_MAX_CONSTS];
    aeth_rng_leaf rng[AETH_TAPE_MAX_RNGS];
    uint32_t n, tiles, units;
};
static_assert(sizeof(TapeArgs) < 4000, "kernel arguments must fit the 4 KiB kernarg segment");

// Arguments of a runtime-compiled tape kernel: every plane a program touches resolved on the host to
// its address (no dependent reads of descriptors before a block's first load), the constants, the
// random leaves and, per program, the leaf index of each RNG instruction.
struct JitArgs {
    const void* plane[AETH_TAPE_MAX_PROGRAMS][1 + AETH_TAPE_MAX_LOADS]; // [p][0] output, [p][1 + q] load q
    double consts[AETH_TAPE_MAX_CONSTS];
    aeth_rng_leaf rng[AETH_TAPE_MAX_RNGS];
    uint8_t rng_arg[AETH_TAPE_MAX_PROGRAMS][AETH_TAPE_MAX_INSNS];
    uint32_t n, tiles, units;
};
static_assert(sizeof(JitArgs) < 4000, "kernel arguments must fit the 4 KiB kernarg segment");

template<class T>
using ext16 = T __attribute__((ext_vector_type(16 / sizeof(T))));

template<bool NT, class T, int VN>
__device__ __forceinline__ void vload(const T* p, T (&v)[VN])
{
    using VT = ext16<T>;
    VT x;
    if constexpr (NT) x = __builtin_nontemporal_load(reinterpret_cast<const VT*>(p));
    else x = *reinterpret_cast<const VT*>(p);
#pragma unroll
    for (int k = 0; k < VN; ++k) v[k] = x[k];
}

template<bool NT, class T, int VN>
__device__ __forceinline__ void vstore(T* p, const T (&v)[VN])
{
    using VT = ext16<T>;
    VT x;
#pragma unroll
    for (int k = 0; k < VN; ++k) x[k] = v[k];
    if constexpr (NT) __builtin_nontemporal_store(x, reinterpret_cast<VT*>(p));
    else *reinterpret_cast<VT*>(p) = x;
}

template<class T, int VN>
__device__ __forceinline__ void get(const T (&S)[kDepth][VN], int d, T (&x)[VN])
{
    switch (d) {
#define AETH_GET(D)                                                                                                   \
    case D:                                                                                                           \
        for (int k = 0; k < VN; ++k) x[k] = S[D][k];                                                                  \
        break;
        AETH_GET(0) AETH_GET(1) AETH_GET(2) AETH_GET(3) AETH_GET(4) AETH_GET(5) AETH_GET(6) AETH_GET(7)
#undef AETH_GET
    default: break;
    }
}

template<class T, int VN>
__device__ __forceinline__ void put(T (&S)[kDepth][VN], int d, const T (&x)[VN])
{
    switch (d) {
#define AETH_PUT(D)                                                                                                   \
    case D:                                                                                                           \
        for (int k = 0; k < VN; ++k) S[D][k] = x[k];                                                                  \
        break;
        AETH_PUT(0) AETH_PUT(1) AETH_PUT(2) AETH_PUT(3) AETH_PUT(4) AETH_PUT(5) AETH_PUT(6) AETH_PUT(7)
#undef AETH_PUT
    default: break;
    }
}

__host__ __device__ constexpr bool is_unary(int op)
{
    return (op >= AETH_OP_NEG && op <= AETH_OP_ABS) || (op >= AETH_OP_SIGN && op <= AETH_OP_TANH);
}
__host__ __device__ constexpr bool is_binary(int op)
{
    return (op >= AETH_OP_ADD && op <= AETH_OP_DIV) || (op >= AETH_OP_FMIN && op <= AETH_OP_POW);
}

template<class T>
__device__ __forceinline__ T unary(int op, T a)
{
    if constexpr (sizeof(T) == 8) {
        switch (op) {
        case AETH_OP_NEG: return -a;
        case AETH_OP_SIN: return dm::sin(a);
        case AETH_OP_COS: return dm::cos(a);
        case AETH_OP_EXP: return dm::exp(a);
        case AETH_OP_LOG: return dm::log(a);
        case AETH_OP_SQRT: return __builtin_sqrt(a);
        case AETH_OP_SIGN: return dm::sign(a);
        case AETH_OP_TAN: return dm::tan(a);
        case AETH_OP_ATAN: return dm::atan(a);
        case AETH_OP_ASIN: return dm::asin(a);
        case AETH_OP_ACOS: return dm::acos(a);
        case AETH_OP_TANH: return dm::tanh(a);
        default: return __builtin_fabs(a);
        }
    } else {
        switch (op) {
        case AETH_OP_NEG: return -a;
        case AETH_OP_SIN: return dm::sinf(a);
        case AETH_OP_COS: return dm::cosf(a);
        case AETH_OP_EXP: return dm::expf(a);
        case AETH_OP_LOG: return dm::logf(a);
        case AETH_OP_SQRT: return __builtin_sqrtf(a);
        case AETH_OP_SIGN: return dm::sign(a);
        case AETH_OP_TAN: return dm::tanf(a);
        case AETH_OP_ATAN: return dm::atanf(a);
        case AETH_OP_ASIN: return dm::asinf(a);
        case AETH_OP_ACOS: return dm::acosf(a);
        case AETH_OP_TANH: return dm::tanhf(a);
        default: return __builtin_fabsf(a);
        }
    }
}

template<class T>
__device__ __forceinline__ T binary(int op, T a, T b)
{
    switch (op) {
    case AETH_OP_ADD: return a + b;
    case AETH_OP_SUB: return a - b;
    case AETH_OP_MUL: return a * b;
    case AETH_OP_FMIN: return dm::fmin(a, b);
    case AETH_OP_FMAX: return dm::fmax(a, b);
    case AETH_OP_SELGT: return dm::selgt(a, b);
    case AETH_OP_POW:
        if constexpr (sizeof(T) == 8) return dm::pow(a, b);
        else return dm::powf(a, b);
    default: return a / b;
    }
}

template<class T>
__device__ __forceinline__ T draw(const aeth_rng_leaf& r, uint32_t i)
{
    const rng::U4 b = rng::block(r.seed, i, r.comp);
    if constexpr (sizeof(T) == 8) {
        if (r.dist == AETH_DIST_UNIFORM) return r.p0 + (r.p1 - r.p0) * rng::u01_from_word_d(b.x);
        return r.p0 + r.p1 * rng::normal_from_block_d(b);
    } else {
        const float p0 = (float)r.p0, p1 = (float)r.p1;
        if (r.dist == AETH_DIST_UNIFORM) return p0 + (p1 - p0) * rng::u01_from_word_f(b.x);
        return p0 + p1 * rng::normal_from_block_f(b);
    }
}

// Run program P on VN samples. L: the program's planes for these samples.
template<class T, int VN, int NL>
__device__ __forceinline__ void run(const aeth_program& P, const TapeArgs& A, const T (&L)[NL][VN],
                                    uint32_t first_sample, T (&out)[VN])
{
    T S[kDepth][VN];
    int d = 0;
    const int ni = P.n_insns;
    for (int pc = 0; pc < ni; ++pc) {
        const aeth_insn w = P.insns[pc];
        const int op = w.op;
        T x[VN];
        if (op == AETH_OP_PUSH) {
            const int j = w.arg;
#pragma unroll
            for (int q = 0; q < NL; ++q)
                if (q == j)
#pragma unroll
                    for (int k = 0; k < VN; ++k) x[k] = L[q][k];
            put(S, d, x);
            ++d;
        } else if (op == AETH_OP_CONST) {
            const T c = (T)A.consts[w.arg];
#pragma unroll
            for (int k = 0; k < VN; ++k) x[k] = c;
            put(S, d, x);
            ++d;
        } else if (op == AETH_OP_RNG) {
            const aeth_rng_leaf& r = A.rng[w.arg];
#pragma unroll
            for (int k = 0; k < VN; ++k) x[k] = draw<T>(r, first_sample + k);
            put(S, d, x);
            ++d;
        } else if (is_unary(op)) {
            get(S, d - 1, x);
#pragma unroll
            for (int k = 0; k < VN; ++k) x[k] = unary<T>(op, x[k]);
            put(S, d - 1, x);
        } else {
            T y[VN];
            get(S, d - 2, x);
            get(S, d - 1, y);
#pragma unroll
            for (int k = 0; k < VN; ++k) x[k] = binary<T>(op, x[k], y[k]);
            put(S, d - 2, x);
            --d;
        }
    }
#pragma unroll
    for (int k = 0; k < VN; ++k) out[k] = S[0][k];
}

template<class T, int M, bool SUM, int NL, bool VECTOR, bool NT>
__global__ __launch_bounds__(kThreads) void tape_kernel(const TapeArgs A, acc::Pair<T>* __restrict__ part)
{
    constexpr int VN = 16 / sizeof(T);
    constexpr int ROUNDS = kTile / (kThreads * VN);
    __shared__ acc::Pair<T> lds[4];
    const uint32_t n = A.n;
    constexpr uint32_t K = units::xcd_chunk(kTile * sizeof(T));
    for (uint32_t v = blockIdx.x; v < A.units; v += gridDim.x) {
        const uint32_t u = units::xcd_unit(v, A.units, K);
        const uint32_t pi = u / A.tiles;
        const uint32_t tile = u - pi * A.tiles;
        const uint32_t s0 = tile * kTile;
        const aeth_program& P = A.prog[pi];
        const int nl = P.n_loads;
        T* outp = static_cast<T*>(const_cast<void*>(A.leaf[P.out_leaf])) + (uint64_t)P.out_comp * A.pitch[P.out_leaf];
        acc::Pair<T> acc{T(0), T(0)};
        const bool full = VECTOR && units::full_tile(s0, kTile, n);
        for (int r = 0; r < ROUNDS; ++r) {
            const uint32_t i0 = s0 + r * (kThreads * VN) + threadIdx.x * VN;
            T L[NL][VN];
#pragma unroll
            for (int q = 0; q < NL; ++q) {
                if (q < nl) {
                    const T* src = static_cast<const T*>(A.leaf[P.load_leaf[q]]) +
                                   (uint64_t)P.load_comp[q] * A.pitch[P.load_leaf[q]];
                    if (full) {
                        vload<NT>(src + i0, L[q]);
                    } else {
#pragma unroll
                        for (int k = 0; k < VN; ++k) L[q][k] = (i0 + k < n) ? src[i0 + k] : T(0);
                    }
                }
            }
            T v[VN];
            run<T, VN, NL>(P, A, L, i0, v);
            if (full) {
                vstore<NT>(outp + i0, v);
                if constexpr (SUM)
#pragma unroll
                    for (int k = 0; k < VN; ++k) acc::accumulate<M, false>(acc, v[k]);
            } else {
#pragma unroll
                for (int k = 0; k < VN; ++k)
                    if (i0 + k < n) {
                        outp[i0 + k] = v[k];
                        if constexpr (SUM) acc::accumulate<M, false>(acc, v[k]);
                    }
            }
        }
        if constexpr (SUM) {
            const acc::Pair<T> rr = acc::block_combine<M>(acc, lds);
            if (threadIdx.x == 0) part[u] = rr;
        }
    }
}

// Group tree of eval.hip (same shape), repeated here to keep the unit self-contained.
template<class T, int M>
__global__ __launch_bounds__(512) void tape_tree_kernel(const acc::Pair<T>* __restrict__ in,
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
    for (uint32_t stride = kGroup / 2; stride >= 1; stride >>= 1) {
        __syncthreads();
        if (t < stride && t + stride < m) s[t] = acc::combine<M>(s[t], s[t + stride]);
    }
    if (t == 0) out[(uint64_t)comp * out_stride + g] = s[0];
}

template<class T>
__global__ void tape_store_pairs(const acc::Pair<T>* __restrict__ in, uint32_t in_stride, T* __restrict__ out, int np)
{
    const int c = threadIdx.x;
    if (c < np) {
        out[2 * c] = in[(uint64_t)c * in_stride].h;
        out[2 * c + 1] = in[(uint64_t)c * in_stride].c;
    }
}

// Host-side validation: stack effects, ranges, one value left.
void validate(const aeth_program& p, int n_leaves, int n_consts, int n_rngs)
{
    if (p.n_loads > AETH_TAPE_MAX_LOADS || p.n_insns > AETH_TAPE_MAX_INSNS || p.n_insns == 0)
        fail("aeth_eval_tape", "program size out of range");
    if (p.out_leaf >= n_leaves) fail("aeth_eval_tape", "output leaf out of range");
    for (int k = 0; k < p.n_loads; ++k)
        if (p.load_leaf[k] >= n_leaves) fail("aeth_eval_tape", "load leaf out of range");
    int d = 0;
    for (int i = 0; i < p.n_insns; ++i) {
        const int op = p.insns[i].op, arg = p.insns[i].arg;
        if (op == AETH_OP_PUSH) {
            if (arg >= p.n_loads) fail("aeth_eval_tape", "PUSH index out of range");
            ++d;
        } else if (op == AETH_OP_CONST) {
            if (arg >= n_consts) fail("aeth_eval_tape", "CONST index out of range");
            ++d;
        } else if (op == AETH_OP_RNG) {
            if (arg >= n_rngs) fail("aeth_eval_tape", "RNG index out of range");
            ++d;
        } else if (is_binary(op)) {
            if (d < 2) fail("aeth_eval_tape", "stack underflow");
            --d;
        } else if (is_unary(op)) {
            if (d < 1) fail("aeth_eval_tape", "stack underflow");
        } else {
            fail("aeth_eval_tape", "unknown opcode");
        }
        if (d > kDepth) fail("aeth_eval_tape", "stack deeper than 8");
    }
    if (d != 1) fail("aeth_eval_tape", "a program must leave exactly one value");
}

template<class T, int M, bool SUM, int NL>
void launch_nl(const TapeArgs& A, bool vec, bool nt, acc::Pair<T>* part, int grid, hipStream_t st)
{
    if (vec && nt)
        hipLaunchKernelGGL((tape_kernel<T, M, SUM, NL, true, true>), dim3(grid), dim3(kThreads), 0, st, A, part);
    else if (vec)
        hipLaunchKernelGGL((tape_kernel<T, M, SUM, NL, true, false>), dim3(grid), dim3(kThreads), 0, st, A, part);
    else
        hipLaunchKernelGGL((tape_kernel<T, M, SUM, NL, false, false>), dim3(grid), dim3(kThreads), 0, st, A, part);
}

// Per-launch timing of the tape kernels themselves (measurement; no counterpart in the reference). While
// enabled, every tape kernel launch outside a stream capture is bracketed by two HIP events on its stream
// (not the sums' group tree after it); aeth_tape_timing_read waits for the pending launches and hands back
// their elapsed milliseconds in launch order.
std::atomic<bool> g_timing{false};
std::mutex g_timing_mu;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_timed;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t pooled_event()
{
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

// A pair joins the pending list only once both events are recorded: a launch that throws in between
// returns its events to the pool and leaves nothing unreadable behind.
struct TimedLaunch {
    hipStream_t st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    explicit TimedLaunch(hipStream_t s) : st(s)
    {
        if (!g_timing.load(std::memory_order_relaxed) || capturing(s)) return;
        std::lock_guard<std::mutex> lock(g_timing_mu);
        e0 = pooled_event();
        e1 = pooled_event();
        if (hipEventRecord(e0, st) != hipSuccess) {
            release();
            fail("hipEventRecord", "timing the tape kernel");
        }
    }
    void done()
    {
        if (!e0) return;
        check(hipEventRecord(e1, st), "hipEventRecord");
        std::lock_guard<std::mutex> lock(g_timing_mu);
        g_timed.emplace_back(e0, e1);
        e0 = e1 = nullptr;
    }
    void release()
    {
        if (e0) g_event_pool.push_back(e0);
        if (e1) g_event_pool.push_back(e1);
        e0 = e1 = nullptr;
    }
    ~TimedLaunch()
    {
        if (!e0) return;
        std::lock_guard<std::mutex> lock(g_timing_mu);
        release();
    }
};

template<class T, int M, bool SUM>
void launch(const TapeArgs& A, int max_loads, int np, bool vec, bool nt, T* sums, hipStream_t st, hipFunction_t jit)
{
    size_t pairs = (size_t)A.tiles * np;
    for (uint32_t t = A.tiles; t > 1;) {
        t = (t + kGroup - 1) / kGroup;
        pairs += (size_t)t * np;
    }
    StreamLease scratch(SUM ? pairs * sizeof(acc::Pair<T>) : 0, st);
    auto* part = static_cast<acc::Pair<T>*>(scratch.get());
    const int grid = (int)std::max<uint32_t>(1u, A.units);
    TimedLaunch timed(st);
    if (jit) {
        JitArgs args;
        std::memset(&args, 0, sizeof args);
        for (int p = 0; p < np; ++p) {
            const aeth_program& P = A.prog[p];
            auto plane = [&](int leaf, int comp) {
                return static_cast<const void*>(static_cast<const T*>(A.leaf[leaf]) + (uint64_t)comp * A.pitch[leaf]);
            };
            args.plane[p][0] = plane(P.out_leaf, P.out_comp);
            for (int q = 0; q < P.n_loads; ++q) args.plane[p][1 + q] = plane(P.load_leaf[q], P.load_comp[q]);
            for (int i = 0; i < P.n_insns; ++i)
                if (P.insns[i].op == AETH_OP_RNG) args.rng_arg[p][i] = P.insns[i].arg;
        }
        std::memcpy(args.consts, A.consts, sizeof args.consts);
        std::memcpy(args.rng, A.rng, sizeof args.rng);
        args.n = A.n;
        args.tiles = A.tiles;
        args.units = A.units;
        void* part_v = part;
        int vec_i = vec ? 1 : 0;
        void* params[] = {&args, &part_v, &vec_i};
        check(hipModuleLaunchKernel(jit, (unsigned)grid, 1, 1, kThreads, 1, 1, 0, st, params, nullptr),
              "hipModuleLaunchKernel(tape)");
    } else if (max_loads <= 4) {
        launch_nl<T, M, SUM, 4>(A, vec, nt, part, grid, st);
    } else if (max_loads <= 9) {
        launch_nl<T, M, SUM, 9>(A, vec, nt, part, grid, st);
    } else {
        launch_nl<T, M, SUM, 16>(A, vec, nt, part, grid, st);
    }
    check(hipGetLastError(), "tape_kernel launch");
    timed.done();
    if constexpr (SUM) {
        acc::Pair<T>* cur = part;
        acc::Pair<T>* next = part + (size_t)A.tiles * np;
        uint32_t count = A.tiles, stride = A.tiles;
        while (count > 1) {
            const uint32_t groups = (count + kGroup - 1) / kGroup;
            hipLaunchKernelGGL((tape_tree_kernel<T, M>), dim3(groups, np), dim3(512), 0, st, cur, next, count, stride,
                               groups);
            cur = next;
            next += (size_t)groups * np;
            stride = groups;
            count = groups;
        }
        hipLaunchKernelGGL((tape_store_pairs<T>), dim3(1), dim3(64), 0, st, cur, stride, sums, np);
        check(hipGetLastError(), "tape tree launch");
    }
}

template<class T>
void dispatch(const TapeArgs& A, int max_loads, int np, bool vec, bool nt, int mode, T* sums, hipStream_t st,
              hipFunction_t jit)
{
    switch (mode) {
    case AETH_SUM_NONE: launch<T, acc::kNeumaier, false>(A, max_loads, np, vec, nt, sums, st, jit); break;
    case AETH_SUM_NEUMAIER: launch<T, acc::kNeumaier, true>(A, max_loads, np, vec, nt, sums, st, jit); break;
    case AETH_SUM_PLAIN: launch<T, acc::kPlain, true>(A, max_loads, np, vec, nt, sums, st, jit); break;
    case AETH_SUM_FLOATFLOAT:
        if constexpr (sizeof(T) == 4) {
            launch<T, acc::kFloatFloat, true>(A, max_loads, np, vec, nt, sums, st, jit);
            break;
        } else {
            fail("aeth_eval_tape", "float-float accumulation needs dtype F32", 0, 2);
        }
    default: fail("aeth_eval_tape", "unknown sum mode");
    }
}


// ---------------------------------------------------------------------------
// Runtime lowering: each distinct tape structure becomes one straight-line
// kernel compiled by hipRTC from the same device headers (kJitPrelude) and
// cached for the life of the process. Generated bodies follow the tape's
// order exactly, so they give the interpreter's bits.
struct JitEntry {
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
};

// The launch shape of a runtime-compiled tape kernel: samples per workgroup tile, bytes per vector access
// and the register allocator's occupancy target. Assignments with sums keep the contract tile and width
// (2048-sample tiles of 16-byte accesses: the reduction order is part of the numerics) and take the
// occupancy target aeth_tape_tune_sums picked, if any; assignments without sums take the shape
// aeth_tape_tune picked for their tape structure on this device, or the contract shape.
struct JitShape {
    uint32_t tile = 2048;
    uint32_t vb = 16;
    uint32_t waves = 0; // minimum waves per SIMD asked of the register allocator (0: the compiler's choice)
};
constexpr JitShape kShapes[] = {{1024, 16, 0}, {2048, 16, 0}, {4096, 16, 0}, {2048, 8, 0}, {4096, 8, 0},
                                {2048, 16, 8}};
// Assignments with sums: the contract tile and access width fix the sums' bits, so only the register
// allocator's occupancy target varies (it changes no operation).
constexpr JitShape kSumShapes[] = {{2048, 16, 0}, {2048, 16, 4}, {2048, 16, 8}};
std::mutex g_jit_mu;
std::map<std::string, JitEntry> g_jit;
std::map<std::string, JitShape> g_tuned; // base tape key -> tuned shape (guarded by g_jit_mu)
std::atomic<uint64_t> g_jit_compiled{0}, g_jit_hits{0}, g_jit_fallbacks{0};

// Programs that differ only in which planes they read and write, and in which random leaves they
// draw from, share one code path: the dataflow (op codes, which loaded plane each PUSH names, which
// constant each CONST names) is compiled in; plane addresses and random leaves come from the kernel
// arguments per program at run time. So the components of a vector expression run the same
// instructions, as the ahead-of-time kernels do.
std::string structure_key(const aeth_program& P)
{
    std::string k = std::to_string(P.n_loads) + ":";
    for (int i = 0; i < P.n_insns; ++i) {
        const int op = P.insns[i].op;
        k += std::to_string(op);
        if (op == AETH_OP_PUSH || op == AETH_OP_CONST) k += "." + std::to_string(P.insns[i].arg);
        k += " ";
    }
    return k;
}

std::string tape_key(int dtype, int mode, int np, const aeth_program* P, bool nt, int dev);
std::string tape_key(int dtype, int mode, int np, const aeth_program* P, bool nt, int dev, JitShape sh)
{
    return tape_key(dtype, mode, np, P, nt, dev) + ":" + std::to_string(sh.tile) + "x" + std::to_string(sh.vb) + "w" +
           std::to_string(sh.waves);
}

std::string tape_key(int dtype, int mode, int np, const aeth_program* P, bool nt, int dev)
{
    std::string k = std::to_string(dev) + ":" + std::to_string(dtype) + ":" + std::to_string(mode) + ":" +
                    std::to_string(np) + (nt ? ":nt" : ":t");
    // the generated code depends on each program's structure only (and on which programs share one)
    for (int p = 0; p < np; ++p) k += "|" + structure_key(P[p]);
    return k;
}

// The tape as straight-line code over one sample: `operand(q)` names loaded plane q's value, `sample` the
// sample index (for random leaves). Appends the statements to `b` and returns the name of the result.
std::string gen_body(const aeth_program& P, bool dbl, const std::function<std::string(int)>& operand,
                     const std::string& sample, const std::string& ind, std::string& b)
{
    std::vector<std::string> st;
    int tmp = 0;
    auto fresh = [&] { return "t" + std::to_string(tmp++); };
    for (int i = 0; i < P.n_insns; ++i) {
        const int op = P.insns[i].op, arg = P.insns[i].arg;
        if (op == AETH_OP_PUSH) {
            st.push_back(operand(arg));
        } else if (op == AETH_OP_CONST) {
            st.push_back("((T)A.consts[" + std::to_string(arg) + "])");
        } else if (op == AETH_OP_RNG) {
            const std::string t = fresh();
            b += ind + "const T " + t + " = jit_draw<T>(A.rng[A.rng_arg[pi][" + std::to_string(i) + "]], " + sample +
                 ");\n";
            st.push_back(t);
        } else if (is_binary(op)) {
            const std::string y = st.back();
            st.pop_back();
            const std::string x = st.back();
            st.pop_back();
            const std::string t = fresh();
            if (op >= AETH_OP_ADD && op <= AETH_OP_DIV) {
                const char* sym = op == AETH_OP_ADD ? "+" : op == AETH_OP_SUB ? "-" : op == AETH_OP_MUL ? "*" : "/";
                b += ind + "const T " + t + " = " + x + " " + sym + " " + y + ";\n";
            } else {
                const char* fn = op == AETH_OP_FMIN    ? "aeth::dm::fmin("
                                 : op == AETH_OP_FMAX  ? "aeth::dm::fmax("
                                 : op == AETH_OP_SELGT ? "aeth::dm::selgt("
                                 : dbl                 ? "aeth::dm::pow("
                                                       : "aeth::dm::powf(";
                b += ind + "const T " + t + " = " + fn + x + ", " + y + ");\n";
            }
            st.push_back(t);
        } else {
            const std::string x = st.back();
            st.pop_back();
            std::string f;
            switch (op) {
            case AETH_OP_NEG: f = "-(" + x + ")"; break;
            case AETH_OP_SIN: f = std::string(dbl ? "aeth::dm::sin(" : "aeth::dm::sinf(") + x + ")"; break;
            case AETH_OP_COS: f = std::string(dbl ? "aeth::dm::cos(" : "aeth::dm::cosf(") + x + ")"; break;
            case AETH_OP_EXP: f = std::string(dbl ? "aeth::dm::exp(" : "aeth::dm::expf(") + x + ")"; break;
            case AETH_OP_LOG: f = std::string(dbl ? "aeth::dm::log(" : "aeth::dm::logf(") + x + ")"; break;
            case AETH_OP_SQRT: f = std::string(dbl ? "__builtin_sqrt(" : "__builtin_sqrtf(") + x + ")"; break;
            case AETH_OP_SIGN: f = "aeth::dm::sign(" + x + ")"; break;
            case AETH_OP_TAN: f = std::string(dbl ? "aeth::dm::tan(" : "aeth::dm::tanf(") + x + ")"; break;
            case AETH_OP_ATAN: f = std::string(dbl ? "aeth::dm::atan(" : "aeth::dm::atanf(") + x + ")"; break;
            case AETH_OP_ASIN: f = std::string(dbl ? "aeth::dm::asin(" : "aeth::dm::asinf(") + x + ")"; break;
            case AETH_OP_ACOS: f = std::string(dbl ? "aeth::dm::acos(" : "aeth::dm::acosf(") + x + ")"; break;
            case AETH_OP_TANH: f = std::string(dbl ? "aeth::dm::tanh(" : "aeth::dm::tanhf(") + x + ")"; break;
            default: f = std::string(dbl ? "__builtin_fabs(" : "__builtin_fabsf(") + x + ")"; break;
            }
            const std::string t = fresh();
            b += ind + "const T " + t + " = " + f + ";\n";
            st.push_back(t);
        }
    }
    return st.back();
}

// One program over one tile, in the structure of the ahead-of-time kernels (units::cfg0_tile): a whole
// tile takes one branch that issues its 16-byte loads (all rounds at once for programs with few loads,
// round by round otherwise), evaluates, accumulates in sample order and stores; the last, partial tile
// goes sample by sample with scalar accesses.
std::string gen_program(const aeth_program& P, bool dbl, uint32_t vb)
{
    const bool v16 = vb == 16; // narrower accesses are element loops the compiler merges
    const char* T = dbl ? "double" : "float";
    std::string b;
    b += "      " + std::string(T) + "* outp = (" + T + "*)A.plane[pi][0];\n";
    for (int q = 0; q < P.n_loads; ++q)
        b += "      const " + std::string(T) + "* L" + std::to_string(q) + " = (const " + T + "*)A.plane[pi][" +
             std::to_string(1 + q) + "];\n";
    const bool upfront = P.n_loads <= 6;
    const std::string slot = upfront ? "r" : "0";
    auto load = [&](int q, const std::string& ind) {
        const std::string Q = std::to_string(q);
        return v16 ? ind + "aeth::units::vload<NT>(L" + Q + " + i0, x" + Q + "[" + slot + "]);\n"
                   : ind + "for (int k = 0; k < VN; ++k) x" + Q + "[" + slot + "][k] = L" + Q + "[i0 + k];\n";
    };
    b += "      if (full) {\n";
    for (int q = 0; q < P.n_loads; ++q)
        b += "        T x" + std::to_string(q) + (upfront ? "[ROUNDS][VN];\n" : "[1][VN];\n");
    if (upfront) {
        b += "#pragma unroll\n        for (int r = 0; r < ROUNDS; ++r) {\n"
             "          const uint32_t i0 = s0 + r * (256 * VN) + threadIdx.x * VN;\n";
        for (int q = 0; q < P.n_loads; ++q) b += load(q, "          ");
        b += "        }\n";
    }
    b += "#pragma unroll\n        for (int r = 0; r < ROUNDS; ++r) {\n"
         "          const uint32_t i0 = s0 + r * (256 * VN) + threadIdx.x * VN;\n";
    if (!upfront)
        for (int q = 0; q < P.n_loads; ++q) b += load(q, "          ");
    b += "          T v[VN];\n#pragma unroll\n          for (int k = 0; k < VN; ++k) {\n";
    const std::string res = gen_body(
        P, dbl, [&](int q) { return "x" + std::to_string(q) + "[" + slot + "][k]"; }, "i0 + k", "            ", b);
    b += "            v[k] = " + res + ";\n            if (SUM) aeth::acc::accumulate<M, false>(acc, v[k]);\n"
         "          }\n";
    b += v16 ? "          aeth::units::vstore<NT>(outp + i0, v);\n"
             : "          for (int k = 0; k < VN; ++k) outp[i0 + k] = v[k];\n";
    b += "        }\n      } else {\n";
    b += "#pragma unroll\n        for (int r = 0; r < ROUNDS; ++r)\n#pragma unroll\n"
         "          for (int k = 0; k < VN; ++k) {\n"
         "            const uint32_t i = s0 + r * (256 * VN) + threadIdx.x * VN + k;\n"
         "            if (i < n) {\n";
    for (int q = 0; q < P.n_loads; ++q)
        b += "              const T y" + std::to_string(q) + " = L" + std::to_string(q) + "[i];\n";
    const std::string res2 =
        gen_body(P, dbl, [&](int q) { return "y" + std::to_string(q); }, "i", "              ", b);
    b += "              outp[i] = " + res2 + ";\n              if (SUM) aeth::acc::accumulate<M, false>(acc, " + res2 +
         ");\n            }\n          }\n      }\n";
    return b;
}

// The generated kernel's symbol: aeth_tape_ and a 64-bit FNV-1a hash of the tape's structure and shape,
// the same in every process and on every device, so profiles name each structure apart.
std::string kernel_name(int dtype, int mode, int np, const aeth_program* P, bool nt, JitShape sh)
{
    const std::string key = tape_key(dtype, mode, np, P, nt, 0, sh);
    uint64_t h = 0xcbf29ce484222325ull;
    for (unsigned char ch : key) {
        h ^= ch;
        h *= 0x100000001b3ull;
    }
    char buf[40];
    std::snprintf(buf, sizeof buf, "aeth_tape_%016llx", (unsigned long long)h);
    return buf;
}

std::string gen_kernel(int dtype, int mode, int np, const aeth_program* P, bool nt, JitShape sh)
{
    const bool dbl = dtype == AETH_F64;
    const int M = mode == AETH_SUM_NONE ? 0 : mode;
    std::string s = kJitPrelude;
    s += R"K(
struct JitArgs {
    const void* plane[AETH_TAPE_MAX_PROGRAMS][1 + AETH_TAPE_MAX_LOADS];
    double consts[AETH_TAPE_MAX_CONSTS];
    aeth_rng_leaf rng[AETH_TAPE_MAX_RNGS];
    uint8_t rng_arg[AETH_TAPE_MAX_PROGRAMS][AETH_TAPE_MAX_INSNS];
    uint32_t n, tiles, units;
};
template<class T>
__device__ __forceinline__ T jit_draw(const aeth_rng_leaf& r, uint32_t i)
{
    const aeth::rng::U4 b = aeth::rng::block(r.seed, i, r.comp);
    if constexpr (sizeof(T) == 8) {
        if (r.dist == AETH_DIST_UNIFORM) return r.p0 + (r.p1 - r.p0) * aeth::rng::u01_from_word_d(b.x);
        return r.p0 + r.p1 * aeth::rng::normal_from_block_d(b);
    } else {
        const float p0 = (float)r.p0, p1 = (float)r.p1;
        if (r.dist == AETH_DIST_UNIFORM) return p0 + (p1 - p0) * aeth::rng::u01_from_word_f(b.x);
        return p0 + p1 * aeth::rng::normal_from_block_f(b);
    }
}
)K";
    s += std::string("extern \"C\" __global__ __launch_bounds__(256") +
         (sh.waves ? ", " + std::to_string(sh.waves) : std::string()) +
         ") void " + kernel_name(dtype, mode, np, P, nt, sh) + "(const JitArgs A, void* part_, int vec)\n{\n";
    s += std::string("  typedef ") + (dbl ? "double" : "float") + " T;\n";
    // nontemporal is compiled in: a run-time choice between two otherwise identical loads is merged
    // by the optimiser into one plain load
    s += std::string("  constexpr bool NT = ") + (nt ? "true" : "false") + ";\n";
    s += "  constexpr int M = " + std::to_string(M) + ";\n";
    s += std::string("  constexpr bool SUM = ") + (mode == AETH_SUM_NONE ? "false" : "true") + ";\n";
    s += "  constexpr uint32_t TILE = " + std::to_string(sh.tile) + ";\n";
    s += "  constexpr int VN = " + std::to_string(sh.vb) + " / sizeof(T);\n  constexpr int ROUNDS = TILE / (256 * VN);\n";
    s += "  __shared__ aeth::acc::Pair<T> lds[4];\n";
    s += "  aeth::acc::Pair<T>* part = (aeth::acc::Pair<T>*)part_;\n  const uint32_t n = A.n;\n";
    s += "  constexpr uint32_t XK = aeth::units::xcd_chunk(TILE * sizeof(T));\n";
    s += "  for (uint32_t v = blockIdx.x; v < A.units; v += gridDim.x) {\n";
    s += "    const uint32_t u = aeth::units::xcd_unit(v, A.units, XK);\n";
    s += "    const uint32_t pi = u / A.tiles;\n    const uint32_t tile = u - pi * A.tiles;\n";
    s += "    const uint32_t s0 = tile * TILE;\n    const bool full = vec && aeth::units::full_tile(s0, TILE, n);\n";
    s += "    aeth::acc::Pair<T> acc{T(0), T(0)};\n    switch (pi) {\n";
    std::vector<bool> done(np, false);
    for (int p = 0; p < np; ++p) {
        if (done[p]) continue;
        const std::string key = structure_key(P[p]);
        for (int q = p; q < np; ++q)
            if (!done[q] && structure_key(P[q]) == key) {
                s += "    case " + std::to_string(q) + ":\n";
                done[q] = true;
            }
        s += "    {\n" + gen_program(P[p], dbl, sh.vb) + "      break;\n    }\n";
    }
    s += "    default: break;\n    }\n";
    s += "    if (SUM) {\n      const aeth::acc::Pair<T> rr = aeth::acc::block_combine<M>(acc, lds);\n"
         "      if (threadIdx.x == 0) part[u] = rr;\n    }\n  }\n}\n";
    return s;
}

// Compile a tape's kernel for `arch_opt` ("--offload-arch=gfx950"); false (log in last error) on failure.
bool jit_compile(const std::string& src, const std::string& arch_opt, std::vector<char>& code)
{
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "aeth_tape_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        set_error("hiprtcCreateProgram failed");
        return false;
    }
    const char* opts[] = {arch_opt.c_str(), "-O3", "-ffp-contract=off", "-std=c++20", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, 8, opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        set_error("hiprtc: " + log.substr(0, 600));
        return false;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return true;
}

// Returns nullptr (and records why) when the runtime compiler is unavailable.
// Compiled tape kernels persist across processes when AETHER_KERNEL_CACHE names a directory: one code object
// per (architecture, generated source), named by two independent 64-bit hashes of both (FNV-1a and a
// multiply-rotate hash), so an accidental collision of names is out of reach. The source embeds the device
// headers, so a library whose device code changed looks for other files instead of loading stale ones.
// Files are written whole under a temporary name and renamed into place, so a concurrent reader never sees
// a partial one; a file that does not load is compiled again and replaced.
std::string cache_file(const std::string& arch, const std::string& src)
{
    static const char* dir = std::getenv("AETHER_KERNEL_CACHE");
    if (!dir || !*dir) return std::string();
    uint64_t h = 0xcbf29ce484222325ull, g = 0x9e3779b97f4a7c15ull;
    for (const std::string* part : {&arch, &src})
        for (unsigned char ch : *part) {
            h ^= ch;
            h *= 0x100000001b3ull;
            g = (g ^ ch) * 0xff51afd7ed558ccdull;
            g = (g << 29) | (g >> 35);
        }
    char buf[40];
    std::snprintf(buf, sizeof buf, "%016llx%016llx", (unsigned long long)h, (unsigned long long)g);
    std::string tag = arch.substr(arch.find('=') + 1);
    return std::string(dir) + "/aeth-" + tag + "-" + buf + ".co";
}

bool read_file(const std::string& path, std::vector<char>& out)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::vector<char> buf;
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    const bool ok = !std::ferror(f) && !buf.empty();
    std::fclose(f);
    if (ok) out.swap(buf);
    return ok;
}

void write_file(const std::string& path, const std::vector<char>& code)
{
    const std::string tmp = path + ".tmp" + std::to_string((unsigned long long)getpid());
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return; // an unwritable cache only costs the next process a compile
    const bool ok = std::fwrite(code.data(), 1, code.size(), f) == code.size();
    if (std::fclose(f) == 0 && ok && std::rename(tmp.c_str(), path.c_str()) == 0) return;
    std::remove(tmp.c_str());
}

hipFunction_t jit_function(int dtype, int mode, int np, const aeth_program* P, bool nt, JitShape sh)
{
    static const bool disabled = std::getenv("AETHER_TAPE_INTERPRET") != nullptr;
    if (disabled) return nullptr;
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    const std::string key = tape_key(dtype, mode, np, P, nt, dev, sh);
    std::lock_guard<std::mutex> lock(g_jit_mu);
    auto it = g_jit.find(key);
    if (it != g_jit.end()) {
        ++g_jit_hits;
        return it->second.fn;
    }
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
    const std::string name = kernel_name(dtype, mode, np, P, nt, sh);
    JitEntry e;
    std::vector<char> code;
    const std::string src = gen_kernel(dtype, mode, np, P, nt, sh);
    const std::string file = cache_file(arch, src);
    if (!file.empty() && read_file(file, code) && hipModuleLoadData(&e.module, code.data()) == hipSuccess) {
        if (hipModuleGetFunction(&e.fn, e.module, name.c_str()) == hipSuccess) {
            g_jit.emplace(key, e);
            ++g_jit_hits;
            return e.fn;
        }
        (void)hipModuleUnload(e.module);
        e = JitEntry{};
    }
    (void)hipGetLastError();
    if (!jit_compile(src, arch, code)) {
        ++g_jit_fallbacks;
        return nullptr;
    }
    check(hipModuleLoadData(&e.module, code.data()), "hipModuleLoadData");
    check(hipModuleGetFunction(&e.fn, e.module, name.c_str()), "hipModuleGetFunction");
    if (!file.empty()) write_file(file, code);
    g_jit.emplace(key, e);
    ++g_jit_compiled;
    return e.fn;
}
} // namespace
} // namespace aeth

using namespace aeth;

namespace {

// Set while aeth_tape_prepare walks an assignment: every kernel the assignment would launch is compiled and
// loaded into this process's cache; nothing is allocated, written or launched.
thread_local bool t_prepare_only = false;

struct PrepareOnly {
    PrepareOnly() { t_prepare_only = true; }
    ~PrepareOnly() { t_prepare_only = false; }
};

// One tape assignment. `force` (the tuner) runs every launch with that shape; otherwise assignments
// without sums use the tuned shape of their structure and assignments with sums the contract shape.
// Returns whether the assignment could run as one launch over all programs (no program reads a plane
// another one writes).
bool eval_tape(int dtype, int n_programs, const aeth_program* programs, int n_leaves, void* const* leaves,
               const uint64_t* leaf_pitch, int n_consts, const double* consts, int n_rngs, const aeth_rng_leaf* rngs,
               uint64_t n, int sum_mode, void* sums, void* stream, const JitShape* force)
{
    {
        if (dtype != AETH_F64 && dtype != AETH_F32) fail("aeth_eval_tape", "unknown dtype");
        if (n_programs < 1 || n_programs > AETH_TAPE_MAX_PROGRAMS) fail("aeth_eval_tape", "1..8 programs");
        if (n_leaves < 1 || n_leaves > AETH_TAPE_MAX_LEAVES) fail("aeth_eval_tape", "1..16 leaves");
        if (n_consts < 0 || n_consts > AETH_TAPE_MAX_CONSTS) fail("aeth_eval_tape", "0..16 constants");
        if (n_rngs < 0 || n_rngs > AETH_TAPE_MAX_RNGS) fail("aeth_eval_tape", "0..8 rng leaves");
        if (n > 0xFFFFFFFFull) fail("aeth_eval_tape", "n beyond the 32-bit offset_t");
        if (sum_mode != AETH_SUM_NONE && !sums && !t_prepare_only) fail("aeth_eval_tape", "sums is NULL");
        const size_t es = dtype == AETH_F64 ? 8 : 4;
        for (int p = 0; p < n_programs; ++p) validate(programs[p], n_leaves, n_consts, n_rngs);
        bool vec = true;
        for (int l = 0; l < n_leaves; ++l) {
            if (leaf_pitch[l] < n) fail("aeth_eval_tape", "leaf pitch < n");
            vec = vec && aligned16(leaves[l]) && (leaf_pitch[l] * es) % 16 == 0;
        }
        if (sum_mode == AETH_SUM_FLOATFLOAT && dtype == AETH_F64)
            fail("aeth_eval_tape", "float-float accumulation needs dtype F32", 0, 2);
        // A program that reads a plane another program writes (a = a.cross(b)): aether stores component
        // 0 of a sample before it evaluates component 1 (expr/Assign.h:124-160), so such programs run as
        // one launch each, in component order -- the same values, sample by sample.
        auto plane = [&](int leaf, int comp) {
            return static_cast<const char*>(leaves[leaf]) + (uint64_t)comp * leaf_pitch[leaf] * es;
        };
        for (int p = 0; p < n_programs; ++p)
            for (int q = p + 1; q < n_programs; ++q)
                if (plane(programs[p].out_leaf, programs[p].out_comp) == plane(programs[q].out_leaf, programs[q].out_comp))
                    fail("aeth_eval_tape", "two programs write the same plane");
        bool serial = false;
        for (int p = 0; p < n_programs && !serial; ++p)
            for (int q = 0; q < n_programs && !serial; ++q) {
                if (q == p) continue;
                const char* out = plane(programs[q].out_leaf, programs[q].out_comp);
                for (int k = 0; k < programs[p].n_loads; ++k)
                    if (plane(programs[p].load_leaf[k], programs[p].load_comp[k]) == out) serial = true;
            }
        hipStream_t st = as_stream(stream);
        auto run = [&](const aeth_program* P, int np, void* out_sums) {
            TapeArgs A;
            std::memset(&A, 0, sizeof A);
            int max_loads = 0;
            uint64_t span_bytes = 0;
            for (int p = 0; p < np; ++p) {
                A.prog[p] = P[p];
                max_loads = std::max<int>(max_loads, P[p].n_loads);
                span_bytes += (uint64_t)(P[p].n_loads + 1) * n * es;
            }
            for (int l = 0; l < n_leaves; ++l) {
                A.leaf[l] = leaves[l];
                A.pitch[l] = leaf_pitch[l];
            }
            for (int c = 0; c < n_consts; ++c) A.consts[c] = consts[c];
            for (int r = 0; r < n_rngs; ++r) A.rng[r] = rngs[r];
            A.n = (uint32_t)n;
            const bool nt = span_bytes >= (512ull << 20);
            JitShape sh{};
            if (force) {
                sh = *force;
            } else {
                // without sums any tuned shape; with sums only occupancy targets were tried (kSumShapes)
                int dev = 0;
                check(hipGetDevice(&dev), "hipGetDevice");
                std::lock_guard<std::mutex> lock(g_jit_mu);
                const auto it = g_tuned.find(tape_key(dtype, sum_mode, np, P, nt, dev));
                if (it != g_tuned.end()) sh = it->second;
            }
            hipFunction_t jit = jit_function(dtype, sum_mode, np, P, nt, sh);
            if (!jit) {
                if (force) fail("aeth_tape_tune", "the runtime compiler is unavailable");
                sh = JitShape{}; // the resident interpreter runs the contract shape
            }
            if (t_prepare_only) return;
            A.tiles = n ? (uint32_t)((n + sh.tile - 1) / sh.tile) : 1u;
            A.units = A.tiles * (uint32_t)np;
            if (dtype == AETH_F64)
                dispatch<double>(A, max_loads, np, vec, nt, sum_mode, static_cast<double*>(out_sums), st, jit);
            else
                dispatch<float>(A, max_loads, np, vec, nt, sum_mode, static_cast<float*>(out_sums), st, jit);
        };
        if (!serial) {
            run(programs, n_programs, sums);
        } else {
            for (int p = 0; p < n_programs; ++p)
                run(&programs[p], 1, sums ? static_cast<char*>(sums) + 2 * es * (size_t)p : nullptr);
        }
        return !serial;
    }
}

// ---------------------------------------------------------------------------
// Trees of any size (aeth_eval_tape_ex). A program that fits one launch's limits runs as before. A larger
// one is cut into pieces that fit: the largest subtrees are evaluated first into temporary planes of T, and
// the pieces above them load those planes. Every node still rounds to T exactly once, so the values are
// the same as one straight-line evaluation; pieces run in dependency order on the caller's stream, and the
// programs of an assignment that has to be cut run one after another in component order (the order in
// which the reference stores a sample's components, expr/Assign.h:124-160).
struct TreeNode {
    int op;
    uint32_t arg;
    int a = -1, b = -1; // children (unary: a)
};

struct PieceCost {
    int insns = 0, depth = 0;
    std::vector<uint64_t> loads; // (leaf << 32 | comp), or temps as (1 << 63 | temp)
    std::vector<uint32_t> consts, rngs;
};

void add_unique(std::vector<uint64_t>& v, uint64_t x)
{
    if (std::find(v.begin(), v.end(), x) == v.end()) v.push_back(x);
}
void add_unique(std::vector<uint32_t>& v, uint32_t x)
{
    if (std::find(v.begin(), v.end(), x) == v.end()) v.push_back(x);
}

struct Splitter {
    const aeth_program_ex& P;
    std::vector<TreeNode> nodes;
    std::vector<int> temp_of; // node -> temp index, or -1
    int root = -1;
    int temps = 0;

    explicit Splitter(const aeth_program_ex& p) : P(p)
    {
        std::vector<int> st;
        for (uint32_t i = 0; i < P.n_insns; ++i) {
            TreeNode t{P.insns[i].op, P.insns[i].arg};
            if (is_binary(t.op)) {
                t.b = st.back();
                st.pop_back();
                t.a = st.back();
                st.pop_back();
            } else if (is_unary(t.op)) {
                t.a = st.back();
                st.pop_back();
            }
            nodes.push_back(t);
            st.push_back((int)nodes.size() - 1);
        }
        root = st.back();
        temp_of.assign(nodes.size(), -1);
    }

    bool leafish(int k) const { return nodes[k].a < 0 || temp_of[k] >= 0; }

    // The cost of evaluating node k's piece: its subtree down to leaves and materialized nodes.
    void cost(int k, PieceCost& c, bool top) const
    {
        const TreeNode& t = nodes[k];
        if (!top && temp_of[k] >= 0) {
            add_unique(c.loads, (1ull << 63) | (uint64_t)temp_of[k]);
            c.insns += 1;
            c.depth = 1;
            return;
        }
        if (t.op == AETH_OP_PUSH) {
            add_unique(c.loads, ((uint64_t)P.load_leaf[t.arg] << 32) | P.load_comp[t.arg]);
            c.insns += 1;
            c.depth = 1;
        } else if (t.op == AETH_OP_CONST) {
            add_unique(c.consts, t.arg);
            c.insns += 1;
            c.depth = 1;
        } else if (t.op == AETH_OP_RNG) {
            add_unique(c.rngs, t.arg);
            c.insns += 1;
            c.depth = 1;
        } else if (t.b < 0) {
            cost(t.a, c, false);
            c.insns += 1;
        } else {
            cost(t.a, c, false);
            const int da = c.depth;
            cost(t.b, c, false);
            c.depth = std::max(da, c.depth + 1);
            c.insns += 1;
        }
    }

    static bool fits(const PieceCost& c)
    {
        // loads leave one leaf slot for the piece's output plane
        return c.insns <= AETH_TAPE_MAX_INSNS && c.depth <= AETH_TAPE_MAX_DEPTH &&
               (int)c.loads.size() < AETH_TAPE_MAX_LEAVES && (int)c.loads.size() <= AETH_TAPE_MAX_LOADS &&
               (int)c.consts.size() <= AETH_TAPE_MAX_CONSTS && (int)c.rngs.size() <= AETH_TAPE_MAX_RNGS;
    }

    int size(int k) const
    {
        if (temp_of[k] >= 0 || nodes[k].a < 0) return 1;
        return 1 + size(nodes[k].a) + (nodes[k].b >= 0 ? size(nodes[k].b) : 0);
    }

    // Make node k's piece fit by materializing its largest non-trivial children, deepest first.
    void fit(int k)
    {
        for (;;) {
            PieceCost c;
            cost(k, c, true);
            if (fits(c)) return;
            const TreeNode& t = nodes[k];
            int pick = -1;
            for (int ch : {t.a, t.b})
                if (ch >= 0 && !leafish(ch) && (pick < 0 || size(ch) > size(pick))) pick = ch;
            if (pick < 0) fail("aeth_eval_tape_ex", "a node cannot be made to fit one program"); // unreachable
            fit(pick);
            temp_of[pick] = temps++;
        }
    }
};

// A tape for node k's piece, its planes in `loads` (leaf pointer, pitch, comp) and constants / random
// leaves remapped into the piece's own tables.
struct Piece {
    std::vector<aeth_insn> insns;
    std::vector<std::pair<const void*, uint64_t>> leaves; // pointer, pitch
    std::vector<std::pair<int, uint32_t>> loads;          // (piece leaf, comp)
    std::vector<uint32_t> consts, rngs;                   // indices into the caller's tables
};

void emit_piece(const Splitter& S, int k, bool top, Piece& pc, void* const* leaves, const uint64_t* pitch,
                const std::vector<void*>& temp_ptr, uint64_t temp_pitch)
{
    const TreeNode& t = S.nodes[k];
    auto load = [&](const void* ptr, uint64_t pt, uint32_t comp) {
        int li = -1;
        for (size_t q = 0; q < pc.leaves.size(); ++q)
            if (pc.leaves[q].first == ptr) li = (int)q;
        if (li < 0) {
            pc.leaves.emplace_back(ptr, pt);
            li = (int)pc.leaves.size() - 1;
        }
        const std::pair<int, uint32_t> key{li, comp};
        size_t j = 0;
        while (j < pc.loads.size() && pc.loads[j] != key) ++j;
        if (j == pc.loads.size()) pc.loads.push_back(key);
        pc.insns.push_back(aeth_insn{AETH_OP_PUSH, (uint8_t)j});
    };
    auto index_of = [](std::vector<uint32_t>& v, uint32_t x) {
        size_t j = 0;
        while (j < v.size() && v[j] != x) ++j;
        if (j == v.size()) v.push_back(x);
        return (uint8_t)j;
    };
    if (!top && S.temp_of[k] >= 0) {
        load(temp_ptr[S.temp_of[k]], temp_pitch, 0);
        return;
    }
    if (t.op == AETH_OP_PUSH) {
        const uint32_t leaf = S.P.load_leaf[t.arg];
        load(leaves[leaf], pitch[leaf], S.P.load_comp[t.arg]);
    } else if (t.op == AETH_OP_CONST) {
        pc.insns.push_back(aeth_insn{AETH_OP_CONST, index_of(pc.consts, t.arg)});
    } else if (t.op == AETH_OP_RNG) {
        pc.insns.push_back(aeth_insn{AETH_OP_RNG, index_of(pc.rngs, t.arg)});
    } else {
        emit_piece(S, t.a, false, pc, leaves, pitch, temp_ptr, temp_pitch);
        if (t.b >= 0) emit_piece(S, t.b, false, pc, leaves, pitch, temp_ptr, temp_pitch);
        pc.insns.push_back(aeth_insn{(uint8_t)t.op, 0});
    }
}

// Run one piece as a one-program assignment into plane (out_ptr, comp 0 of a leaf of pitch out_pitch).
void run_piece(int dtype, Piece& pc, void* out_ptr, uint64_t out_pitch, uint32_t out_comp, const double* consts,
               const aeth_rng_leaf* rngs, uint64_t n, int sum_mode, void* sums, hipStream_t st)
{
    int out_leaf = -1;
    for (size_t q = 0; q < pc.leaves.size(); ++q)
        if (pc.leaves[q].first == out_ptr) out_leaf = (int)q;
    if (out_leaf < 0) {
        pc.leaves.emplace_back(out_ptr, out_pitch);
        out_leaf = (int)pc.leaves.size() - 1;
    }
    aeth_program prog;
    std::memset(&prog, 0, sizeof prog);
    prog.n_loads = (uint8_t)pc.loads.size();
    prog.n_insns = (uint8_t)pc.insns.size();
    prog.out_leaf = (uint8_t)out_leaf;
    prog.out_comp = (uint8_t)out_comp;
    for (size_t q = 0; q < pc.loads.size(); ++q) {
        prog.load_leaf[q] = (uint8_t)pc.loads[q].first;
        prog.load_comp[q] = (uint8_t)pc.loads[q].second;
    }
    for (size_t i = 0; i < pc.insns.size(); ++i) prog.insns[i] = pc.insns[i];
    std::vector<void*> lp;
    std::vector<uint64_t> lpitch;
    for (const auto& l : pc.leaves) {
        lp.push_back(const_cast<void*>(l.first));
        lpitch.push_back(l.second);
    }
    std::vector<double> cs;
    for (uint32_t c : pc.consts) cs.push_back(consts[c]);
    std::vector<aeth_rng_leaf> rs;
    for (uint32_t r : pc.rngs) rs.push_back(rngs[r]);
    eval_tape(dtype, 1, &prog, (int)lp.size(), lp.data(), lpitch.data(), (int)cs.size(), cs.data(), (int)rs.size(),
              rs.data(), n, sum_mode, sums, st, nullptr);
}

void validate_ex(const aeth_program_ex& p, int n_leaves, int n_consts, int n_rngs)
{
    if (p.n_insns == 0 || !p.insns) fail("aeth_eval_tape_ex", "a program needs instructions");
    if (p.n_loads && (!p.load_leaf || !p.load_comp)) fail("aeth_eval_tape_ex", "load tables are NULL");
    if ((int)p.out_leaf >= n_leaves) fail("aeth_eval_tape_ex", "output leaf out of range");
    for (uint32_t k = 0; k < p.n_loads; ++k)
        if ((int)p.load_leaf[k] >= n_leaves) fail("aeth_eval_tape_ex", "load leaf out of range");
    int64_t d = 0;
    for (uint32_t i = 0; i < p.n_insns; ++i) {
        const int op = p.insns[i].op;
        const uint32_t arg = p.insns[i].arg;
        if (op == AETH_OP_PUSH) {
            if (arg >= p.n_loads) fail("aeth_eval_tape_ex", "PUSH index out of range");
            ++d;
        } else if (op == AETH_OP_CONST) {
            if ((int64_t)arg >= n_consts) fail("aeth_eval_tape_ex", "CONST index out of range");
            ++d;
        } else if (op == AETH_OP_RNG) {
            if ((int64_t)arg >= n_rngs) fail("aeth_eval_tape_ex", "RNG index out of range");
            ++d;
        } else if (is_binary(op)) {
            if (d < 2) fail("aeth_eval_tape_ex", "stack underflow");
            --d;
        } else if (is_unary(op)) {
            if (d < 1) fail("aeth_eval_tape_ex", "stack underflow");
        } else {
            fail("aeth_eval_tape_ex", "unknown opcode");
        }
    }
    if (d != 1) fail("aeth_eval_tape_ex", "a program must leave exactly one value");
}

bool fits_one_launch(const aeth_program_ex& p)
{
    if (p.n_loads > AETH_TAPE_MAX_LOADS || p.n_insns > AETH_TAPE_MAX_INSNS) return false;
    int d = 0, dmax = 0;
    for (uint32_t i = 0; i < p.n_insns; ++i) {
        const int op = p.insns[i].op;
        if (p.insns[i].arg > 255) return false;
        if (op == AETH_OP_PUSH || op == AETH_OP_CONST || op == AETH_OP_RNG) ++d;
        else if (is_binary(op)) --d;
        dmax = std::max(dmax, d);
    }
    return dmax <= AETH_TAPE_MAX_DEPTH;
}

void eval_tape_ex(int dtype, int n_programs, const aeth_program_ex* programs, int n_leaves, void* const* leaves,
                  const uint64_t* leaf_pitch, int n_consts, const double* consts, int n_rngs, const aeth_rng_leaf* rngs,
                  uint64_t n, int sum_mode, void* sums, void* stream)
{
    if (dtype != AETH_F64 && dtype != AETH_F32) fail("aeth_eval_tape_ex", "unknown dtype");
    if (n_programs < 1 || n_programs > 64) fail("aeth_eval_tape_ex", "1..64 programs");
    if (n_leaves < 1 || !leaves || !leaf_pitch) fail("aeth_eval_tape_ex", "no leaves");
    if (n_consts < 0 || n_rngs < 0) fail("aeth_eval_tape_ex", "negative table size");
    if (n > 0xFFFFFFFFull) fail("aeth_eval_tape_ex", "n beyond the 32-bit offset_t");
    if (sum_mode != AETH_SUM_NONE && !sums && !t_prepare_only) fail("aeth_eval_tape_ex", "sums is NULL");
    for (int p = 0; p < n_programs; ++p) validate_ex(programs[p], n_leaves, n_consts, n_rngs);
    bool one = n_programs <= AETH_TAPE_MAX_PROGRAMS && n_leaves <= AETH_TAPE_MAX_LEAVES &&
               n_consts <= AETH_TAPE_MAX_CONSTS && n_rngs <= AETH_TAPE_MAX_RNGS;
    for (int p = 0; p < n_programs && one; ++p) one = fits_one_launch(programs[p]) && programs[p].out_leaf < 256 &&
                                                      programs[p].out_comp < 256;
    if (one) { // the plain tape: one launch over all programs
        std::vector<aeth_program> ps(n_programs);
        for (int p = 0; p < n_programs; ++p) {
            const aeth_program_ex& e = programs[p];
            aeth_program& q = ps[p];
            std::memset(&q, 0, sizeof q);
            q.n_loads = (uint8_t)e.n_loads;
            q.n_insns = (uint8_t)e.n_insns;
            q.out_leaf = (uint8_t)e.out_leaf;
            q.out_comp = (uint8_t)e.out_comp;
            for (uint32_t k = 0; k < e.n_loads; ++k) {
                q.load_leaf[k] = (uint8_t)e.load_leaf[k];
                q.load_comp[k] = (uint8_t)e.load_comp[k];
            }
            for (uint32_t i = 0; i < e.n_insns; ++i) q.insns[i] = aeth_insn{e.insns[i].op, (uint8_t)e.insns[i].arg};
        }
        eval_tape(dtype, n_programs, ps.data(), n_leaves, leaves, leaf_pitch, n_consts, consts, n_rngs, rngs, n,
                  sum_mode, sums, stream, nullptr);
        return;
    }
    // cut into pieces; the programs run in component order
    const size_t es = dtype == AETH_F64 ? 8 : 4;
    const uint64_t temp_pitch = (n + 31) / 32 * 32;
    hipStream_t st = as_stream(stream);
    for (int p = 0; p < n_programs; ++p) {
        const aeth_program_ex& P = programs[p];
        Splitter S(P);
        S.fit(S.root);
        std::vector<std::unique_ptr<StreamScratch>> temps;
        std::vector<void*> temp_ptr(S.temps, nullptr);
        for (int t = 0; t < S.temps; ++t) {
            if (t_prepare_only) { // no planes: a piece's kernel depends on the tape's structure, not its pointers
                temp_ptr[t] = reinterpret_cast<void*>((uintptr_t)256 * (t + 1));
                continue;
            }
            temps.emplace_back(new StreamScratch(std::max<uint64_t>(temp_pitch, 32) * es, st));
            temp_ptr[t] = temps.back()->get();
        }
        // materialized nodes in the order they were cut (children before parents)
        std::vector<int> order(S.temps, -1);
        for (size_t k = 0; k < S.nodes.size(); ++k)
            if (S.temp_of[k] >= 0) order[S.temp_of[k]] = (int)k;
        for (int t = 0; t < S.temps; ++t) {
            Piece pc;
            emit_piece(S, order[t], true, pc, leaves, leaf_pitch, temp_ptr, temp_pitch);
            run_piece(dtype, pc, temp_ptr[t], temp_pitch, 0, consts, rngs, n, AETH_SUM_NONE, nullptr, st);
        }
        Piece top;
        emit_piece(S, S.root, true, top, leaves, leaf_pitch, temp_ptr, temp_pitch);
        void* out = static_cast<char*>(leaves[P.out_leaf]);
        void* psums = sums ? static_cast<char*>(sums) + 2 * es * (size_t)p : nullptr;
        // the output plane goes in as comp P.out_comp of its leaf
        run_piece(dtype, top, out, leaf_pitch[P.out_leaf], P.out_comp, consts, rngs, n, sum_mode, psums, st);
    }
}

} // namespace

extern "C" int aeth_eval_tape(int dtype, int n_programs, const aeth_program* programs, int n_leaves,
                              void* const* leaves, const uint64_t* leaf_pitch, int n_consts, const double* consts,
                              int n_rngs, const aeth_rng_leaf* rngs, uint64_t n, int sum_mode, void* sums,
                              void* stream)
{
    return guarded([&] {
        eval_tape(dtype, n_programs, programs, n_leaves, leaves, leaf_pitch, n_consts, consts, n_rngs, rngs, n,
                  sum_mode, sums, stream, nullptr);
    });
}

extern "C" int aeth_eval_tape_ex(int dtype, int n_programs, const aeth_program_ex* programs, int n_leaves,
                                 void* const* leaves, const uint64_t* leaf_pitch, int n_consts, const double* consts,
                                 int n_rngs, const aeth_rng_leaf* rngs, uint64_t n, int sum_mode, void* sums,
                                 void* stream)
{
    return guarded([&] {
        eval_tape_ex(dtype, n_programs, programs, n_leaves, leaves, leaf_pitch, n_consts, consts, n_rngs, rngs, n,
                     sum_mode, sums, stream);
    });
}

extern "C" int aeth_tape_timing(int enable)
{
    return guarded([&] { g_timing.store(enable != 0); });
}

extern "C" int aeth_tape_timing_read(float* ms, int capacity, int* count)
{
    return guarded([&] {
        if (!count || (capacity > 0 && !ms)) fail("aeth_tape_timing_read", "NULL output");
        std::lock_guard<std::mutex> lock(g_timing_mu);
        const size_t take = std::min<size_t>(g_timed.size(), (size_t)std::max(capacity, 0));
        *count = (int)g_timed.size();
        size_t k = 0;
        // the pairs read so far, and one that fails to read, leave the list
        struct Drop {
            size_t& k;
            ~Drop()
            {
                for (size_t j = 0; j < k; ++j) {
                    g_event_pool.push_back(g_timed[j].first);
                    g_event_pool.push_back(g_timed[j].second);
                }
                g_timed.erase(g_timed.begin(), g_timed.begin() + (std::ptrdiff_t)k);
            }
        } drop{k};
        while (k < take) {
            const auto pr = g_timed[k++];
            float t = 0.0f;
            check(hipEventSynchronize(pr.second), "hipEventSynchronize");
            check(hipEventElapsedTime(&t, pr.first, pr.second), "hipEventElapsedTime");
            ms[k - 1] = t;
        }
    });
}

extern "C" int aeth_tape_prepare(int dtype, int n_programs, const aeth_program_ex* programs, int n_leaves,
                                 void* const* leaves, const uint64_t* leaf_pitch, int n_consts, const double* consts,
                                 int n_rngs, const aeth_rng_leaf* rngs, uint64_t n, int sum_mode, void* stream)
{
    return guarded([&] {
        PrepareOnly scope;
        eval_tape_ex(dtype, n_programs, programs, n_leaves, leaves, leaf_pitch, n_consts, consts, n_rngs, rngs, n,
                     sum_mode, nullptr, stream);
    });
}

namespace {

// Time every candidate shape of `dest = tape` (with `sum_mode`) on the caller's arrays, keep the fastest for
// the tape's structure on this device. Every candidate writes the same values.
void tune_tape(const char* who, int dtype, int n_programs, const aeth_program* programs, int n_leaves,
               void* const* leaves, const uint64_t* leaf_pitch, int n_consts, const double* consts, int n_rngs,
               const aeth_rng_leaf* rngs, uint64_t n, int sum_mode, void* sums, void* stream,
               const JitShape* cands, size_t ncands, aeth_tape_shape* chosen)
{
    hipStream_t st = as_stream(stream);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        fail(who, "tuning is not allowed inside a stream capture");
    // every candidate writes the same values, so the destination ends holding the assignment's result; an
    // assignment that reads any plane it writes (a = 2a + b) would not, and is refused
    if (n_programs < 1 || n_programs > AETH_TAPE_MAX_PROGRAMS) fail(who, "1..8 programs");
    if (n_leaves < 1 || n_leaves > AETH_TAPE_MAX_LEAVES) fail(who, "1..16 leaves");
    if (sum_mode != AETH_SUM_NONE && !sums) fail(who, "sums is NULL");
    const size_t es = dtype == AETH_F64 ? 8 : 4;
    for (int p = 0; p < n_programs; ++p) {
        validate(programs[p], n_leaves, n_consts, n_rngs);
        for (int q = 0; q < n_programs; ++q) {
            const auto& Q = programs[q];
            const char* out = static_cast<const char*>(leaves[Q.out_leaf]) + (uint64_t)Q.out_comp * leaf_pitch[Q.out_leaf] * es;
            for (int k = 0; k < programs[p].n_loads; ++k) {
                const auto& P = programs[p];
                const char* in = static_cast<const char*>(leaves[P.load_leaf[k]]) +
                                 (uint64_t)P.load_comp[k] * leaf_pitch[P.load_leaf[k]] * es;
                if (in == out) fail(who, "the destination is also read: tune an assignment into another array");
            }
        }
    }
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    JitShape best = cands[0];
    float best_ms = -1.0f;
    try {
        for (size_t i = 0; i < ncands; ++i) {
            const JitShape& sh = cands[i];
            auto once = [&] {
                eval_tape(dtype, n_programs, programs, n_leaves, leaves, leaf_pitch, n_consts, consts, n_rngs, rngs,
                          n, sum_mode, sums, stream, &sh);
            };
            once(); // compiles and warms the variant
            constexpr int kReps = 3;
            check(hipEventRecord(e0, st), "hipEventRecord");
            for (int r = 0; r < kReps; ++r) once();
            check(hipEventRecord(e1, st), "hipEventRecord");
            check(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0.0f;
            check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            ms /= kReps;
            if (best_ms < 0.0f || ms < best_ms) {
                best_ms = ms;
                best = sh;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    uint64_t span_bytes = 0;
    for (int p = 0; p < n_programs; ++p) span_bytes += (uint64_t)(programs[p].n_loads + 1) * n * es;
    const bool nt = span_bytes >= (512ull << 20);
    {
        std::lock_guard<std::mutex> lock(g_jit_mu);
        g_tuned[tape_key(dtype, sum_mode, n_programs, programs, nt, dev)] = best;
    }
    if (chosen) {
        chosen->tile = best.tile;
        chosen->vector_bytes = best.vb;
        chosen->waves_per_simd = best.waves;
        chosen->ms = best_ms;
    }
}

} // namespace

extern "C" int aeth_tape_tune(int dtype, int n_programs, const aeth_program* programs, int n_leaves,
                              void* const* leaves, const uint64_t* leaf_pitch, int n_consts, const double* consts,
                              int n_rngs, const aeth_rng_leaf* rngs, uint64_t n, void* stream, aeth_tape_shape* chosen)
{
    return guarded([&] {
        tune_tape("aeth_tape_tune", dtype, n_programs, programs, n_leaves, leaves, leaf_pitch, n_consts, consts,
                  n_rngs, rngs, n, AETH_SUM_NONE, nullptr, stream, kShapes, std::size(kShapes), chosen);
    });
}

extern "C" int aeth_tape_tune_sums(int dtype, int n_programs, const aeth_program* programs, int n_leaves,
                                   void* const* leaves, const uint64_t* leaf_pitch, int n_consts,
                                   const double* consts, int n_rngs, const aeth_rng_leaf* rngs, uint64_t n,
                                   int sum_mode, void* sums, void* stream, aeth_tape_shape* chosen)
{
    return guarded([&] {
        if (sum_mode == AETH_SUM_NONE) fail("aeth_tape_tune_sums", "a sum mode is required (aeth_tape_tune tunes without)");
        tune_tape("aeth_tape_tune_sums", dtype, n_programs, programs, n_leaves, leaves, leaf_pitch, n_consts, consts,
                  n_rngs, rngs, n, sum_mode, sums, stream, kSumShapes, std::size(kSumShapes), chosen);
    });
}

extern "C" int aeth_tape_jit_stats(uint64_t* compiled, uint64_t* cache_hits, uint64_t* fallbacks)
{
    return guarded([&] {
        if (compiled) *compiled = g_jit_compiled.load();
        if (cache_hits) *cache_hits = g_jit_hits.load();
        if (fallbacks) *fallbacks = g_jit_fallbacks.load();
    });
}

extern "C" int aeth_tape_jit_build(int dtype, int sum_mode, int n_programs, const aeth_program* programs,
                                   const char* arch, uint64_t* code_bytes)
{
    return guarded([&] {
        if (n_programs < 1 || n_programs > AETH_TAPE_MAX_PROGRAMS) fail("aeth_tape_jit_build", "1..8 programs");
        for (int p = 0; p < n_programs; ++p)
            validate(programs[p], AETH_TAPE_MAX_LEAVES, AETH_TAPE_MAX_CONSTS, AETH_TAPE_MAX_RNGS);
        std::vector<char> code;
        if (!jit_compile(gen_kernel(dtype, sum_mode, n_programs, programs, true, JitShape{}),
                         std::string("--offload-arch=") + arch, code))
            throw Failure{1};
        if (code_bytes) *code_bytes = code.size();
    });
}
