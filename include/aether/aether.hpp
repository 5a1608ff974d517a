// aether/aether.hpp -- the C++ front end of libaether_hip.so, in aether's spelling.
//
// A host translation unit (g++ or clang, C++20, no HIP headers) keeps the
// reference's API for the hot path and gets one gfx950 kernel per assignment:
//
//   aether::Array<double, 3> a(n), b(n), c(n), d(n);
//   const aether::random::Generator g(42);
//   a.deviceView() = g.uniform<double, 3>(-1.0, 1.0);                  // Philox leaf, on the device
//   auto dv = d.deviceView();
//   dv = a.deviceView() + b.deviceView() * c.deviceView()
//        + aether::cwiseSin(a.deviceView()) * aether::cwiseExp(-1.0 * (b.deviceView() * b.deviceView()));
//   d.download();
//
// References: Array (array/Array.h:82-364), views and SampleIndex
// (view/View.h:150-424, index/SampleIndex.h:36-61), lazy assignment
// (expr/Assign.h:124-160,251-346), operators and nodes (expr/Operations.h:
// 25-71, expr/nodes/Arithmetic.h:33-143, expr/nodes/Elementwise.h:245-533,
// expr/nodes/Geometric.h:32-73, expr/Reduce.h:150-190), Generator and leaves
// (random/Generator.h:56-195, random/detail/DistributionLeaf.h:111-236,
// random/free.h:41-44), FFT (fft/Fft.h:58-236).
//
// Device assignment lowers the tree to the tape of aeth_eval_tape (one program
// per component; node semantics: -x is x*-1, x/s is x*(1/s) with the
// reciprocal rounded in T, folds right-nested). The per-sample host spelling
// `hv[SampleIndex::make(i)] = expr` evaluates on the host with std:: math,
// like the reference's host route. In a HIP translation unit the same
// spelling works inside the caller's own kernels over device views
// (`dv[SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x)] = expr`) and
// writes the bits the whole-array assignment writes.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <span>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../aether_hip.h"
#include "device/flag.hpp"

// AETHER_DEVICEHOST() (the reference's aether/macros.h): the expression nodes, views, SampleIndex, SampleRef
// and the distribution leaves compile for the host and, in a HIP translation unit, for the device as well, so
// a caller's own kernel can evaluate `view[SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x)] = expr` per
// sample. On the device the nodes call this library's device math and Philox (aether_amd/csrc/dmath.hpp,
// rng.hpp: the functions the runtime-compiled tape kernels call), so such a kernel writes the same bits as the
// whole-array assignment `dv = expr`; on the host they keep std:: math (the reference's host route).
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>

#include "../../aether_amd/csrc/rng.hpp"
#define AETHER_DEVICEHOST() __host__ __device__
#else
#define AETHER_DEVICEHOST()
#endif

namespace aether {

class Error : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

using Stream = void*; // a hipStream_t; nullptr = the null stream
using offset_t = std::uint32_t;
inline constexpr std::size_t kCapacityAlignment = 32;

namespace detail {

inline void check(int rc, const char* op)
{
    if (rc != 0) throw Error(std::string(op) + ": " + aeth_last_error());
}

inline const AetherMemoryV1& seam()
{
    static const AetherMemoryV1* t = aether_memory_v1();
    return *t;
}

inline int current_device()
{
    int d = 0;
    check(seam().current_device(&d), "current_device");
    return d;
}

template<class T>
constexpr int dtype_code()
{
    static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>, "aether: T must be float or double");
    return std::is_same_v<T, double> ? AETH_F64 : AETH_F32;
}

// random/detail/Backend.h:75-81
constexpr std::uint64_t splitmix64(std::uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Philox4x32-10 block at (seed, global, counter) (random/detail/Philox.h:95-133, Backend.h:144-150)
inline std::array<std::uint32_t, 4> philox_block(std::uint64_t seed, std::uint32_t global, std::uint64_t counter)
{
    std::uint32_t x = (std::uint32_t)counter, y = (std::uint32_t)(counter >> 32), z = global, w = 0;
    std::uint32_t k0 = (std::uint32_t)seed, k1 = (std::uint32_t)(seed >> 32);
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const std::uint64_t p0 = (std::uint64_t)0xD2511F53u * x, p1 = (std::uint64_t)0xCD9E8D57u * z;
        const std::uint32_t nx = (std::uint32_t)(p1 >> 32) ^ y ^ k0, nz = (std::uint32_t)(p0 >> 32) ^ w ^ k1;
        x = nx;
        y = (std::uint32_t)p1;
        z = nz;
        w = (std::uint32_t)p0;
    }
    return {x, y, z, w};
}

// ---- lowering context for aeth_eval_tape / aeth_eval_tape_ex
struct TapeCtx {
    int dtype = AETH_F64;
    std::vector<const void*> leaves;
    std::vector<std::uint64_t> pitch;
    std::vector<double> consts;
    std::vector<aeth_rng_leaf> rngs;
    std::vector<std::pair<std::uint32_t, std::uint32_t>> loads; // (leaf, comp) of the current program
    std::vector<aeth_insn_ex> insns;                            // the current program's tape

    struct Program {
        std::vector<std::uint32_t> load_leaf, load_comp;
        std::vector<aeth_insn_ex> insns;
        std::uint32_t out_leaf = 0, out_comp = 0;
    };
    std::vector<Program> progs;         // one per destination component, any size
    std::vector<aeth_program> programs; // the same as one launch's programs, when they fit (else empty)

    int leaf(const void* p, std::uint64_t pt)
    {
        for (std::size_t k = 0; k < leaves.size(); ++k)
            if (leaves[k] == p) return (int)k;
        leaves.push_back(p);
        pitch.push_back(pt);
        return (int)leaves.size() - 1;
    }
    void emit(int op, std::size_t arg) { insns.push_back(aeth_insn_ex{(std::uint8_t)op, {0, 0, 0}, (std::uint32_t)arg}); }
    void push(const void* p, std::uint64_t pt, std::size_t comp)
    {
        const std::pair<std::uint32_t, std::uint32_t> key{(std::uint32_t)leaf(p, pt), (std::uint32_t)comp};
        std::size_t j = 0;
        while (j < loads.size() && loads[j] != key) ++j;
        if (j == loads.size()) loads.push_back(key);
        emit(AETH_OP_PUSH, j);
    }
    void constant(double v)
    {
        std::size_t j = 0;
        while (j < consts.size() && std::memcmp(&consts[j], &v, sizeof v) != 0) ++j;
        if (j == consts.size()) consts.push_back(v);
        emit(AETH_OP_CONST, j);
    }
    void rng(std::uint64_t seed, double p0, double p1, std::uint32_t comp, std::uint32_t dist)
    {
        std::size_t j = 0;
        for (; j < rngs.size(); ++j) {
            const auto& r = rngs[j];
            if (r.seed == seed && r.p0 == p0 && r.p1 == p1 && r.comp == comp && r.dist == dist) break;
        }
        if (j == rngs.size()) rngs.push_back(aeth_rng_leaf{seed, p0, p1, comp, dist});
        emit(AETH_OP_RNG, j);
    }
    void op(int o) { emit(o, 0); }

    /** The programs in aeth_eval_tape_ex's form (pointing into progs: valid while this context lives). */
    std::vector<aeth_program_ex> ex() const
    {
        std::vector<aeth_program_ex> out(progs.size());
        for (std::size_t c = 0; c < progs.size(); ++c) {
            const Program& p = progs[c];
            out[c] = aeth_program_ex{(std::uint32_t)p.load_leaf.size(), (std::uint32_t)p.insns.size(), p.out_leaf,
                                     p.out_comp, p.load_leaf.data(), p.load_comp.data(), p.insns.data()};
        }
        return out;
    }
};

inline int stack_depth(const std::vector<aeth_insn_ex>& insns)
{
    int d = 0, peak = 0;
    for (const auto& i : insns) {
        if (i.op == AETH_OP_PUSH || i.op == AETH_OP_CONST || i.op == AETH_OP_RNG) ++d;
        else if ((i.op >= AETH_OP_ADD && i.op <= AETH_OP_DIV) || (i.op >= AETH_OP_FMIN && i.op <= AETH_OP_POW)) --d;
        peak = std::max(peak, d);
    }
    return peak;
}

/**
 * Lower `dest = e` (dest: N planes `pitch` elements apart at `dest`) to one
 * program per component, of any size. Leaves are numbered in order of first
 * use, the destination after the first program's operands; constants and
 * random leaves are shared across programs. When every program fits one
 * launch (aeth_eval_tape's limits) `programs` holds that form too.
 */
template<class T, std::size_t N, class E>
TapeCtx lower(const void* dest, std::size_t pitch, const E& e)
{
    static_assert(E::ncomp == N || E::ncomp == 1, "aether: expression and view component counts differ");
    static_assert(N <= 64, "aether: too many components for one assignment");
    TapeCtx x;
    x.dtype = dtype_code<T>();
    x.progs.resize(N);
    bool fits = N <= AETH_TAPE_MAX_PROGRAMS;
    for (std::size_t c = 0; c < N; ++c) {
        x.loads.clear();
        x.insns.clear();
        e.emit(E::ncomp > 1 ? c : 0, x);
        TapeCtx::Program& p = x.progs[c];
        p.out_leaf = (std::uint32_t)x.leaf(dest, pitch);
        p.out_comp = (std::uint32_t)c;
        for (const auto& l : x.loads) {
            p.load_leaf.push_back(l.first);
            p.load_comp.push_back(l.second);
        }
        p.insns = x.insns;
        fits = fits && x.loads.size() <= AETH_TAPE_MAX_LOADS && x.insns.size() <= AETH_TAPE_MAX_INSNS &&
               stack_depth(x.insns) <= AETH_TAPE_MAX_DEPTH;
    }
    fits = fits && x.leaves.size() <= AETH_TAPE_MAX_LEAVES && x.consts.size() <= AETH_TAPE_MAX_CONSTS &&
           x.rngs.size() <= AETH_TAPE_MAX_RNGS;
    if (fits) {
        x.programs.resize(N);
        for (std::size_t c = 0; c < N; ++c) {
            const TapeCtx::Program& p = x.progs[c];
            aeth_program& q = x.programs[c];
            std::memset(&q, 0, sizeof q);
            q.n_loads = (std::uint8_t)p.load_leaf.size();
            q.n_insns = (std::uint8_t)p.insns.size();
            q.out_leaf = (std::uint8_t)p.out_leaf;
            q.out_comp = (std::uint8_t)p.out_comp;
            for (std::size_t k = 0; k < p.load_leaf.size(); ++k) {
                q.load_leaf[k] = (std::uint8_t)p.load_leaf[k];
                q.load_comp[k] = (std::uint8_t)p.load_comp[k];
            }
            for (std::size_t k = 0; k < p.insns.size(); ++k)
                q.insns[k] = aeth_insn{p.insns[k].op, (std::uint8_t)p.insns[k].arg};
        }
    }
    return x;
}

// One node operation on one value: std:: math on the host, on the device the tape kernels' own device math.
template<int FN, class T>
AETHER_DEVICEHOST() inline T unary_op(T a)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (FN == AETH_OP_SQRT) {
        if constexpr (sizeof(T) == 8) return __builtin_sqrt(a);
        else return __builtin_sqrtf(a);
    } else if constexpr (FN == AETH_OP_ABS) {
        if constexpr (sizeof(T) == 8) return __builtin_fabs(a);
        else return __builtin_fabsf(a);
    } else if constexpr (FN == AETH_OP_SIGN) {
        return aeth::dm::sign(a);
    } else if constexpr (FN == AETH_OP_NEG) {
        return -a;
    } else if constexpr (sizeof(T) == 8) {
        if constexpr (FN == AETH_OP_SIN) return aeth::dm::sin(a);
        else if constexpr (FN == AETH_OP_COS) return aeth::dm::cos(a);
        else if constexpr (FN == AETH_OP_EXP) return aeth::dm::exp(a);
        else if constexpr (FN == AETH_OP_LOG) return aeth::dm::log(a);
        else if constexpr (FN == AETH_OP_TAN) return aeth::dm::tan(a);
        else if constexpr (FN == AETH_OP_ATAN) return aeth::dm::atan(a);
        else if constexpr (FN == AETH_OP_ASIN) return aeth::dm::asin(a);
        else if constexpr (FN == AETH_OP_ACOS) return aeth::dm::acos(a);
        else return aeth::dm::tanh(a);
    } else {
        if constexpr (FN == AETH_OP_SIN) return aeth::dm::sinf(a);
        else if constexpr (FN == AETH_OP_COS) return aeth::dm::cosf(a);
        else if constexpr (FN == AETH_OP_EXP) return aeth::dm::expf(a);
        else if constexpr (FN == AETH_OP_LOG) return aeth::dm::logf(a);
        else if constexpr (FN == AETH_OP_TAN) return aeth::dm::tanf(a);
        else if constexpr (FN == AETH_OP_ATAN) return aeth::dm::atanf(a);
        else if constexpr (FN == AETH_OP_ASIN) return aeth::dm::asinf(a);
        else if constexpr (FN == AETH_OP_ACOS) return aeth::dm::acosf(a);
        else return aeth::dm::tanhf(a);
    }
#else
    if constexpr (FN == AETH_OP_SIN) return std::sin(a);
    else if constexpr (FN == AETH_OP_COS) return std::cos(a);
    else if constexpr (FN == AETH_OP_EXP) return std::exp(a);
    else if constexpr (FN == AETH_OP_LOG) return std::log(a);
    else if constexpr (FN == AETH_OP_SQRT) return std::sqrt(a);
    else if constexpr (FN == AETH_OP_ABS) return std::fabs(a);
    else if constexpr (FN == AETH_OP_SIGN) return a > 0 ? T(1) : (a < 0 ? T(-1) : T(0));
    else if constexpr (FN == AETH_OP_TAN) return std::tan(a);
    else if constexpr (FN == AETH_OP_ATAN) return std::atan(a);
    else if constexpr (FN == AETH_OP_ASIN) return std::asin(a);
    else if constexpr (FN == AETH_OP_ACOS) return std::acos(a);
    else if constexpr (FN == AETH_OP_TANH) return std::tanh(a);
    else return -a;
#endif
}

template<int OP, class T>
AETHER_DEVICEHOST() inline T binary_op(T a, T b)
{
    if constexpr (OP == AETH_OP_ADD) return a + b;
    else if constexpr (OP == AETH_OP_SUB) return a - b;
    else if constexpr (OP == AETH_OP_MUL) return a * b;
    else if constexpr (OP == AETH_OP_DIV) return a / b;
#if defined(__HIP_DEVICE_COMPILE__)
    else if constexpr (OP == AETH_OP_FMIN) return aeth::dm::fmin(a, b);
    else if constexpr (OP == AETH_OP_FMAX) return aeth::dm::fmax(a, b);
    else if constexpr (OP == AETH_OP_SELGT) return aeth::dm::selgt(a, b);
    else if constexpr (sizeof(T) == 8) return aeth::dm::pow(a, b);
    else return aeth::dm::powf(a, b);
#else
    else if constexpr (OP == AETH_OP_FMIN) return std::fmin(a, b);
    else if constexpr (OP == AETH_OP_FMAX) return std::fmax(a, b);
    else if constexpr (OP == AETH_OP_SELGT) return a > b ? a : b;
    else return std::pow(a, b);
#endif
}

} // namespace detail

// ---------------------------------------------------------------- index
/** index/SampleIndex.h:36-61 */
struct SampleIndex {
    offset_t g = 0, w = 0;
    AETHER_DEVICEHOST() static constexpr SampleIndex make(std::size_t index)
    {
        return SampleIndex{(offset_t)index, (offset_t)index};
    }
    /** From a kernel's built-in indices: global = blockIdx * blockDim + threadIdx, work = threadIdx. */
    AETHER_DEVICEHOST() static constexpr SampleIndex make(offset_t thread, offset_t block, offset_t block_dim)
    {
        return SampleIndex{block * block_dim + thread, thread};
    }
    AETHER_DEVICEHOST() constexpr offset_t global() const { return g; }
    AETHER_DEVICEHOST() constexpr offset_t work() const { return w; }
};

/** index/BundleIndex.h:56-112: W consecutive samples [base, base + W) handled by one thread. Lane k is the
 *  sample base + k (its work index is k). A bundle always names a whole W-wide span; the part of a tail
 *  bundle inside an n-sample array is asked for separately with mask(n). */
template<std::size_t W>
struct BundleIndex {
    static constexpr std::size_t width = W;
    offset_t base_ = 0;

    /** Thread `thread` of block `block` (blocks of `block_dim`) takes samples W (block block_dim + thread) on. */
    AETHER_DEVICEHOST() static constexpr BundleIndex make(offset_t thread, offset_t block, offset_t block_dim)
    {
        return BundleIndex{(offset_t)W * (block * block_dim + thread)};
    }
    /** Bundle `i` of a host loop: samples W i on. */
    AETHER_DEVICEHOST() static constexpr BundleIndex make(std::size_t i) { return BundleIndex{(offset_t)(W * i)}; }
    AETHER_DEVICEHOST() constexpr offset_t base() const { return base_; }
    AETHER_DEVICEHOST() constexpr SampleIndex lane(std::size_t k) const
    {
        return SampleIndex{base_ + (offset_t)k, (offset_t)k};
    }

    /** The lanes of a bundle that lie inside the array: the first `active_` ones. */
    struct Mask {
        offset_t active_;
        AETHER_DEVICEHOST() constexpr bool lane(std::size_t k) const { return (offset_t)k < active_; }
        AETHER_DEVICEHOST() constexpr bool full() const { return active_ == (offset_t)W; }
    };
    AETHER_DEVICEHOST() constexpr Mask mask(offset_t n) const
    {
        const offset_t left = n > base_ ? n - base_ : offset_t(0);
        return Mask{left < (offset_t)W ? left : (offset_t)W};
    }
};

/** index/SampleIndex.h:78-99: `Width` consecutive samples from base_, of which the first active_ are in
 *  use (a tail packet has fewer). */
template<std::size_t Width>
struct PacketIndex {
    static constexpr std::size_t width = Width;
    std::size_t base_;
    std::size_t active_;
    AETHER_DEVICEHOST() static constexpr PacketIndex make(std::size_t base) { return PacketIndex{base, Width}; }
    AETHER_DEVICEHOST() static constexpr PacketIndex makeTail(std::size_t base, std::size_t remaining)
    {
        return PacketIndex{base, remaining};
    }
    AETHER_DEVICEHOST() constexpr bool full() const { return active_ == Width; }
    AETHER_DEVICEHOST() constexpr SampleIndex scalar(std::size_t k) const { return SampleIndex::make(base_ + k); }
};

// ---------------------------------------------------------------- expressions
template<class D>
struct Expr;

template<class E>
concept aether_expression = std::is_base_of_v<Expr<E>, E>;

template<class E>
using elem_t = typename E::element_type;

template<class L, class R, int OP>
struct BinaryNode;
template<class E, int FN>
struct UnaryNode;
template<class L, class R>
struct DotNode;
template<class L, class R>
struct CrossNode;
template<class T>
struct ScalarNode;
template<class E, bool MAX>
struct FoldNode;
template<class L, class R>
struct QuatMulNode;
template<class E>
struct QuatConjNode;
template<class Q, class V>
struct QuatRotateNode;
template<class E>
struct PureQuatNode;
template<class E, std::size_t OFF, std::size_t LEN>
struct SegmentNode;

/** CRTP base (expr/Expression.h:145-354): the per-sample reductions, geometric, quaternion and slicing
 *  members. Each is a lazy node evaluated per sample (on the device: one tape program per component);
 *  the folds keep the reference's order (expr/Reduce.h:150-164: acc = op(T(x_{k+1}), acc)). */
template<class D>
struct Expr {
    /** Whether the tree lowers to a tape (every built-in node does; a node type of the caller's own, which
     *  derives from Expr and provides `ncomp`, `element_type` and an AETHER_DEVICEHOST() `eval(c, i)`,
     *  does not: it assigns to a whole device view through a kernel of the caller's HIP translation unit). */
    static constexpr bool lowerable = false;
    AETHER_DEVICEHOST() const D& self() const { return static_cast<const D&>(*this); }
    template<class E>
    AETHER_DEVICEHOST() auto dot(const E& o) const
    {
        return DotNode<D, E>(self(), o);
    }
    template<class E>
    AETHER_DEVICEHOST() auto cross(const E& o) const
    {
        return CrossNode<D, E>(self(), o);
    }
    AETHER_DEVICEHOST() auto squaredNorm() const { return DotNode<D, D>(self(), self()); }
    AETHER_DEVICEHOST() auto norm() const { return UnaryNode<DotNode<D, D>, AETH_OP_SQRT>(squaredNorm()); }
    /** 1 / sqrt(|v|^2) (Expression.h:202-206) */
    AETHER_DEVICEHOST() auto rNorm() const { return BinaryNode<ScalarNode<typename D::element_type>, decltype(norm()), AETH_OP_DIV>(one(), norm()); }
    /** 1 / |v|^2 */
    AETHER_DEVICEHOST() auto rSquaredNorm() const
    {
        return BinaryNode<ScalarNode<typename D::element_type>, DotNode<D, D>, AETH_OP_DIV>(one(), squaredNorm());
    }
    /** |v|^2 sqrt(|v|^2) */
    AETHER_DEVICEHOST() auto cubedNorm() const
    {
        return BinaryNode<DotNode<D, D>, decltype(norm()), AETH_OP_MUL>(squaredNorm(), norm());
    }
    /** 1 / (|v|^2 sqrt(|v|^2)) */
    AETHER_DEVICEHOST() auto rCubedNorm() const
    {
        return BinaryNode<ScalarNode<typename D::element_type>, decltype(cubedNorm()), AETH_OP_DIV>(one(), cubedNorm());
    }
    /** max over components of (x < 0 ? -x : x), the ordered max (expr/Reduce.h:74-130) */
    AETHER_DEVICEHOST() auto maxNorm() const { return FoldNode<D, true>(self()); }
    /** sum of the components */
    AETHER_DEVICEHOST() auto sum() const { return FoldNode<D, false>(self()); }
    /** v * rNorm() (expr/nodes/Geometric.h UnitVector) */
    AETHER_DEVICEHOST() auto unitVector() const { return BinaryNode<D, decltype(rNorm()), AETH_OP_MUL>(self(), rNorm()); }
    /** the Hadamard product (expr/nodes/Product.h) */
    template<class E>
    AETHER_DEVICEHOST() auto cwiseMul(const E& o) const
    {
        return BinaryNode<D, E, AETH_OP_MUL>(self(), o);
    }
    /** Hamilton product (expr/nodes/Quaternion.h:35-80) */
    template<class E>
    AETHER_DEVICEHOST() auto quatMul(const E& o) const
    {
        return QuatMulNode<D, E>(self(), o);
    }
    AETHER_DEVICEHOST() auto quatConj() const { return QuatConjNode<D>(self()); }
    AETHER_DEVICEHOST() auto quatReciprocal() const
    {
        return BinaryNode<QuatConjNode<D>, decltype(rSquaredNorm()), AETH_OP_MUL>(quatConj(), rSquaredNorm());
    }
    /** rotation of the 3-vector v by this unit quaternion (Quaternion.h:117-170) */
    template<class V>
    AETHER_DEVICEHOST() auto quatRotate(const V& v) const
    {
        return QuatRotateNode<D, V>(self(), v);
    }
    AETHER_DEVICEHOST() auto asPureQuaternion() const { return PureQuatNode<D>(self()); }
    template<std::size_t OFF, std::size_t LEN>
    AETHER_DEVICEHOST() auto segment() const
    {
        return SegmentNode<D, OFF, LEN>(self());
    }
    template<std::size_t N>
    AETHER_DEVICEHOST() auto head() const
    {
        return segment<0, N>();
    }
    template<std::size_t N>
    AETHER_DEVICEHOST() auto tail() const
    {
        return segment<D::ncomp - N, N>();
    }
    AETHER_DEVICEHOST() auto asBack3DVector() const { return tail<3>(); }

private:
    AETHER_DEVICEHOST() static auto one() { return ScalarNode<typename D::element_type>(1); }
};

template<class T>
struct ScalarNode : Expr<ScalarNode<T>> {
    using element_type = T;
    static constexpr std::size_t ncomp = 1;
    static constexpr bool lowerable = true;
    T v;
    AETHER_DEVICEHOST() explicit ScalarNode(T x) : v(x) {}
    void emit(std::size_t, detail::TapeCtx& c) const { c.constant((double)v); }
    AETHER_DEVICEHOST() T eval(std::size_t, std::size_t) const { return v; }
};

/**
 * Item<T, N> (view/Item.h:60-330, vector rank): N values held by value, a register-resident vector.
 * Constructed from N values, filled(v), Zeros(), Ones(), or the value of an expression at sample 0 (a
 * sample-free tree of Items and constants; `SampleRef::get()` gives a view's sample as one). As an
 * expression leaf it is the same N values at every sample (a whole-array assignment lowers them to
 * constants). Assigning a tree stores component c before it evaluates component c + 1, as a view's sample
 * assignment does.
 */
template<class T, std::size_t N>
class Item : public Expr<Item<T, N>> {
public:
    using element_type = T;
    static constexpr std::size_t ncomp = N;
    static constexpr std::size_t Size = N;
    static constexpr bool lowerable = true;

    constexpr Item() = default;
    template<class... Vs>
        requires(sizeof...(Vs) == N && N >= 2 && (std::is_arithmetic_v<Vs> && ...))
    AETHER_DEVICEHOST() constexpr Item(Vs... vs) : data_{static_cast<T>(vs)...}
    {
    }
    AETHER_DEVICEHOST() constexpr explicit Item(const T& v)
        requires(N == 1)
        : data_{v}
    {
    }
    template<class E>
        requires(std::is_base_of_v<Expr<E>, E> && !std::is_same_v<E, Item>)
    AETHER_DEVICEHOST() Item(const E& e)
    {
        *this = e;
    }
    template<class E>
        requires(std::is_base_of_v<Expr<E>, E> && !std::is_same_v<E, Item>)
    AETHER_DEVICEHOST() Item& operator=(const E& e)
    {
        static_assert(E::ncomp == N || E::ncomp == 1, "aether: expression and item component counts differ");
        for (std::size_t c = 0; c < N; ++c) data_[c] = (T)e.eval(E::ncomp > 1 ? c : 0, 0);
        return *this;
    }

    AETHER_DEVICEHOST() static constexpr Item filled(const T& v)
    {
        Item r;
        for (std::size_t c = 0; c < N; ++c) r.data_[c] = v;
        return r;
    }
    AETHER_DEVICEHOST() static constexpr Item Zeros() { return filled(T(0)); }
    AETHER_DEVICEHOST() static constexpr Item Ones() { return filled(T(1)); }

    AETHER_DEVICEHOST() constexpr T& operator()(std::size_t c) { return data_[c]; }
    AETHER_DEVICEHOST() constexpr const T& operator()(std::size_t c) const { return data_[c]; }
    template<std::size_t I>
    AETHER_DEVICEHOST() constexpr T& get()
    {
        static_assert(I < N, "aether: Item::get<I>() out of range");
        return data_[I];
    }
    template<std::size_t I>
    AETHER_DEVICEHOST() constexpr const T& get() const
    {
        static_assert(I < N, "aether: Item::get<I>() out of range");
        return data_[I];
    }
    AETHER_DEVICEHOST() constexpr T* data() { return data_; }
    AETHER_DEVICEHOST() constexpr const T* data() const { return data_; }
    AETHER_DEVICEHOST() static constexpr std::size_t size() { return N; }

    void emit(std::size_t c, detail::TapeCtx& x) const { x.constant((double)data_[c]); }
    AETHER_DEVICEHOST() T eval(std::size_t c, std::size_t) const { return data_[c]; }

private:
    T data_[N] = {};
};

template<class L, class R, int OP>
struct BinaryNode : Expr<BinaryNode<L, R, OP>> {
    using element_type = elem_t<L>;
    static constexpr std::size_t ncomp = L::ncomp > R::ncomp ? L::ncomp : R::ncomp;
    static constexpr bool lowerable = L::lowerable && R::lowerable;
    static_assert(L::ncomp == R::ncomp || L::ncomp == 1 || R::ncomp == 1, "aether: component counts differ");
    static_assert(std::is_same_v<elem_t<L>, elem_t<R>>, "aether: element types differ");
    L l;
    R r;
    AETHER_DEVICEHOST() BinaryNode(const L& a, const R& b) : l(a), r(b) {}
    void emit(std::size_t c, detail::TapeCtx& x) const
    {
        l.emit(L::ncomp > 1 ? c : 0, x);
        r.emit(R::ncomp > 1 ? c : 0, x);
        x.op(OP);
    }
    AETHER_DEVICEHOST() element_type eval(std::size_t c, std::size_t i) const
    {
        const element_type a = l.eval(L::ncomp > 1 ? c : 0, i), b = r.eval(R::ncomp > 1 ? c : 0, i);
        return detail::binary_op<OP>(a, b);
    }
};

template<class E, int FN>
struct UnaryNode : Expr<UnaryNode<E, FN>> {
    using element_type = elem_t<E>;
    static constexpr std::size_t ncomp = E::ncomp;
    static constexpr bool lowerable = E::lowerable;
    E e;
    AETHER_DEVICEHOST() explicit UnaryNode(const E& x) : e(x) {}
    void emit(std::size_t c, detail::TapeCtx& x) const
    {
        e.emit(c, x);
        x.op(FN);
    }
    AETHER_DEVICEHOST() element_type eval(std::size_t c, std::size_t i) const { return detail::unary_op<FN>(e.eval(c, i)); }
};

/** expr/Reduce.h:150-164: acc = T(x_0), then acc = op(T(x_{k+1}), acc); sum: T = x, op = +; maxNorm:
 *  T(x) = x < 0 ? -x : x (as SELGT(-x, x)), op = the ordered max SELGT. All N terms are pushed first
 *  (highest component first), so each op's left operand is the new term. */
template<class E, bool MAX>
struct FoldNode : Expr<FoldNode<E, MAX>> {
    using element_type = elem_t<E>;
    static constexpr std::size_t ncomp = 1;
    static constexpr bool lowerable = E::lowerable;
    E e;
    AETHER_DEVICEHOST() explicit FoldNode(const E& x) : e(x) {}
    void emit(std::size_t, detail::TapeCtx& x) const
    {
        for (std::size_t k = E::ncomp; k-- > 0;) {
            e.emit(k, x);
            if constexpr (MAX) {
                x.op(AETH_OP_NEG);
                e.emit(k, x);
                x.op(AETH_OP_SELGT);
            }
        }
        for (std::size_t k = 1; k < E::ncomp; ++k) x.op(MAX ? AETH_OP_SELGT : AETH_OP_ADD);
    }
    AETHER_DEVICEHOST() element_type term(std::size_t k, std::size_t i) const
    {
        const element_type a = e.eval(k, i);
        if constexpr (MAX) return -a > a ? -a : a;
        else return a;
    }
    AETHER_DEVICEHOST() element_type eval(std::size_t, std::size_t i) const
    {
        element_type acc = term(0, i);
        for (std::size_t k = 1; k < E::ncomp; ++k) {
            const element_type t = term(k, i);
            if constexpr (MAX) acc = t > acc ? t : acc;
            else acc = t + acc;
        }
        return acc;
    }
};

/** expr/nodes/Quaternion.h:35-80: Hamilton product, left to right as written there */
template<class L, class R>
struct QuatMulNode : Expr<QuatMulNode<L, R>> {
    using element_type = elem_t<L>;
    static constexpr std::size_t ncomp = 4;
    static constexpr bool lowerable = L::lowerable && R::lowerable;
    static_assert(L::ncomp == 4 && R::ncomp == 4, "aether: quatMul needs quaternions");
    L l;
    R r;
    AETHER_DEVICEHOST() QuatMulNode(const L& a, const R& b) : l(a), r(b) {}
    // component c = l[A0] r[B0] s1 l[A1] r[B1] s2 l[A2] r[B2] s3 l[A3] r[B3], s = +1 / -1
    static constexpr int A[4][4] = {{0, 1, 2, 3}, {0, 1, 2, 3}, {0, 1, 2, 3}, {0, 1, 2, 3}};
    static constexpr int B[4][4] = {{0, 1, 2, 3}, {1, 0, 3, 2}, {2, 3, 0, 1}, {3, 2, 1, 0}};
    static constexpr int S[4][3] = {{-1, -1, -1}, {1, 1, -1}, {-1, 1, 1}, {1, -1, 1}};
    void emit(std::size_t c, detail::TapeCtx& x) const
    {
        l.emit(A[c][0], x);
        r.emit(B[c][0], x);
        x.op(AETH_OP_MUL);
        for (int t = 1; t < 4; ++t) {
            l.emit(A[c][t], x);
            r.emit(B[c][t], x);
            x.op(AETH_OP_MUL);
            x.op(S[c][t - 1] > 0 ? AETH_OP_ADD : AETH_OP_SUB);
        }
    }
    AETHER_DEVICEHOST() element_type eval(std::size_t c, std::size_t i) const
    {
        element_type acc = l.eval(A[c][0], i) * r.eval(B[c][0], i);
        for (int t = 1; t < 4; ++t) {
            const element_type p = l.eval(A[c][t], i) * r.eval(B[c][t], i);
            acc = S[c][t - 1] > 0 ? acc + p : acc - p;
        }
        return acc;
    }
};

/** Quaternion.h:83-110: (x0, -x1, -x2, -x3), true negation */
template<class E>
struct QuatConjNode : Expr<QuatConjNode<E>> {
    using element_type = elem_t<E>;
    static constexpr std::size_t ncomp = 4;
    static constexpr bool lowerable = E::lowerable;
    static_assert(E::ncomp == 4, "aether: quatConj needs a quaternion");
    E e;
    AETHER_DEVICEHOST() explicit QuatConjNode(const E& x) : e(x) {}
    void emit(std::size_t c, detail::TapeCtx& x) const
    {
        e.emit(c, x);
        if (c > 0) x.op(AETH_OP_NEG);
    }
    AETHER_DEVICEHOST() element_type eval(std::size_t c, std::size_t i) const { return c > 0 ? -e.eval(c, i) : e.eval(c, i); }
};

/** Quaternion.h:117-170: t = 2 (u x v); v' = v + w t + u x t, each sum as written there */
template<class Q, class V>
struct QuatRotateNode : Expr<QuatRotateNode<Q, V>> {
    using element_type = elem_t<Q>;
    static constexpr std::size_t ncomp = 3;
    static constexpr bool lowerable = Q::lowerable && V::lowerable;
    static_assert(Q::ncomp == 4 && V::ncomp == 3, "aether: quatRotate needs a quaternion and a 3-vector");
    Q q;
    V v;
    AETHER_DEVICEHOST() QuatRotateNode(const Q& a, const V& b) : q(a), v(b) {}
    // t_c = 2 * (u[I] v[J] - u[J] v[I]) with u = q[1..3]
    static constexpr std::size_t I[3] = {1, 2, 0}, J[3] = {2, 0, 1};
    void emit_t(std::size_t c, detail::TapeCtx& x) const
    {
        x.constant(2.0);
        q.emit(1 + I[c], x);
        v.emit(J[c], x);
        x.op(AETH_OP_MUL);
        q.emit(1 + J[c], x);
        v.emit(I[c], x);
        x.op(AETH_OP_MUL);
        x.op(AETH_OP_SUB);
        x.op(AETH_OP_MUL);
    }
    void emit(std::size_t c, detail::TapeCtx& x) const
    {
        v.emit(c, x);
        q.emit(0, x);
        emit_t(c, x);
        x.op(AETH_OP_MUL);
        x.op(AETH_OP_ADD);
        q.emit(1 + I[c], x);
        emit_t(J[c], x);
        x.op(AETH_OP_MUL);
        q.emit(1 + J[c], x);
        emit_t(I[c], x);
        x.op(AETH_OP_MUL);
        x.op(AETH_OP_SUB);
        x.op(AETH_OP_ADD);
    }
    AETHER_DEVICEHOST() element_type t(std::size_t c, std::size_t i) const
    {
        return element_type(2) * (q.eval(1 + I[c], i) * v.eval(J[c], i) - q.eval(1 + J[c], i) * v.eval(I[c], i));
    }
    AETHER_DEVICEHOST() element_type eval(std::size_t c, std::size_t i) const
    {
        return v.eval(c, i) + q.eval(0, i) * t(c, i) + (q.eval(1 + I[c], i) * t(J[c], i) - q.eval(1 + J[c], i) * t(I[c], i));
    }
};

/** Quaternion.h:173-205: (0, x0, x1, x2) */
template<class E>
struct PureQuatNode : Expr<PureQuatNode<E>> {
    using element_type = elem_t<E>;
    static constexpr std::size_t ncomp = 4;
    static constexpr bool lowerable = E::lowerable;
    static_assert(E::ncomp == 3, "aether: asPureQuaternion needs a 3-vector");
    E e;
    AETHER_DEVICEHOST() explicit PureQuatNode(const E& x) : e(x) {}
    void emit(std::size_t c, detail::TapeCtx& x) const
    {
        if (c == 0) x.constant(0.0);
        else e.emit(c - 1, x);
    }
    AETHER_DEVICEHOST() element_type eval(std::size_t c, std::size_t i) const { return c == 0 ? element_type(0) : e.eval(c - 1, i); }
};

/** expr/nodes/Structural.h Segment: components OFF .. OFF + LEN - 1 */
template<class E, std::size_t OFF, std::size_t LEN>
struct SegmentNode : Expr<SegmentNode<E, OFF, LEN>> {
    using element_type = elem_t<E>;
    static constexpr std::size_t ncomp = LEN;
    static constexpr bool lowerable = E::lowerable;
    static_assert(OFF + LEN <= E::ncomp && LEN > 0, "aether: segment out of range");
    E e;
    AETHER_DEVICEHOST() explicit SegmentNode(const E& x) : e(x) {}
    void emit(std::size_t c, detail::TapeCtx& x) const { e.emit(OFF + c, x); }
    AETHER_DEVICEHOST() element_type eval(std::size_t c, std::size_t i) const { return e.eval(OFF + c, i); }
};

/** expr/Reduce.h:172-188: (x_{N-1} y_{N-1}) + (... + (x_0 y_0)) */
template<class L, class R>
struct DotNode : Expr<DotNode<L, R>> {
    using element_type = elem_t<L>;
    static constexpr std::size_t ncomp = 1;
    static constexpr bool lowerable = L::lowerable && R::lowerable;
    static_assert(L::ncomp == R::ncomp, "aether: dot operands differ in size");
    L l;
    R r;
    AETHER_DEVICEHOST() DotNode(const L& a, const R& b) : l(a), r(b) {}
    void emit(std::size_t, detail::TapeCtx& x) const
    {
        for (std::size_t k = L::ncomp; k-- > 0;) {
            l.emit(k, x);
            r.emit(k, x);
            x.op(AETH_OP_MUL);
        }
        for (std::size_t k = 1; k < L::ncomp; ++k) x.op(AETH_OP_ADD);
    }
    AETHER_DEVICEHOST() element_type eval(std::size_t, std::size_t i) const
    {
        element_type acc = l.eval(0, i) * r.eval(0, i);
        for (std::size_t k = 1; k < L::ncomp; ++k) acc = l.eval(k, i) * r.eval(k, i) + acc;
        return acc;
    }
};

/** expr/nodes/Geometric.h:55-71 */
template<class L, class R>
struct CrossNode : Expr<CrossNode<L, R>> {
    using element_type = elem_t<L>;
    static constexpr std::size_t ncomp = 3;
    static constexpr bool lowerable = L::lowerable && R::lowerable;
    static_assert(L::ncomp == 3 && R::ncomp == 3, "aether: cross needs 3-vectors");
    L l;
    R r;
    AETHER_DEVICEHOST() CrossNode(const L& a, const R& b) : l(a), r(b) {}
    static constexpr std::size_t I[3] = {1, 2, 0}, J[3] = {2, 0, 1};
    void emit(std::size_t c, detail::TapeCtx& x) const
    {
        l.emit(I[c], x);
        r.emit(J[c], x);
        x.op(AETH_OP_MUL);
        l.emit(J[c], x);
        r.emit(I[c], x);
        x.op(AETH_OP_MUL);
        x.op(AETH_OP_SUB);
    }
    AETHER_DEVICEHOST() element_type eval(std::size_t c, std::size_t i) const
    {
        return l.eval(I[c], i) * r.eval(J[c], i) - l.eval(J[c], i) * r.eval(I[c], i);
    }
};

// Scalars wrap into ScalarNode of the other operand's element type.
template<class E, class S>
AETHER_DEVICEHOST() auto wrap_scalar(const S& s)
{
    return ScalarNode<elem_t<E>>((elem_t<E>)s);
}

template<class L, class R>
    requires aether_expression<L> && aether_expression<R>
AETHER_DEVICEHOST() auto operator+(const L& l, const R& r)
{
    return BinaryNode<L, R, AETH_OP_ADD>(l, r);
}
template<class L, class R>
    requires aether_expression<L> && aether_expression<R>
AETHER_DEVICEHOST() auto operator-(const L& l, const R& r)
{
    return BinaryNode<L, R, AETH_OP_SUB>(l, r);
}
template<class L, class R>
    requires aether_expression<L> && aether_expression<R>
AETHER_DEVICEHOST() auto operator*(const L& l, const R& r)
{
    return BinaryNode<L, R, AETH_OP_MUL>(l, r);
}
template<class L, class R>
    requires aether_expression<L> && aether_expression<R>
AETHER_DEVICEHOST() auto operator/(const L& l, const R& r)
{
    return BinaryNode<L, R, AETH_OP_DIV>(l, r);
}
template<class E, class S>
    requires aether_expression<E> && std::is_arithmetic_v<S>
AETHER_DEVICEHOST() auto operator+(const E& e, S s)
{
    return e + wrap_scalar<E>(s);
}
template<class E, class S>
    requires aether_expression<E> && std::is_arithmetic_v<S>
AETHER_DEVICEHOST() auto operator+(S s, const E& e)
{
    return wrap_scalar<E>(s) + e;
}
template<class E, class S>
    requires aether_expression<E> && std::is_arithmetic_v<S>
AETHER_DEVICEHOST() auto operator-(const E& e, S s)
{
    return e - wrap_scalar<E>(s);
}
template<class E, class S>
    requires aether_expression<E> && std::is_arithmetic_v<S>
AETHER_DEVICEHOST() auto operator-(S s, const E& e)
{
    return wrap_scalar<E>(s) - e;
}
/** CWiseScale (expr/Operations.h:49-59): s * e */
template<class E, class S>
    requires aether_expression<E> && std::is_arithmetic_v<S>
AETHER_DEVICEHOST() auto operator*(S s, const E& e)
{
    return wrap_scalar<E>(s) * e;
}
template<class E, class S>
    requires aether_expression<E> && std::is_arithmetic_v<S>
AETHER_DEVICEHOST() auto operator*(const E& e, S s)
{
    return e * wrap_scalar<E>(s);
}
/** expr/Operations.h:62-71: e / s is e * (1 / s), the reciprocal rounded in T */
template<class E, class S>
    requires aether_expression<E> && std::is_arithmetic_v<S>
AETHER_DEVICEHOST() auto operator/(const E& e, S s)
{
    return e * wrap_scalar<E>(elem_t<E>(1) / (elem_t<E>)s);
}
/** -x is x * -1 (docs/content/numerics_contract.rst:66) */
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto operator-(const E& e)
{
    return e * wrap_scalar<E>(-1);
}

template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseSin(const E& e) { return UnaryNode<E, AETH_OP_SIN>(e); }
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseCos(const E& e) { return UnaryNode<E, AETH_OP_COS>(e); }
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseExp(const E& e) { return UnaryNode<E, AETH_OP_EXP>(e); }
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseLog(const E& e) { return UnaryNode<E, AETH_OP_LOG>(e); }
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseSqrt(const E& e) { return UnaryNode<E, AETH_OP_SQRT>(e); }
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseAbs(const E& e) { return UnaryNode<E, AETH_OP_ABS>(e); }
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto sqrt(const E& e) { return UnaryNode<E, AETH_OP_SQRT>(e); }
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseSign(const E& e) { return UnaryNode<E, AETH_OP_SIGN>(e); }
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseTan(const E& e) { return UnaryNode<E, AETH_OP_TAN>(e); }
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseAtan(const E& e) { return UnaryNode<E, AETH_OP_ATAN>(e); }
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseAsin(const E& e) { return UnaryNode<E, AETH_OP_ASIN>(e); }
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseAcos(const E& e) { return UnaryNode<E, AETH_OP_ACOS>(e); }
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseTanh(const E& e) { return UnaryNode<E, AETH_OP_TANH>(e); }
/** math/root.h:74 (host): T(1) / sqrt(x), two correctly rounded operations */
template<class E>
    requires aether_expression<E>
AETHER_DEVICEHOST() auto cwiseRsqrt(const E& e) { return wrap_scalar<E>(1) / UnaryNode<E, AETH_OP_SQRT>(e); }
/** Elementwise.h:198-210: min / max against a broadcast bound */
template<class E, class S>
    requires aether_expression<E> && std::is_arithmetic_v<S>
AETHER_DEVICEHOST() auto cwiseMin(const E& e, S bound)
{
    return BinaryNode<E, ScalarNode<elem_t<E>>, AETH_OP_FMIN>(e, wrap_scalar<E>(bound));
}
template<class E, class S>
    requires aether_expression<E> && std::is_arithmetic_v<S>
AETHER_DEVICEHOST() auto cwiseMax(const E& e, S bound)
{
    return BinaryNode<E, ScalarNode<elem_t<E>>, AETH_OP_FMAX>(e, wrap_scalar<E>(bound));
}
/** Elementwise.h:761-775: min / max of two expressions */
template<class L, class R>
    requires aether_expression<L> && aether_expression<R>
AETHER_DEVICEHOST() auto cwiseMin(const L& l, const R& r)
{
    return BinaryNode<L, R, AETH_OP_FMIN>(l, r);
}
template<class L, class R>
    requires aether_expression<L> && aether_expression<R>
AETHER_DEVICEHOST() auto cwiseMax(const L& l, const R& r)
{
    return BinaryNode<L, R, AETH_OP_FMAX>(l, r);
}
/** Elementwise.h:219-224: fmin(fmax(e, lo), hi) */
template<class E, class S>
    requires aether_expression<E> && std::is_arithmetic_v<S>
AETHER_DEVICEHOST() auto clamp(const E& e, S lo, S hi)
{
    return cwiseMin(cwiseMax(e, lo), hi);
}
/** Elementwise.h:742-746: l / r, both expressions */
template<class L, class R>
    requires aether_expression<L> && aether_expression<R>
AETHER_DEVICEHOST() auto cwiseDiv(const L& l, const R& r)
{
    return BinaryNode<L, R, AETH_OP_DIV>(l, r);
}
/** Elementwise.h:749-760: e ^ exponent (a broadcast scalar) or e ^ r */
template<class E, class S>
    requires aether_expression<E> && std::is_arithmetic_v<S>
AETHER_DEVICEHOST() auto cwisePow(const E& e, S exponent)
{
    return BinaryNode<E, ScalarNode<elem_t<E>>, AETH_OP_POW>(e, wrap_scalar<E>(exponent));
}
template<class L, class R>
    requires aether_expression<L> && aether_expression<R>
AETHER_DEVICEHOST() auto cwisePow(const L& l, const R& r)
{
    return BinaryNode<L, R, AETH_OP_POW>(l, r);
}

// ---------------------------------------------------------------- views
template<class T, std::size_t N>
class View;

/** A view as an expression leaf (view/View.h:150-424). */
template<class T, std::size_t N>
struct ViewLeaf : Expr<ViewLeaf<T, N>> {
    using element_type = std::remove_const_t<T>;
    static constexpr std::size_t ncomp = N;
    static constexpr bool lowerable = true;
    T* p;
    std::size_t pitch;
    void emit(std::size_t c, detail::TapeCtx& x) const { x.push(p, pitch, c); }
    AETHER_DEVICEHOST() element_type eval(std::size_t c, std::size_t i) const { return p[c * pitch + i]; }
};

enum class SumMode : int { None = AETH_SUM_NONE, Neumaier = AETH_SUM_NEUMAIER, Plain = AETH_SUM_PLAIN,
                           FloatFloat = AETH_SUM_FLOATFLOAT };

/** SampleRef (expr/Assign.h:251-346): host per-sample assignment. */
template<class T, std::size_t N>
class SampleRef {
public:
    AETHER_DEVICEHOST() SampleRef(T* p, std::size_t pitch, std::size_t i) : p_(p), pitch_(pitch), i_(i) {}
    template<class E>
        requires aether_expression<E>
    AETHER_DEVICEHOST() SampleRef& operator=(const E& e)
    {
        for (std::size_t c = 0; c < N; ++c) p_[c * pitch_ + i_] = (T)e.eval(E::ncomp > 1 ? c : 0, i_);
        return *this;
    }
    template<class E>
        requires aether_expression<E>
    AETHER_DEVICEHOST() SampleRef& operator+=(const E& e)
    {
        for (std::size_t c = 0; c < N; ++c) p_[c * pitch_ + i_] += (T)e.eval(E::ncomp > 1 ? c : 0, i_);
        return *this;
    }
    template<class E>
        requires aether_expression<E>
    AETHER_DEVICEHOST() SampleRef& operator-=(const E& e)
    {
        for (std::size_t c = 0; c < N; ++c) p_[c * pitch_ + i_] -= (T)e.eval(E::ncomp > 1 ? c : 0, i_);
        return *this;
    }
    /** A scalar broadcast to every component (expr/Assign.h:296-306): = sets, += adds, -= subtracts. */
    template<class S>
        requires(!aether_expression<S> && std::is_arithmetic_v<S>)
    AETHER_DEVICEHOST() SampleRef& operator=(const S& s)
    {
        for (std::size_t c = 0; c < N; ++c) p_[c * pitch_ + i_] = (T)s;
        return *this;
    }
    template<class S>
        requires(!aether_expression<S> && std::is_arithmetic_v<S>)
    AETHER_DEVICEHOST() SampleRef& operator+=(const S& s)
    {
        for (std::size_t c = 0; c < N; ++c) p_[c * pitch_ + i_] += (T)s;
        return *this;
    }
    template<class S>
        requires(!aether_expression<S> && std::is_arithmetic_v<S>)
    AETHER_DEVICEHOST() SampleRef& operator-=(const S& s)
    {
        for (std::size_t c = 0; c < N; ++c) p_[c * pitch_ + i_] -= (T)s;
        return *this;
    }
    /** Not rebindable: `v[i] = w[j]` is not an assignment of values (the reference's SampleRef holds its
     *  view by reference); assign `w[j].get()` instead. */
    SampleRef& operator=(const SampleRef&) = delete;
    SampleRef(const SampleRef&) = default;
    /** The sample's components as an Item (expr/Assign.h:336-341). */
    AETHER_DEVICEHOST() Item<T, N> get() const
    {
        Item<T, N> r;
        for (std::size_t c = 0; c < N; ++c) r(c) = p_[c * pitch_ + i_];
        return r;
    }

private:
    T* p_;
    std::size_t pitch_, i_;
};

namespace detail {

/** Store the W values of one component of a bundle: whole 16-byte stores when the bundle is full and its
 *  first sample 16-byte aligned, lane by lane otherwise (a tail bundle, or a view whose planes start off a
 *  16-byte boundary). */
template<class T, std::size_t W>
AETHER_DEVICEHOST() inline void store_lanes(T* dst, const T (&v)[W], std::size_t active)
{
    if constexpr ((W * sizeof(T)) % 16 == 0) {
        if (active == W && reinterpret_cast<std::uintptr_t>(dst) % 16 == 0) {
            __builtin_memcpy(static_cast<T*>(__builtin_assume_aligned(dst, 16)), v, W * sizeof(T));
            return;
        }
    }
    for (std::size_t k = 0; k < active; ++k) dst[k] = v[k];
}

} // namespace detail

/** View[BundleIndex<W>] (backend/cuda/bundle/Assign.h:56-265): the W samples of a bundle, assigned one
 *  component at a time over its lanes. Each sample gets the bits of the per-sample assignment (the same
 *  node arithmetic; a sample's component c is stored before its component c + 1 is evaluated). */
template<class T, std::size_t N, std::size_t W>
class BundleRef {
public:
    AETHER_DEVICEHOST() BundleRef(T* p, std::size_t pitch, std::size_t base, std::size_t active)
        : p_(p), pitch_(pitch), base_(base), active_(active)
    {
    }
    template<class E>
    AETHER_DEVICEHOST() BundleRef& operator=(const E& e)
    {
        for (std::size_t c = 0; c < N; ++c) {
            T v[W];
            for (std::size_t k = 0; k < W; ++k)
                v[k] = k < active_ ? (T)e.eval(E::ncomp > 1 ? c : 0, base_ + k) : T(0);
            detail::store_lanes<T, W>(p_ + c * pitch_ + base_, v, active_);
        }
        return *this;
    }

private:
    T* p_;
    std::size_t pitch_, base_, active_;
};

namespace detail {

template<class E>
inline constexpr bool always_false = false;

#if defined(__HIPCC__)
/** dest = e for a tree holding a node type of the caller's own: one sample per lane, grid-stride, its
 *  components stored in order; instantiated in the caller's HIP translation unit. */
template<class T, std::size_t N, class E>
__global__ __launch_bounds__(256) void assign_kernel(T* p, std::size_t pitch, std::uint32_t n, E e)
{
    // component c is stored before component c + 1 is evaluated, as the reference assigns a sample
    // (expr/Assign.h:124-160): a tree that reads the destination (v = f(v).cross(w)) sees the components
    // of this sample already stored
    for (std::uint64_t i = (std::uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (std::uint64_t)gridDim.x * 256u)
        for (std::size_t c = 0; c < N; ++c) p[c * pitch + i] = (T)e.eval(E::ncomp > 1 ? c : 0, (std::size_t)i);
}
#endif

} // namespace detail

/**
 * View<T, N>: `samples` items of N components, plane pitch `pitch`
 * (layout_stride with strides {pitch, 1}). A device view assigned an
 * expression runs one kernel over all samples; a host view supports the
 * per-sample spelling and element access.
 */
template<class T, std::size_t N>
class View : public Expr<View<T, N>> {
public:
    using element_type = T;
    static constexpr std::size_t ncomp = N;
    static constexpr bool lowerable = true;

    AETHER_DEVICEHOST() View(T* p, std::size_t samples, std::size_t pitch, bool device)
        : p_(p), samples_(samples), pitch_(pitch), device_(device)
    {
    }
    View(const View&) = default;
    AETHER_DEVICEHOST() T* data() const { return p_; }
    AETHER_DEVICEHOST() std::size_t samples() const { return samples_; }
    AETHER_DEVICEHOST() std::size_t pitch() const { return pitch_; }
    AETHER_DEVICEHOST() bool on_device() const { return device_; }
    AETHER_DEVICEHOST() T& operator()(std::size_t c, std::size_t i) const { return p_[c * pitch_ + i]; }
    AETHER_DEVICEHOST() T& operator()(std::size_t i) const { return p_[i]; }
    AETHER_DEVICEHOST() SampleRef<T, N> operator[](SampleIndex i) const { return SampleRef<T, N>(p_, pitch_, i.global()); }
    /** A whole bundle (every lane inside the view); a tail bundle goes through bundleAssign with its mask. */
    template<std::size_t W>
    AETHER_DEVICEHOST() BundleRef<T, N, W> operator[](BundleIndex<W> b) const
    {
        return BundleRef<T, N, W>(p_, pitch_, b.base(), W);
    }

    void emit(std::size_t c, detail::TapeCtx& x) const { x.push(p_, pitch_, c); }
    AETHER_DEVICEHOST() T eval(std::size_t c, std::size_t i) const { return p_[c * pitch_ + i]; }

    /** Whole-array lazy assignment: one kernel (device views only). */
    template<class E>
        requires aether_expression<E>
    View& operator=(const E& e)
    {
        assign(e, SumMode::None, nullptr, nullptr);
        return *this;
    }
    View& operator=(const View& o)
    {
        if (this != &o) assign(o, SumMode::None, nullptr, nullptr);
        return *this;
    }

    /**
     * dest = e, fused with per-component sums of dest (sums_host: N values
     * hi + comp when mode != None). Synchronises the stream when sums are asked.
     */
    template<class E>
    void assign(const E& e, SumMode mode, double* sums_host, Stream stream)
    {
        T pairs[2 * N];
        assign_pairs(e, mode, pairs, stream);
        if (mode != SumMode::None)
            for (std::size_t c = 0; c < N; ++c) sums_host[c] = (double)pairs[2 * c] + (double)pairs[2 * c + 1];
    }

    /** dest = e with the sums as (hi, comp) pairs of T per component in pairs_host (2N values, when
     *  mode != None): the form sums of partial arrays combine in (combine_ranks). */
    template<class E>
    void assign_pairs(const E& e, SumMode mode, T* pairs_host, Stream stream)
    {
        static_assert(E::ncomp == N || E::ncomp == 1, "aether: expression and view component counts differ");
        if (!device_) throw Error("aether: whole-array assignment needs a device view");
        if constexpr (!E::lowerable) {
            assign_generic(e, mode, pairs_host, stream);
        } else {
            const detail::TapeCtx x = detail::lower<T, N>(p_, pitch_, e);
            void* sums_dev = nullptr;
            void* chunk = nullptr;
            const int dev = detail::current_device();
            if (mode != SumMode::None)
                detail::check(detail::seam().chunk_alloc(AETHER_MEM_DEVICE, dev, 2 * N * sizeof(T), stream, &chunk,
                                                         &sums_dev),
                              "chunk_alloc");
            const std::vector<aeth_program_ex> progs = x.ex();
            const int rc = aeth_eval_tape_ex(x.dtype, (int)N, progs.data(), (int)x.leaves.size(),
                                             const_cast<void* const*>(x.leaves.data()), x.pitch.data(),
                                             (int)x.consts.size(), x.consts.data(), (int)x.rngs.size(), x.rngs.data(),
                                             samples_, (int)mode, sums_dev, stream);
            if (rc == 0 && mode != SumMode::None)
                detail::check(detail::seam().copy(AETHER_MEM_CPU, dev, pairs_host, AETHER_MEM_DEVICE, dev, sums_dev,
                                                  2 * N * sizeof(T), stream),
                              "copy");
            if (chunk) detail::seam().chunk_release(chunk);
            detail::check(rc, "aeth_eval_tape");
        }
    }

    /** Compile and load every kernel that assigning e (with `mode`) would launch, without running anything
     *  (aeth_tape_prepare): the first assignment then pays no runtime compile. A tree holding a node type of
     *  the caller's own has nothing to compile at run time. */
    template<class E>
    void prepare(const E& e, SumMode mode = SumMode::None, Stream stream = nullptr) const
    {
        static_assert(E::ncomp == N || E::ncomp == 1, "aether: expression and view component counts differ");
        if (!device_) throw Error("aether: preparing an assignment needs a device view");
        if constexpr (E::lowerable) {
            const detail::TapeCtx x = detail::lower<T, N>(p_, pitch_, e);
            const std::vector<aeth_program_ex> progs = x.ex();
            detail::check(aeth_tape_prepare(x.dtype, (int)N, progs.data(), (int)x.leaves.size(),
                                            const_cast<void* const*>(x.leaves.data()), x.pitch.data(),
                                            (int)x.consts.size(), x.consts.data(), (int)x.rngs.size(),
                                            x.rngs.data(), samples_, (int)mode, stream),
                          "aeth_tape_prepare");
        }
    }

    /** dest = e for a tree holding a node type of the caller's own (no tape lowering): a kernel of the
     *  caller's HIP translation unit evaluates it per sample, and the sums (when asked) are the library's
     *  reduction of dest in the fixed shape, which is bit for bit the fused sum of the same values. */
    template<class E>
    void assign_generic(const E& e, SumMode mode, T* pairs_host, Stream stream)
    {
#if defined(__HIPCC__)
        if (samples_ > 0xFFFFFFFFull) throw Error("aether: a view of more than 2^32-1 samples");
        if (samples_) {
            const std::size_t blocks = std::min<std::size_t>((samples_ + 255) / 256, 16384);
            hipLaunchKernelGGL((detail::assign_kernel<T, N, E>), dim3((unsigned)blocks), dim3(256), 0,
                               reinterpret_cast<hipStream_t>(stream), p_, pitch_, (std::uint32_t)samples_, e);
            const hipError_t err = hipGetLastError();
            if (err != hipSuccess) throw Error(std::string("aether: assignment kernel launch: ") + hipGetErrorString(err));
        }
        if (mode == SumMode::None) return;
        void* sums_dev = nullptr;
        void* chunk = nullptr;
        const int dev = detail::current_device();
        detail::check(detail::seam().chunk_alloc(AETHER_MEM_DEVICE, dev, 2 * N * sizeof(T), stream, &chunk, &sums_dev),
                      "chunk_alloc");
        const int rc = aeth_reduce(detail::dtype_code<T>(), (int)mode, 0, p_, (int)N, samples_, pitch_, sums_dev, stream);
        if (rc == 0)
            detail::check(detail::seam().copy(AETHER_MEM_CPU, dev, pairs_host, AETHER_MEM_DEVICE, dev, sums_dev,
                                              2 * N * sizeof(T), stream),
                          "copy");
        detail::seam().chunk_release(chunk);
        detail::check(rc, "aeth_reduce");
#else
        (void)e, (void)mode, (void)pairs_host, (void)stream;
        static_assert(detail::always_false<E>, "aether: an expression holding a node type of the caller's own "
                                               "assigns to a whole device view only in a HIP translation unit "
                                               "(hipcc); a host translation unit lowers built-in nodes only");
#endif
    }

    /** Autotune the launch shape of dest = e (aeth_tape_tune); this view holds the result afterwards. With a
     *  sum mode (aeth_tape_tune_sums) only the occupancy target varies and later evaluate() calls of the
     *  structure with that mode use the pick. */
    template<class E>
    aeth_tape_shape tune(const E& e, Stream stream = nullptr, SumMode mode = SumMode::None)
    {
        static_assert(E::ncomp == N || E::ncomp == 1, "aether: expression and view component counts differ");
        static_assert(E::lowerable, "aether: tuning takes a tree of built-in nodes (it tunes the runtime-compiled kernel)");
        if (!device_) throw Error("aether: tuning needs a device view");
        const detail::TapeCtx x = detail::lower<T, N>(p_, pitch_, e);
        if (x.programs.empty()) throw Error("aether: tuning takes an expression that fits one launch");
        aeth_tape_shape chosen{};
        if (mode == SumMode::None) {
            detail::check(aeth_tape_tune(x.dtype, (int)N, x.programs.data(), (int)x.leaves.size(),
                                         const_cast<void* const*>(x.leaves.data()), x.pitch.data(),
                                         (int)x.consts.size(), x.consts.data(), (int)x.rngs.size(), x.rngs.data(),
                                         samples_, stream, &chosen),
                          "aeth_tape_tune");
            return chosen;
        }
        void* sums_dev = nullptr;
        void* chunk = nullptr;
        const int dev = detail::current_device();
        detail::check(detail::seam().chunk_alloc(AETHER_MEM_DEVICE, dev, 2 * N * sizeof(T), stream, &chunk, &sums_dev),
                      "chunk_alloc");
        const int rc = aeth_tape_tune_sums(x.dtype, (int)N, x.programs.data(), (int)x.leaves.size(),
                                           const_cast<void* const*>(x.leaves.data()), x.pitch.data(),
                                           (int)x.consts.size(), x.consts.data(), (int)x.rngs.size(), x.rngs.data(),
                                           samples_, (int)mode, sums_dev, stream, &chosen);
        detail::seam().chunk_release(chunk);
        detail::check(rc, "aeth_tape_tune_sums");
        return chosen;
    }

private:
    T* p_;
    std::size_t samples_, pitch_;
    bool device_;
};

// ---------------------------------------------------------------- Array
/**
 * Array<T, N> (array/Array.h:82-364): N components per sample, SoA planes of
 * capacity() elements (samples rounded up to 32), one pinned host copy and
 * one device copy, moved with upload()/download().
 */
template<class T, std::size_t N = 1>
class Array {
public:
    explicit Array(std::size_t n_samples) : samples_(n_samples), capacity_(quantize(n_samples)) { allocate(); }
    Array(const Array&) = delete;
    Array& operator=(const Array&) = delete;
    Array(Array&& o) noexcept { swap(o); }
    Array& operator=(Array&& o) noexcept
    {
        swap(o);
        return *this;
    }
    ~Array() { release(); }

    /** array/Array.h:89-94: the view types (planes at pitch capacity(); the packed form at pitch samples()). */
    using ViewT = View<T, N>;
    using PackedViewT = View<T, N>;

    std::size_t samples() const { return samples_; }
    std::size_t size() const { return N * samples_; }
    std::size_t capacity() const { return capacity_; }
    View<T, N> hostView() const { return View<T, N>(static_cast<T*>(host_), samples_, capacity_, false); }
    View<T, N> deviceView() const { return View<T, N>(static_cast<T*>(dev_), samples_, capacity_, true); }
    /** array/Array.h:176-218: the compact view (plane pitch = samples()), only for an Array whose capacity is
     *  its sample count (throws otherwise). */
    View<T, N> hostPacked() const
    {
        require_packed();
        return View<T, N>(static_cast<T*>(host_), samples_, samples_, false);
    }
    View<T, N> devicePacked() const
    {
        require_packed();
        return View<T, N>(static_cast<T*>(dev_), samples_, samples_, true);
    }
    void upload(Stream s = nullptr)
    {
        const int dev = detail::current_device();
        detail::check(detail::seam().copy(AETHER_MEM_DEVICE, dev, dev_, AETHER_MEM_PINNED, dev, host_, bytes(),
                                          s),
                      "upload");
        detail::check(detail::seam().stream_sync(dev, s), "upload");
    }
    void download(Stream s = nullptr)
    {
        const int dev = detail::current_device();
        detail::check(detail::seam().copy(AETHER_MEM_PINNED, dev, host_, AETHER_MEM_DEVICE, dev, dev_, bytes(),
                                          s),
                      "download");
        detail::check(detail::seam().stream_sync(dev, s), "download");
    }
    /** Grow capacity (array/Array.h:120-135): pitched copy of the samples on both sides. */
    void reserve(std::size_t n)
    {
        if (n <= capacity_) return;
        Array next(samples_, quantize(n));
        const int dev = detail::current_device();
        for (std::size_t c = 0; c < N && samples_ > 0; ++c) {
            detail::check(detail::seam().copy(AETHER_MEM_DEVICE, dev, static_cast<T*>(next.dev_) + c * next.capacity_,
                                              AETHER_MEM_DEVICE, dev, static_cast<T*>(dev_) + c * capacity_,
                                              samples_ * sizeof(T), nullptr),
                          "reserve");
            std::memcpy(static_cast<T*>(next.host_) + c * next.capacity_, static_cast<T*>(host_) + c * capacity_,
                        samples_ * sizeof(T));
        }
        detail::check(detail::seam().stream_sync(dev, nullptr), "reserve");
        swap(next);
    }
    void resize(std::size_t n)
    {
        reserve(n);
        samples_ = n;
    }
    /** Append one sample on the host copy, doubling capacity when full (array/Array.h:155-163);
     *  upload() afterwards to reach the device copy. */
    void push_back(const std::array<T, N>& item)
    {
        if (samples_ == capacity_) reserve(capacity_ == 0 ? kCapacityAlignment : capacity_ * 2);
        for (std::size_t c = 0; c < N; ++c) static_cast<T*>(host_)[c * capacity_ + samples_] = item[c];
        ++samples_;
    }
    void clear() { samples_ = 0; }
    /** The reserved tail [samples(), capacity()) (array/Array.h:196,221). */
    View<T, N> hostSpare() const
    {
        return View<T, N>(static_cast<T*>(host_) + samples_, capacity_ - samples_, capacity_, false);
    }
    View<T, N> deviceSpare() const
    {
        return View<T, N>(static_cast<T*>(dev_) + samples_, capacity_ - samples_, capacity_, true);
    }
    /** Adopt `count` samples written into the spare tail (array/Array.h:238-247). */
    void commit(std::size_t count)
    {
        if (count > capacity_ - samples_)
            throw Error("Array::commit: count=" + std::to_string(count) + " exceeds spare capacity (" +
                        std::to_string(capacity_ - samples_) + ")");
        samples_ += count;
    }

private:
    void require_packed() const
    {
        if (capacity_ != samples_)
            throw Error("aether: a packed view needs capacity() == samples() (" + std::to_string(capacity_) + " != " +
                        std::to_string(samples_) + ")");
    }
    Array(std::size_t n_samples, std::size_t capacity) : samples_(n_samples), capacity_(capacity) { allocate(); }
    static std::size_t quantize(std::size_t n)
    {
        return (n + kCapacityAlignment - 1) / kCapacityAlignment * kCapacityAlignment;
    }
    std::size_t bytes() const { return N * capacity_ * sizeof(T); }
    void allocate()
    {
        const int dev = detail::current_device();
        const std::size_t b = bytes() ? bytes() : sizeof(T);
        detail::check(detail::seam().chunk_alloc(AETHER_MEM_PINNED, dev, b, nullptr, &host_chunk_, &host_),
                      "Array host allocation");
        detail::check(detail::seam().chunk_alloc(AETHER_MEM_DEVICE, dev, b, nullptr, &dev_chunk_, &dev_),
                      "Array device allocation");
        detail::check(detail::seam().stream_sync(dev, nullptr), "Array allocation");
    }
    void release()
    {
        if (dev_chunk_) detail::seam().chunk_release(dev_chunk_);
        if (host_chunk_) detail::seam().chunk_release(host_chunk_);
        dev_chunk_ = host_chunk_ = nullptr;
        dev_ = host_ = nullptr;
    }
    void swap(Array& o)
    {
        std::swap(samples_, o.samples_);
        std::swap(capacity_, o.capacity_);
        std::swap(host_chunk_, o.host_chunk_);
        std::swap(dev_chunk_, o.dev_chunk_);
        std::swap(host_, o.host_);
        std::swap(dev_, o.dev_);
    }
    std::size_t samples_ = 0, capacity_ = 0;
    void* host_chunk_ = nullptr;
    void* dev_chunk_ = nullptr;
    void* host_ = nullptr;
    void* dev_ = nullptr;
};

// aether.h:234-260: Vec is a register-resident Item; the *View aliases are batched views
template<class T, std::size_t N>
using Vec = Item<T, N>;
using Vec3d = Vec<double, 3>;
using Vec4d = Vec<double, 4>;
using Vec6d = Vec<double, 6>;
template<class T>
using ScalarView = View<T, 1>;
template<class T, std::size_t N>
using VecView = View<T, N>;
using Vec3dView = VecView<double, 3>;

/** dest[b] = e over the lanes of bundle `b` that `m` marks as inside the view (a tail bundle:
 *  bi.mask(n)), the spelling of backend/cuda/bundle/Assign.h's bundleAssign; lanes past the mask are
 *  neither evaluated nor stored. */
template<class T, std::size_t N, std::size_t W, class E>
AETHER_DEVICEHOST() void bundleAssign(const View<T, N>& dest, const E& e, BundleIndex<W> b,
                                      typename BundleIndex<W>::Mask m)
{
    BundleRef<T, N, W>(dest.data(), dest.pitch(), b.base(), m.active_) = e;
}

/** dest[b] = e over a whole bundle. */
template<class T, std::size_t N, std::size_t W, class E>
AETHER_DEVICEHOST() void bundleAssign(const View<T, N>& dest, const E& e, BundleIndex<W> b)
{
    dest[b] = e;
}

/** dest = e with fused per-component sums (hi + comp per component) copied to the host. */
template<class T, std::size_t N, class E>
std::array<double, N> evaluate(View<T, N> dest, const E& e, SumMode mode = SumMode::Neumaier, Stream s = nullptr)
{
    std::array<double, N> sums{};
    dest.assign(e, mode, sums.data(), s);
    return sums;
}

/** dest = e with fused per-component sums kept as (hi, comp) pairs of T. */
template<class T, std::size_t N, class E>
std::array<std::array<T, 2>, N> evaluate_pairs(View<T, N> dest, const E& e, SumMode mode = SumMode::Neumaier,
                                               Stream s = nullptr)
{
    T flat[2 * N];
    dest.assign_pairs(e, mode, flat, s);
    std::array<std::array<T, 2>, N> out{};
    for (std::size_t c = 0; c < N; ++c) out[c] = {flat[2 * c], flat[2 * c + 1]};
    return out;
}

/** Compile and load the kernels of `dest = e` (with `mode`) without running them (no counterpart in the
 *  reference, whose kernels are compiled ahead of time). */
template<class T, std::size_t N, class E>
void prepare(View<T, N> dest, const E& e, SumMode mode = SumMode::None, Stream s = nullptr)
{
    dest.prepare(e, mode, s);
}

/** Pick the fastest launch shape for this expression's structure on this device (no counterpart in the
 *  reference); later assignments of the same structure without sums use it. dest holds the result. */
template<class T, std::size_t N, class E>
aeth_tape_shape tune(View<T, N> dest, const E& e, Stream s = nullptr, SumMode mode = SumMode::None)
{
    return dest.tune(e, s, mode);
}

// ---------------------------------------------------------------- accumulation
/**
 * The owning sticky device error word (err/DeviceFlag.h:86-160): one 32-bit word in device memory that
 * kernels set through deviceView() (DeviceFlagView::set, an atomic OR) and a pinned host mirror the host
 * reads it through. Sticky: a read never clears it; clear() or checkAndClear() does. Move-only.
 */
class DeviceFlag {
public:
    DeviceFlag() : dev_(detail::current_device())
    {
        detail::check(detail::seam().chunk_alloc(AETHER_MEM_DEVICE, dev_, 16, nullptr, &dchunk_, &dword_),
                      "DeviceFlag allocation");
        if (detail::seam().chunk_alloc(AETHER_MEM_PINNED, dev_, 16, nullptr, &hchunk_, &hword_) != 0) {
            detail::seam().chunk_release(dchunk_);
            dchunk_ = nullptr;
            detail::check(1, "DeviceFlag allocation");
        }
        clear();
    }
    DeviceFlag(const DeviceFlag&) = delete;
    DeviceFlag& operator=(const DeviceFlag&) = delete;
    DeviceFlag(DeviceFlag&& o) noexcept
        : dev_(o.dev_), dchunk_(std::exchange(o.dchunk_, nullptr)), hchunk_(std::exchange(o.hchunk_, nullptr)),
          dword_(std::exchange(o.dword_, nullptr)), hword_(std::exchange(o.hword_, nullptr))
    {
    }
    DeviceFlag& operator=(DeviceFlag&& o) noexcept
    {
        if (this != &o) {
            release();
            dev_ = o.dev_;
            dchunk_ = std::exchange(o.dchunk_, nullptr);
            hchunk_ = std::exchange(o.hchunk_, nullptr);
            dword_ = std::exchange(o.dword_, nullptr);
            hword_ = std::exchange(o.hword_, nullptr);
        }
        return *this;
    }
    ~DeviceFlag() { release(); }

    /** The view a kernel takes by value. */
    DeviceFlagView deviceView() const { return DeviceFlagView(static_cast<std::uint32_t*>(dword_)); }
    /** Is the flag set? Waits for the device's work, reads the device word; does not clear it. */
    bool isSet() const
    {
        detail::check(detail::seam().stream_sync(dev_, nullptr), "DeviceFlag::isSet");
        detail::check(detail::seam().copy(AETHER_MEM_PINNED, dev_, hword_, AETHER_MEM_DEVICE, dev_, dword_,
                                          sizeof(std::uint32_t), nullptr),
                      "DeviceFlag::isSet");
        detail::check(detail::seam().stream_sync(dev_, nullptr), "DeviceFlag::isSet");
        std::uint32_t v = 0;
        std::memcpy(&v, hword_, sizeof v);
        return v != 0;
    }
    /** Reset the flag on the host and device copies. */
    void clear()
    {
        const std::uint32_t z = 0;
        std::memcpy(hword_, &z, sizeof z);
        detail::check(detail::seam().stream_sync(dev_, nullptr), "DeviceFlag::clear");
        detail::check(detail::seam().copy(AETHER_MEM_DEVICE, dev_, dword_, AETHER_MEM_PINNED, dev_, hword_,
                                          sizeof(std::uint32_t), nullptr),
                      "DeviceFlag::clear");
        detail::check(detail::seam().stream_sync(dev_, nullptr), "DeviceFlag::clear");
    }
    /** If the flag is set: clear it (so the caller may keep using it) and throw aether::Error naming
     *  `context`, shaped like err::fail (err/Error.h:63). */
    void checkAndClear(const std::string& context)
    {
        if (!isSet()) return;
        clear();
        throw Error(context + " failed on DeviceFlag, size=" + std::to_string(sizeof(std::uint32_t)) +
                    " bytes (device flag was set)");
    }

private:
    void release()
    {
        if (dchunk_) detail::seam().chunk_release(dchunk_);
        if (hchunk_) detail::seam().chunk_release(hchunk_);
        dchunk_ = hchunk_ = nullptr;
    }
    int dev_ = 0;
    void* dchunk_ = nullptr;
    void* hchunk_ = nullptr;
    void* dword_ = nullptr;
    void* hword_ = nullptr;
};

namespace accum {

inline constexpr std::uint32_t npos = 0xFFFFFFFFu; // accum/AppendCounter.h:26-29

enum class Policy { CompensatedAtomic = AETH_PLANE_COMPENSATED, Serialized = AETH_PLANE_SERIALIZED };

namespace detail {
inline void zero_device(void* p, std::size_t bytes, Stream s)
{
    if (!bytes) return;
    const std::uint32_t z = 0;
    aether::detail::check(aether::detail::seam().fill(AETHER_MEM_DEVICE, aether::detail::current_device(), p, bytes,
                                                      &z, sizeof z, s),
                          "fill");
}
} // namespace detail

/**
 * Plane<VD> (accum/AccumPlane.h:405-470): per-sample accumulator of VD lanes of double with a
 * companion residual lane. add() folds a term array into every sample with the compensated atomic
 * (concurrent producers on any streams) or a plain add (Serialized); delivered() writes value + comp.
 * Device kernels of a caller's own use aether::accum::PlaneView (include/aether/device/accum.hpp) on
 * valueView()/compView().
 */
template<std::size_t VD = 3, Policy P = Policy::CompensatedAtomic>
class Plane {
public:
    explicit Plane(std::size_t n) : value_(n), comp_(P == Policy::CompensatedAtomic ? n : 0) { zeroDeviceAsync(); }
    std::size_t samples() const { return value_.samples(); }
    View<double, VD> valueView() const { return value_.deviceView(); }
    View<double, VD> compView() const { return comp_.deviceView(); }
    /** Clear both lanes of the whole allocation, stream-ordered. */
    void zeroDeviceAsync(Stream s = nullptr)
    {
        detail::zero_device(value_.deviceView().data(), VD * value_.capacity() * sizeof(double), s);
        detail::zero_device(comp_.deviceView().data(), VD * comp_.capacity() * sizeof(double), s);
    }
    /** plane[i] += terms[i] for every sample, components in ascending order. */
    void add(const Array<double, VD>& terms, Stream s = nullptr)
    {
        if (terms.samples() != samples()) throw Error("AccumPlane::add: terms and plane differ in samples");
        aether::detail::check(aeth_plane_add((int)P, (int)VD, terms.deviceView().data(), terms.capacity(), samples(),
                                             value_.deviceView().data(),
                                             P == Policy::CompensatedAtomic ? comp_.deviceView().data() : nullptr,
                                             value_.capacity(), s),
                              "aeth_plane_add");
    }
    /** out = value + comp (one kernel). */
    void delivered(Array<double, VD>& out, Stream s = nullptr) const
    {
        if (out.samples() != samples()) throw Error("AccumPlane::delivered: output and plane differ in samples");
        if constexpr (P == Policy::CompensatedAtomic) out.deviceView().assign(valueView() + compView(), SumMode::None, nullptr, s);
        else out.deviceView().assign(valueView() * 1.0, SumMode::None, nullptr, s);
    }

private:
    Array<double, VD> value_;
    Array<double, VD> comp_;
};

/** A device slot counter over `capacity` places with a DeviceFlag for overflow (accum/AppendCounter.h:82-192). */
class AppendCounter {
public:
    explicit AppendCounter(std::size_t capacity) : capacity_(capacity)
    {
        const int dev = aether::detail::current_device();
        aether::detail::check(aether::detail::seam().chunk_alloc(AETHER_MEM_DEVICE, dev, 16, nullptr, &chunk_, &mem_),
                              "AppendCounter allocation");
        reset();
        aether::detail::check(aether::detail::seam().stream_sync(dev, nullptr), "AppendCounter");
    }
    AppendCounter(const AppendCounter&) = delete;
    AppendCounter& operator=(const AppendCounter&) = delete;
    ~AppendCounter()
    {
        if (chunk_) aether::detail::seam().chunk_release(chunk_);
    }
    std::size_t capacity() const { return capacity_; }
    std::uint32_t* counter() const { return static_cast<std::uint32_t*>(mem_); }
    std::uint32_t* overflowFlag() const { return overflow_.deviceView().data(); }
    const DeviceFlag& overflow() const { return overflow_; }
    /** The raw count (may exceed capacity: every overflowing claim still incremented it). */
    std::size_t rawCount() const { return read(0); }
    /** Non-destructive peek at the overflow flag. */
    bool overflowed() const { return overflow_.isSet(); }
    /** Zero the counter, stream-ordered (and the overflow flag when asked). */
    void reset(bool clearOverflow = false, Stream s = nullptr)
    {
        detail::zero_device(mem_, 16, s);
        if (clearOverflow) overflow_.clear();
    }
    /** Throw if the overflow flag is set, and clear it. */
    void checkOverflow(const std::string& context = "AppendCounter::checkOverflow")
    {
        overflow_.checkAndClear(context);
    }
    /** Adopt the claimed slots into `arr` (clamped to its spare capacity) and reset for the next step. */
    template<class T, std::size_t N>
    std::size_t commit(Array<T, N>& arr)
    {
        const std::size_t k = std::min(std::min(rawCount(), capacity_), arr.capacity() - arr.samples());
        arr.commit(k);
        reset();
        return k;
    }

private:
    std::size_t read(int word) const
    {
        std::uint32_t v = 0;
        const int dev = aether::detail::current_device();
        aether::detail::check(aether::detail::seam().copy(AETHER_MEM_CPU, dev, &v, AETHER_MEM_DEVICE, dev,
                                                          static_cast<std::uint32_t*>(mem_) + word, sizeof v, nullptr),
                              "AppendCounter read");
        aether::detail::check(aether::detail::seam().stream_sync(dev, nullptr), "AppendCounter read");
        return v;
    }
    std::size_t capacity_;
    void* chunk_ = nullptr;
    void* mem_ = nullptr;
    DeviceFlag overflow_;
};

/**
 * Append the index of every sample of `x` above `threshold` into the spare tail of `out` (its device
 * copy), claiming slots from `counter` (whose capacity must not exceed out's spare capacity); one
 * atomic per tile of 16384 samples, which keep their order. Commit with counter.commit(out).
 */
template<class T>
void appendWhere(View<T, 1> x, double threshold, AppendCounter& counter, Array<std::uint32_t>& out,
                 Stream s = nullptr)
{
    if (counter.capacity() > out.capacity() - out.samples())
        throw Error("appendWhere: the counter's capacity exceeds the array's spare capacity");
    aether::detail::check(aeth_append_where(aether::detail::dtype_code<T>(), x.data(), x.samples(), threshold,
                                            out.deviceSpare().data(), counter.capacity(), counter.counter(),
                                            counter.overflowFlag(), AETH_CLAIM_TILE, s),
                          "aeth_append_where");
}

/** run forms of keyedSum (accum/run.h:1-40). */
enum class Run { Aligned = AETH_RUN_ALIGNED, Tested = AETH_RUN_TESTED, Sorted = AETH_RUN_SORTED };

/**
 * target[keys[l]] += terms[l] for every lane with ONE atomic per run of a wavefront's lanes that
 * share a slot; with `comp`, each commit is compensated (delivered: target + comp). All pointers are
 * device memory.
 */
template<class T>
void keyedSum(Run form, const std::int64_t* keys, const T* terms, std::size_t n, T* target, std::size_t slots,
              T* comp = nullptr, int width = 64, std::int64_t own = 0, Stream s = nullptr)
{
    aether::detail::check(aeth_keyed_sum(aether::detail::dtype_code<T>(), (int)form, width, comp != nullptr ? 1 : 0,
                                         keys, terms, n, own, slots, target, comp, nullptr, s),
                          "aeth_keyed_sum");
}

} // namespace accum

template<std::size_t VD = 3, accum::Policy P = accum::Policy::CompensatedAtomic>
using AccumPlane = accum::Plane<VD, P>;

// ---------------------------------------------------------------- runtime views
inline constexpr std::size_t RuntimeRankCap = AETH_RUNTIME_RANK_CAP; // view/RuntimeView.h:71

/** RuntimeView (view/RuntimeView.h:77-120): untyped data, dtype, rank, extents, element strides. */
struct RuntimeView {
    aeth_runtime_view v{};
    std::size_t size() const
    {
        std::size_t t = 1;
        for (int k = 0; k < v.rank; ++k) t *= v.extents[k];
        return t;
    }
    int rank() const { return v.rank; }
};

/** A RuntimeView of one component plane of a View: rank 1 over the samples. */
template<class T, std::size_t N>
RuntimeView fromView(const View<T, N>& view, std::size_t component = 0)
{
    RuntimeView r;
    r.v.data = view.data() + component * view.pitch();
    r.v.dtype = detail::dtype_code<T>();
    r.v.rank = 1;
    r.v.extents[0] = (std::uint32_t)view.samples();
    r.v.strides[0] = 1;
    return r;
}

/** A RuntimeView over `shape` (row-major) with explicit element strides; strides empty = dense. */
template<class T>
RuntimeView makeRuntimeView(T* data, std::span<const std::size_t> shape, std::span<const std::int64_t> strides = {})
{
    if (shape.size() > RuntimeRankCap) throw Error("RuntimeView: rank exceeds " + std::to_string(RuntimeRankCap));
    RuntimeView r;
    r.v.data = const_cast<std::remove_const_t<T>*>(data);
    r.v.dtype = detail::dtype_code<std::remove_const_t<T>>();
    r.v.rank = (int)shape.size();
    std::int64_t dense = 1;
    for (std::size_t k = shape.size(); k-- > 0;) {
        r.v.extents[k] = (std::uint32_t)shape[k];
        r.v.strides[k] = strides.empty() ? dense : strides[k];
        dense *= (std::int64_t)shape[k];
    }
    return r;
}

namespace interop {
/**
 * fromDLPack (interop/DLPack.h:221-260) for any DLTensor-shaped struct (the DLPack field names: data,
 * device.device_type, ndim, dtype.{code,bits,lanes}, shape, strides, byte_offset): float32/float64,
 * one lane, zero byte_offset, device memory (ROCm or CUDA device types); strides NULL = dense.
 */
template<class DLT>
RuntimeView fromDLPack(const DLT& t)
{
    if (t.dtype.code != 2 || t.dtype.lanes != 1 || (t.dtype.bits != 32 && t.dtype.bits != 64))
        throw Error("aether::interop::fromDLPack: float32/float64 single-lane tensors only");
    if (t.byte_offset != 0) throw Error("aether::interop::fromDLPack: byte_offset != 0 is not supported");
    if (t.ndim < 0 || (std::size_t)t.ndim > RuntimeRankCap) throw Error("aether::interop::fromDLPack: rank exceeds the cap");
    const int dt = (int)t.device.device_type;
    if (dt != 10 && dt != 2 && dt != 13) // kDLROCM, kDLCUDA, kDLCUDAManaged
        throw Error("aether::interop::fromDLPack: not device memory");
    RuntimeView r;
    r.v.data = t.data;
    r.v.dtype = t.dtype.bits == 64 ? AETH_F64 : AETH_F32;
    r.v.rank = t.ndim;
    std::int64_t dense = 1;
    for (int k = t.ndim; k-- > 0;) {
        r.v.extents[k] = (std::uint32_t)t.shape[k];
        r.v.strides[k] = t.strides ? (std::int64_t)t.strides[k] : dense;
        dense *= (std::int64_t)t.shape[k];
    }
    return r;
}
} // namespace interop

namespace eval {

enum class RuntimeOp : int { Assign = AETH_RT_ASSIGN, Add = AETH_RT_ADD, Sub = AETH_RT_SUB, Scale = AETH_RT_SCALE,
                             AddScaled = AETH_RT_ADD_SCALED, SubScaled = AETH_RT_SUB_SCALED };

/** runtimeEval (eval/Runtime.h:250-300): validates shapes and dtypes, then one kernel on the device. */
inline void runtimeEval(RuntimeOp op, const RuntimeView& a, const RuntimeView& b, const RuntimeView& out,
                        double scalar = 0.0, Stream s = nullptr)
{
    aether::detail::check(aeth_runtime_eval((int)op, &a.v, &b.v, &out.v, scalar, s), "aether::eval::runtimeEval");
}

} // namespace eval

// ---------------------------------------------------------------- random
namespace random {

inline constexpr std::uint64_t DEFAULT_SEED = 0xFE7A5EED0C0FFEE5ULL; // random/free.h:44

/** A distribution leaf (random/detail/DistributionLeaf.h:111-236). */
template<class Real, std::size_t VD>
struct DistLeaf : Expr<DistLeaf<Real, VD>> {
    using element_type = Real;
    static constexpr std::size_t ncomp = VD;
    static constexpr bool lowerable = true;
    std::uint64_t seed; // salted
    double p0, p1;
    std::uint32_t dist;
    void emit(std::size_t c, detail::TapeCtx& x) const { x.rng(seed, p0, p1, (std::uint32_t)c, dist); }
    AETHER_DEVICEHOST() Real eval(std::size_t c, std::size_t i) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        // the tape kernels' draw: the same Philox block, conversions and device math
        const aeth::rng::U4 d = aeth::rng::block(seed, (std::uint32_t)i, c);
        if constexpr (std::is_same_v<Real, double>) {
            if (dist == AETH_DIST_UNIFORM) return p0 + (p1 - p0) * aeth::rng::u01_from_word_d(d.x);
            return p0 + p1 * aeth::rng::normal_from_block_d(d);
        } else {
            const float q0 = (float)p0, q1 = (float)p1;
            if (dist == AETH_DIST_UNIFORM) return q0 + (q1 - q0) * aeth::rng::u01_from_word_f(d.x);
            return q0 + q1 * aeth::rng::normal_from_block_f(d);
        }
#else
        const auto b = detail::philox_block(seed, (std::uint32_t)i, c);
        if (dist == AETH_DIST_UNIFORM) {
            if constexpr (std::is_same_v<Real, double>) return p0 + (p1 - p0) * ((double)b[0] * 0x1p-32 + 0x1p-33);
            else {
                const float u = (float)(((b[0] >> 9) << 1) | 1u) * 0x1p-24f;
                return (float)p0 + ((float)p1 - (float)p0) * u;
            }
        }
        if constexpr (std::is_same_v<Real, double>) {
            const std::uint64_t zx = (std::uint64_t)b[0] ^ ((std::uint64_t)b[1] << 21);
            const std::uint64_t zy = (std::uint64_t)b[2] ^ ((std::uint64_t)b[3] << 21);
            const double u = (double)zx * 0x1p-53 + 0x1p-54, v = (double)zy * (0x1p-53 * 2.0) + 0x1p-53;
            return p0 + p1 * (std::sqrt(-2.0 * std::log(u)) * std::sin(v * 3.1415926535897932));
        } else {
            const float k = 2.3283064e-10f * 6.2831855f;
            const float u = (float)(((b[0] >> 9) << 1) | 1u) * 0x1p-24f, v = (float)b[1] * k + (k / 2.0f);
            return (float)p0 + (float)p1 * (std::sqrt(-2.0f * std::log(u)) * std::sin(v));
        }
#endif
    }
};

/** random::Generator (random/Generator.h:56-195). */
class Generator {
public:
    constexpr explicit Generator(std::uint64_t seed, std::uint64_t stream_id = 0)
        : key_(detail::splitmix64(seed ^ detail::splitmix64(stream_id)))
    {
    }
    constexpr Generator substream(std::uint64_t k) const
    {
        Generator g(0);
        g.key_ = detail::splitmix64(key_ ^ detail::splitmix64(k ^ 0x2545F4914F6CDD1DULL));
        return g;
    }
    template<class Real, std::size_t VD>
    AETHER_DEVICEHOST() DistLeaf<Real, VD> uniform(Real lo, Real hi) const
    {
        DistLeaf<Real, VD> l;
        l.seed = key_ ^ 0x243F6A8885A308D3ULL;
        l.p0 = lo;
        l.p1 = hi;
        l.dist = AETH_DIST_UNIFORM;
        return l;
    }
    template<class Real, std::size_t VD>
    AETHER_DEVICEHOST() DistLeaf<Real, VD> normal(Real mean, Real sd) const
    {
        DistLeaf<Real, VD> l;
        l.seed = key_ ^ 0x13198A2E03707344ULL;
        l.p0 = mean;
        l.p1 = sd;
        l.dist = AETH_DIST_NORMAL;
        return l;
    }
    /** Per-component bounds and parameters (random/Generator.h:90-111): lo + (hi - lo) u and mean + sd z, the
     *  same draws as the scalar forms with the parameters of component c (a leaf on [0, 1) gives u exactly,
     *  one of mean 0 and sd 1 gives z exactly). */
    template<class Real, std::size_t VD>
    AETHER_DEVICEHOST() auto uniform(const Item<Real, VD>& lo, const Item<Real, VD>& hi) const
    {
        return lo + (hi - lo) * uniform<Real, VD>(Real(0), Real(1));
    }
    template<class Real, std::size_t VD>
    AETHER_DEVICEHOST() auto normal(const Item<Real, VD>& mean, const Item<Real, VD>& sd) const
    {
        return mean + sd * normal<Real, VD>(Real(0), Real(1));
    }
    /** random/detail/DistributionLeaf.h:240-300: exp(mu + sigma z), z the standard normal of the lognormal
     *  stream's own salt. */
    template<class Real, std::size_t VD>
    AETHER_DEVICEHOST() auto lognormal(Real mu, Real sigma) const
    {
        return cwiseExp(leaf<Real, VD>(0xA4093822299F31D0ULL, AETH_DIST_NORMAL, mu, sigma));
    }
    template<class Real, std::size_t VD>
    AETHER_DEVICEHOST() auto lognormal(const Item<Real, VD>& mu, const Item<Real, VD>& sigma) const
    {
        return cwiseExp(mu + sigma * leaf<Real, VD>(0xA4093822299F31D0ULL, AETH_DIST_NORMAL, Real(0), Real(1)));
    }
    /** DistributionLeaf.h:305-365: -log(u) / lambda, a division (not a product by 1/lambda). */
    template<class Real, std::size_t VD>
    AETHER_DEVICEHOST() auto exponential(Real lambda) const
    {
        return cwiseDiv(-cwiseLog(leaf<Real, VD>(0x082EFA98EC4E6C89ULL, AETH_DIST_UNIFORM, Real(0), Real(1))),
                        ScalarNode<Real>(lambda));
    }
    template<class Real, std::size_t VD>
    AETHER_DEVICEHOST() auto exponential(const Item<Real, VD>& lambda) const
    {
        return cwiseDiv(-cwiseLog(leaf<Real, VD>(0x082EFA98EC4E6C89ULL, AETH_DIST_UNIFORM, Real(0), Real(1))), lambda);
    }
    /** DistributionLeaf.h:370-430: 1 when u < p, else 0. The difference of two floats is 0 only when they are
     *  equal, so sign(p - u) is +1 exactly when u < p, and max(sign(p - u), 0) is the draw (0 for a NaN p). */
    template<class Real, std::size_t VD>
    AETHER_DEVICEHOST() auto bernoulli(Real p) const
    {
        return cwiseMax(
            cwiseSign(ScalarNode<Real>(p) - leaf<Real, VD>(0x452821E638D01377ULL, AETH_DIST_UNIFORM, Real(0), Real(1))),
            ScalarNode<Real>(Real(0)));
    }
    template<class Real, std::size_t VD>
    AETHER_DEVICEHOST() auto bernoulli(const Item<Real, VD>& p) const
    {
        return cwiseMax(cwiseSign(p - leaf<Real, VD>(0x452821E638D01377ULL, AETH_DIST_UNIFORM, Real(0), Real(1))),
                        ScalarNode<Real>(Real(0)));
    }
    constexpr std::uint64_t key() const { return key_; }

private:
    template<class Real, std::size_t VD>
    AETHER_DEVICEHOST() DistLeaf<Real, VD> leaf(std::uint64_t salt, std::uint32_t dist, Real p0, Real p1) const
    {
        DistLeaf<Real, VD> l;
        l.seed = key_ ^ salt;
        l.p0 = p0;
        l.p1 = p1;
        l.dist = dist;
        return l;
    }
    std::uint64_t key_;
};

inline Generator seed(std::uint64_t s) { return Generator(s); }

// random/free.h:46-79: the distributions of Generator(DEFAULT_SEED)
template<class Real, std::size_t VD>
AETHER_DEVICEHOST() auto uniform(Real lo, Real hi)
{
    return Generator(DEFAULT_SEED).uniform<Real, VD>(lo, hi);
}
template<class Real, std::size_t VD>
AETHER_DEVICEHOST() auto normal(Real mean, Real sd)
{
    return Generator(DEFAULT_SEED).normal<Real, VD>(mean, sd);
}
template<class Real, std::size_t VD>
AETHER_DEVICEHOST() auto lognormal(Real mu, Real sigma)
{
    return Generator(DEFAULT_SEED).lognormal<Real, VD>(mu, sigma);
}
template<class Real, std::size_t VD>
AETHER_DEVICEHOST() auto exponential(Real lambda)
{
    return Generator(DEFAULT_SEED).exponential<Real, VD>(lambda);
}
template<class Real, std::size_t VD>
AETHER_DEVICEHOST() auto bernoulli(Real p)
{
    return Generator(DEFAULT_SEED).bernoulli<Real, VD>(p);
}

} // namespace random

// ---------------------------------------------------------------- fft
namespace fft {

enum class Norm { Backward = AETHER_FFT_BACKWARD, Forward = AETHER_FFT_FORWARD, Ortho = AETHER_FFT_ORTHO };

namespace detail {

// One transform through the seam's plans. `in_values`/`out_values`: T values on each side. Host buffers are
// staged through device chunks; the call returns when the result is in `out`.
template<class T>
void run(int kind, const T* in, std::size_t in_values, T* out, std::size_t out_values,
         std::span<const std::size_t> shape, std::span<const std::size_t> axes, bool forward, Norm norm, Stream s)
{
    static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>);
    const auto& t = aether::detail::seam();
    const int dev = aether::detail::current_device();
    const std::vector<std::int64_t> sh(shape.begin(), shape.end());
    const std::vector<std::int32_t> ax(axes.begin(), axes.end());
    int pdev = -1;
    const bool host = t.pointer_device(in, &pdev) != 0;
    struct Staging {
        const AetherMemoryV1& t;
        void* chunk[2] = {nullptr, nullptr};
        void* ptr[2] = {nullptr, nullptr};
        void* plan = nullptr;
        ~Staging()
        {
            if (plan) t.fft_destroy(plan);
            for (void* c : chunk)
                if (c) t.chunk_release(c);
        }
    } st{t};
    const void* din = in;
    void* dout = out;
    if (host) {
        const std::size_t bytes[2] = {in_values * sizeof(T), out_values * sizeof(T)};
        for (int k = 0; k < 2; ++k)
            aether::detail::check(t.chunk_alloc(AETHER_MEM_DEVICE, dev, bytes[k] ? bytes[k] : sizeof(T), s,
                                                &st.chunk[k], &st.ptr[k]),
                                  "fft staging");
        aether::detail::check(t.copy(AETHER_MEM_DEVICE, dev, st.ptr[0], AETHER_MEM_CPU, dev, in, bytes[0], s),
                              "fft upload");
        din = st.ptr[0];
        dout = st.ptr[1];
    }
    aether::detail::check(t.fft_plan(dev, kind, std::is_same_v<T, double> ? 1 : 0, (int)norm, forward ? 1 : 0,
                                     (int)sh.size(), sh.data(), (int)ax.size(), ax.data(), s, &st.plan),
                          "fft_plan");
    aether::detail::check(t.fft_exec(st.plan, din, dout, s), "fft_exec");
    if (host)
        aether::detail::check(
            t.copy(AETHER_MEM_CPU, dev, out, AETHER_MEM_DEVICE, dev, st.ptr[1], out_values * sizeof(T), s),
            "fft download");
    aether::detail::check(t.stream_sync(dev, s), "stream_sync");
}

inline std::size_t count(std::span<const std::size_t> shape)
{
    std::size_t total = 1;
    for (auto v : shape) total *= v;
    return total;
}

inline std::size_t half_count(std::span<const std::size_t> shape, std::span<const std::size_t> axes)
{
    if (axes.empty() || axes.back() >= shape.size()) throw Error("aether::fft: bad axes");
    const std::size_t n = shape[axes.back()];
    return n ? count(shape) / n * (n / 2 + 1) : 0;
}

} // namespace detail

/**
 * Complex to complex over `axes` of a dense row-major array of `shape`
 * (fft/Fft.h:186-197). `in`/`out` hold prod(shape) interleaved (re, im)
 * pairs, in device memory or in host memory (staged through the device).
 */
template<class T>
void c2c(const T* in, T* out, std::span<const std::size_t> shape, std::span<const std::size_t> axes, bool forward,
         Norm norm = Norm::Backward, Stream s = nullptr)
{
    const std::size_t v = 2 * detail::count(shape);
    detail::run<T>(AETHER_FFT_C2C, in, v, out, v, shape, axes, forward, norm, s);
}

/**
 * Real to complex over `axes` (fft/Fft.h:198-213): `shape` is the real
 * input's; `out` holds that shape with the last axis of `axes` cut to
 * n/2 + 1, as (re, im) pairs.
 */
template<class T>
void r2c(const T* in, T* out, std::span<const std::size_t> shape, std::span<const std::size_t> axes,
         Norm norm = Norm::Backward, Stream s = nullptr)
{
    detail::run<T>(AETHER_FFT_R2C, in, detail::count(shape), out, 2 * detail::half_count(shape, axes), shape, axes,
                   true, norm, s);
}

/**
 * Complex to real over `axes` (fft/Fft.h:215-236): `shape` is the real
 * output's; `in` holds the half spectrum and is left unchanged.
 */
template<class T>
void c2r(const T* in, T* out, std::span<const std::size_t> shape, std::span<const std::size_t> axes,
         Norm norm = Norm::Backward, Stream s = nullptr)
{
    detail::run<T>(AETHER_FFT_C2R, in, 2 * detail::half_count(shape, axes), out, detail::count(shape), shape, axes,
                   false, norm, s);
}

} // namespace fft
} // namespace aether
