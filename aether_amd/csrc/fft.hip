// fft.hip -- batched power-of-two complex FFTs, LDS-resident, one launch.
//
// The reference's host FFT is pocketfft (fft/Fft.h:98-236) and its GPU FFT is
// cuFFT behind the plugin seam (dsc/src/cufft_plan.h:122-320). This is a
// from-scratch gfx950 design: every transform of up to 8192 points lives in
// LDS for its whole life -- read from HBM once, written once. A work
// descriptor names a chunk of signals of one length (as many as fit in 8192
// points); one workgroup of 1024 threads runs the Stockham autosort passes
// (radix 4, a final radix 2 when log2 N is odd) on all signals of its chunk
// at once, so small and large transforms are handled by the same launch and
// every pass has 2048 (radix 4) or 4096 (radix 2) butterflies in flight.
// LDS rows are padded by one element every 16 so the stride-4 writes of the
// first passes do not collide in the banks.
//
// Precisions: complex128; complex64; complex64 in memory with float-float
// (Ff2-style, banded/Ff2.h:132-145,246-321) arithmetic inside, the
// compensated FP32 mode; and complex64 in memory with FP64 arithmetic inside
// (FP64 is full rate on this GPU: the double transform's arithmetic at the
// float transform's memory traffic). Twiddles come from one FP64 table
// exp(-2 pi i m / 8192) generated on the device (float-float pairs for the
// compensated mode).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/aether_hip.h"
#include "accum.hpp"
#include "common.hpp"
#include "dmath.hpp"
#include "fft.hpp"

namespace aeth {
namespace fft {

constexpr int kMaxLog2 = 13;
constexpr int kMaxN = 1 << kMaxLog2;
constexpr int kThreads = 1024;

__host__ __device__ constexpr uint32_t pad(uint32_t e) { return e + (e >> 4); }
constexpr uint32_t kLdsElems = pad(kMaxN);

// ---- working types --------------------------------------------------------
struct C128 {
    double re, im;
};
struct C64 {
    float re, im;
};
// complex64 in memory, float-float arithmetic: (re_hi, re_lo, im_hi, im_lo)
struct CFF {
    float rh, rl, ih, il;
};
// complex64 in memory, FP64 arithmetic (the same arithmetic as C128, so equal inputs give equal bits)
struct CW {
    double re, im;
};

__device__ __forceinline__ C128 add(C128 a, C128 b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ C128 sub(C128 a, C128 b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ C64 add(C64 a, C64 b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ C64 sub(C64 a, C64 b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ CW add(CW a, CW b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ CW sub(CW a, CW b) { return {a.re - b.re, a.im - b.im}; }

// float-float sum for the butterflies: TwoSum of the high parts, the low parts added to its error (8 flops).
// The pair is left unnormalised: the high part is the rounded sum of the high parts and the low part carries
// everything else, so the pair's value is exact up to the rounding of the low part (2^-24 of it, below
// 2^-47 (|a| + |b|) while the low parts stay below 2^-23 of their high parts, which a butterfly network of
// at most 13 levels keeps). TwoSum is exact for any pair of floats, so later sums need no normalised input;
// the value is rounded once to float at the store. A transform's error is measured against the norm of its
// input, so cancellation in a + b costs nothing there (the accurate Ff2 add of accum.hpp, 23 flops, stays on
// the reductions).
__device__ __forceinline__ acc::Pair<float> ffadd(float ah, float al, float bh, float bl)
{
    const float s = ah + bh;
    const float bb = s - ah;
    const float e = ((ah - (s - bb)) + (bh - bb)) + (al + bl);
    return acc::Pair<float>{s, e};
}
// a - b: ffadd(a, -b) written without the negations (the same operations, so the same bits)
__device__ __forceinline__ acc::Pair<float> ffsub(float ah, float al, float bh, float bl)
{
    const float s = ah - bh;
    const float bb = s - ah;
    const float e = ((ah - (s - bb)) - (bh + bb)) + (al - bl);
    return acc::Pair<float>{s, e};
}
__device__ __forceinline__ CFF add(CFF a, CFF b)
{
    const auto r = ffadd(a.rh, a.rl, b.rh, b.rl), i = ffadd(a.ih, a.il, b.ih, b.il);
    return {r.h, r.c, i.h, i.c};
}
__device__ __forceinline__ CFF sub(CFF a, CFF b)
{
    const auto r = ffsub(a.rh, a.rl, b.rh, b.rl), i = ffsub(a.ih, a.il, b.ih, b.il);
    return {r.h, r.c, i.h, i.c};
}

// -i * x (forward) or +i * x (inverse): exact
template<class C>
__device__ __forceinline__ C rot(C x, bool fwd);
template<>
__device__ __forceinline__ C128 rot(C128 x, bool fwd) { return fwd ? C128{x.im, -x.re} : C128{-x.im, x.re}; }
template<>
__device__ __forceinline__ C64 rot(C64 x, bool fwd) { return fwd ? C64{x.im, -x.re} : C64{-x.im, x.re}; }
template<>
__device__ __forceinline__ CW rot(CW x, bool fwd) { return fwd ? CW{x.im, -x.re} : CW{-x.im, x.re}; }
template<>
__device__ __forceinline__ CFF rot(CFF x, bool fwd)
{
    return fwd ? CFF{x.ih, x.il, -x.rh, -x.rl} : CFF{-x.ih, -x.il, x.rh, x.rl};
}

// a + rot(t) and a - rot(t) (rot: -i t forward, +i t inverse)
template<class C>
__device__ __forceinline__ C addrot(C a, C t, bool fwd) { return add(a, rot(t, fwd)); }
template<class C>
__device__ __forceinline__ C subrot(C a, C t, bool fwd) { return sub(a, rot(t, fwd)); }
// the compensated type with the rotation folded into the sums (no negated pairs)
template<>
__device__ __forceinline__ CFF addrot(CFF a, CFF t, bool fwd)
{
    // forward: (a.re + t.im, a.im - t.re); inverse: (a.re - t.im, a.im + t.re)
    const auto r = fwd ? ffadd(a.rh, a.rl, t.ih, t.il) : ffsub(a.rh, a.rl, t.ih, t.il);
    const auto i = fwd ? ffsub(a.ih, a.il, t.rh, t.rl) : ffadd(a.ih, a.il, t.rh, t.rl);
    return {r.h, r.c, i.h, i.c};
}
template<>
__device__ __forceinline__ CFF subrot(CFF a, CFF t, bool fwd) { return addrot(a, t, !fwd); }

// float-float product of (ah + al)(bh + bl), 2Prod + cross terms, fast2Sum renormalised
__device__ __forceinline__ acc::Pair<float> ffmul(float ah, float al, float bh, float bl)
{
    const float p = ah * bh;
    float e = __builtin_fmaf(ah, bh, -p);
    e = __builtin_fmaf(ah, bl, e);
    e = __builtin_fmaf(al, bh, e);
    const float hi = p + e;
    return acc::Pair<float>{hi, e - (hi - p)};
}

// twiddle tables: W[m] = exp(-2 pi i m / 8192)
struct TwTables {
    const double2* d = nullptr;
    const float2* f = nullptr;
    const float4* ff = nullptr; // (re_hi, re_lo, im_hi, im_lo)
};

__device__ __forceinline__ C128 twiddle(const TwTables& t, uint32_t m, bool fwd, C128*)
{
    const double2 w = t.d[m];
    return {w.x, fwd ? w.y : -w.y};
}
__device__ __forceinline__ C64 twiddle(const TwTables& t, uint32_t m, bool fwd, C64*)
{
    const float2 w = t.f[m];
    return {w.x, fwd ? w.y : -w.y};
}
__device__ __forceinline__ CFF twiddle(const TwTables& t, uint32_t m, bool fwd, CFF*)
{
    const float4 w = t.ff[m];
    return fwd ? CFF{w.x, w.y, w.z, w.w} : CFF{w.x, w.y, -w.z, -w.w};
}
__device__ __forceinline__ CW twiddle(const TwTables& t, uint32_t m, bool fwd, CW*)
{
    const double2 w = t.d[m];
    return {w.x, fwd ? w.y : -w.y};
}

// complex products with one rounding less per part (fused multiply-add)
__device__ __forceinline__ C128 cmul(C128 a, C128 w)
{
    return {__builtin_fma(a.re, w.re, -(a.im * w.im)), __builtin_fma(a.re, w.im, a.im * w.re)};
}
__device__ __forceinline__ C64 cmul(C64 a, C64 w)
{
    return {__builtin_fmaf(a.re, w.re, -(a.im * w.im)), __builtin_fmaf(a.re, w.im, a.im * w.re)};
}
// float-float x*y + u*v as one dot product: two exact products (FMA residuals), TwoSum of their high
// parts, every low-order term (residuals, cross products) gathered into one float; left unnormalised like
// ffadd's pairs (16 flops)
__device__ __forceinline__ acc::Pair<float> ffdot2(float xh, float xl, float yh, float yl, float uh, float ul,
                                                   float vh, float vl)
{
    const float p = xh * yh, q = uh * vh;
    const float pe = __builtin_fmaf(xh, yh, -p), qe = __builtin_fmaf(uh, vh, -q);
    const float s = p + q;
    const float bb = s - p;
    float e = (p - (s - bb)) + (q - bb);
    e += pe + qe;
    e = __builtin_fmaf(xh, yl, e);
    e = __builtin_fmaf(xl, yh, e);
    e = __builtin_fmaf(uh, vl, e);
    e = __builtin_fmaf(ul, vh, e);
    return acc::Pair<float>{s, e};
}
__device__ __forceinline__ CFF cmul(CFF a, CFF w)
{
    const auto re = ffdot2(a.rh, a.rl, w.rh, w.rl, -a.ih, -a.il, w.ih, w.il);
    const auto im = ffdot2(a.rh, a.rl, w.ih, w.il, a.ih, a.il, w.rh, w.rl);
    return {re.h, re.c, im.h, im.c};
}
__device__ __forceinline__ CW cmul(CW a, CW w)
{
    return {__builtin_fma(a.re, w.re, -(a.im * w.im)), __builtin_fma(a.re, w.im, a.im * w.re)};
}

// global I/O: complex128 as double2, complex64 (plain, float-float and wide) as float2
__device__ __forceinline__ C128 gload(const void* p, uint64_t i, C128*)
{
    const double2 v = static_cast<const double2*>(p)[i];
    return {v.x, v.y};
}
__device__ __forceinline__ C64 gload(const void* p, uint64_t i, C64*)
{
    const float2 v = static_cast<const float2*>(p)[i];
    return {v.x, v.y};
}
__device__ __forceinline__ CFF gload(const void* p, uint64_t i, CFF*)
{
    const float2 v = static_cast<const float2*>(p)[i];
    return {v.x, 0.0f, v.y, 0.0f};
}
__device__ __forceinline__ CW gload(const void* p, uint64_t i, CW*)
{
    const float2 v = static_cast<const float2*>(p)[i];
    return {(double)v.x, (double)v.y};
}
__device__ __forceinline__ void gstore(void* p, uint64_t i, C128 v, double s)
{
    static_cast<double2*>(p)[i] = make_double2(v.re * s, v.im * s);
}
__device__ __forceinline__ void gstore(void* p, uint64_t i, C64 v, double s)
{
    const float fs = (float)s;
    static_cast<float2*>(p)[i] = make_float2(v.re * fs, v.im * fs);
}
__device__ __forceinline__ void gstore(void* p, uint64_t i, CFF v, double s)
{
    // scale in float-float (s as a float-float pair), then round once to float
    const float sh = (float)s, sl = (float)(s - (double)sh);
    const auto r = ffmul(v.rh, v.rl, sh, sl), im = ffmul(v.ih, v.il, sh, sl);
    static_cast<float2*>(p)[i] = make_float2(r.h + r.c, im.h + im.c);
}
__device__ __forceinline__ void gstore(void* p, uint64_t i, CW v, double s)
{
    // scale in FP64, then round once to float
    static_cast<float2*>(p)[i] = make_float2((float)(v.re * s), (float)(v.im * s));
}

// One Stockham pass of radix R over all signals of the chunk (P points).
template<int R, class C>
__device__ __forceinline__ void pass(C* buf, uint32_t log2n, uint32_t P, uint32_t Ns, const TwTables& tw, bool fwd)
{
    constexpr int MAXB = kMaxN / (R * kThreads);
    const uint32_t N = 1u << log2n;
    const uint32_t NR = N / R;
    const uint32_t nb = P / R;
    C v[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const uint32_t b = threadIdx.x + q * kThreads;
        if (b < nb) {
            const uint32_t s = b / NR, j = b - s * NR;
            const uint32_t base = s * N;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = buf[pad(base + j + r * NR)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const uint32_t b = threadIdx.x + q * kThreads;
        if (b < nb) {
            const uint32_t s = b / NR, j = b - s * NR;
            const uint32_t base = s * N;
            const uint32_t k = j & (Ns - 1);
            const uint32_t m = k * (kMaxN / (R * Ns));
            C* x = v[q];
#pragma unroll
            for (int r = 1; r < R; ++r) x[r] = cmul(x[r], twiddle(tw, (r * m) & (kMaxN - 1), fwd, (C*)nullptr));
            C y[R];
            if constexpr (R == 4) {
                const C s02 = add(x[0], x[2]), d02 = sub(x[0], x[2]);
                const C s13 = add(x[1], x[3]), d13 = rot(sub(x[1], x[3]), fwd);
                y[0] = add(s02, s13);
                y[1] = add(d02, d13);
                y[2] = sub(s02, s13);
                y[3] = sub(d02, d13);
            } else {
                y[0] = add(x[0], x[1]);
                y[1] = sub(x[0], x[1]);
            }
            const uint32_t idx = base + (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[pad(idx + r * Ns)] = y[r];
        }
    }
    __syncthreads();
}

template<class C>
__global__ __launch_bounds__(kThreads) void fft_kernel(const Desc* __restrict__ descs, const Sig* __restrict__ sigs,
                                                       const void* __restrict__ in, void* __restrict__ out,
                                                       TwTables tw, int fwd_flag, int norm, double fixed_scale)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C* buf = reinterpret_cast<C*>(lds_raw);
    const bool fwd = fwd_flag != 0;
    const Desc d = descs[blockIdx.x];
    const uint32_t log2n = d.log2n, N = 1u << log2n, P = (uint32_t)d.count << log2n;
    // numpy norms (fft/Fft.h:73-76): Ortho 1/sqrt(N); Backward scales the inverse, Forward the forward
    // norm 3: a caller-fixed scale (multi-axis plans scale once, by the product of the lengths)
    const double scale = norm == 3 ? fixed_scale
                         : norm == 2 ? 1.0 / __builtin_sqrt((double)N)
                                     : ((fwd == (norm == 1)) ? 1.0 / (double)N : 1.0);
    const uint32_t stride = d.stride;
    if (stride == 1) {
        for (uint32_t e = threadIdx.x; e < P; e += kThreads) {
            const uint32_t s = e >> log2n, j = e & (N - 1);
            buf[pad(e)] = gload(in, sigs[d.first_sig + s].in_off + j, (C*)nullptr);
        }
    } else {
        // strided lines: neighbouring lanes take neighbouring signals at the same point
        const uint32_t cnt = d.count;
        for (uint32_t e = threadIdx.x; e < P; e += kThreads) {
            const uint32_t j = e / cnt, s = e - j * cnt;
            buf[pad(s * N + j)] = gload(in, sigs[d.first_sig + s].in_off + (uint64_t)j * stride, (C*)nullptr);
        }
    }
    __syncthreads();
    uint32_t Ns = 1;
    for (; Ns * 4 <= N; Ns *= 4) pass<4>(buf, log2n, P, Ns, tw, fwd);
    if (Ns < N) pass<2>(buf, log2n, P, Ns, tw, fwd);
    if (stride == 1) {
        for (uint32_t e = threadIdx.x; e < P; e += kThreads) {
            const uint32_t s = e >> log2n, j = e & (N - 1);
            gstore(out, sigs[d.first_sig + s].out_off + j, buf[pad(e)], scale);
        }
    } else {
        const uint32_t cnt = d.count;
        for (uint32_t e = threadIdx.x; e < P; e += kThreads) {
            const uint32_t j = e / cnt, s = e - j * cnt;
            gstore(out, sigs[d.first_sig + s].out_off + (uint64_t)j * stride, buf[pad(s * N + j)], scale);
        }
    }
}

// ---- register-resident transforms (complex128 and complex64 on contiguous signals) -------------------
// One workgroup per chunk (up to 8192 points of one length); 512 threads hold 16 points each. A transform
// of 2^L points is L/4 Stockham passes of radix 16, then one of radix 2, 4 or 8 when 4 does not divide L;
// each pass is a DFT in registers (radix 16 as 4 x 4, radix 8 as 2 x 4, constant twiddles inside). The
// first pass reads its inputs straight from HBM and the last writes its outputs straight to HBM, both
// coalesced (consecutive lanes take consecutive butterflies), so a point crosses LDS only between passes,
// as one word (16 bytes for complex128 and the compensated types, 8 for complex64) at a swizzled place
// (xpos). The twiddles of a pass, W^(r m) for r < R, are products of the table entries
// W^m, W^2m, W^4m and W^8m (at most four factors), formed where they are used: a radix-16 butterfly loads
// four entries instead of fifteen. Complex products use fused multiply-adds.
constexpr int kRT = 512; // threads
constexpr int kRP = 16;  // points per thread

template<class C>
struct Words;
// The FP64 types cross LDS as two 8-byte words (real parts, then imaginary parts): half the LDS of whole
// 16-byte points, so two workgroups of 8192 points (or four of 4096) share a CU.
template<>
struct Words<C128> {
    static constexpr int n = 2;
    using W = double;
    __device__ static W get(const C128& v, int w) { return w == 0 ? v.re : v.im; }
    __device__ static void put(C128& v, int w, W x)
    {
        if (w == 0) v.re = x;
        else v.im = x;
    }
};
// the compensated type as two 8-byte words: the real pair, then the imaginary pair
template<>
struct Words<CFF> {
    static constexpr int n = 2;
    using W = float2;
    __device__ static W get(const CFF& v, int w) { return w == 0 ? make_float2(v.rh, v.rl) : make_float2(v.ih, v.il); }
    __device__ static void put(CFF& v, int w, W x)
    {
        if (w == 0) {
            v.rh = x.x;
            v.rl = x.y;
        } else {
            v.ih = x.x;
            v.il = x.y;
        }
    }
};
template<>
struct Words<CW> {
    static constexpr int n = 2;
    using W = double;
    __device__ static W get(const CW& v, int w) { return w == 0 ? v.re : v.im; }
    __device__ static void put(CW& v, int w, W x)
    {
        if (w == 0) v.re = x;
        else v.im = x;
    }
};
template<>
struct Words<C64> {
    static constexpr int n = 1;
    using W = float2;
    __device__ static W get(const C64& v, int) { return make_float2(v.re, v.im); }
    __device__ static void put(C64& v, int, W w) { v = C64{w.x, w.y}; }
};

// cos(2 pi k / 16)
__device__ constexpr double kCos16[16] = {1.0,
                                          0.92387953251128675613,
                                          0.70710678118654752440,
                                          0.38268343236508977173,
                                          0.0,
                                          -0.38268343236508977173,
                                          -0.70710678118654752440,
                                          -0.92387953251128675613,
                                          -1.0,
                                          -0.92387953251128675613,
                                          -0.70710678118654752440,
                                          -0.38268343236508977173,
                                          0.0,
                                          0.38268343236508977173,
                                          0.70710678118654752440,
                                          0.92387953251128675613};

// x * W16^E (forward W16 = exp(-2 pi i / 16); the inverse conjugates); E is a compile-time constant
template<int E, class C>
__device__ __forceinline__ C w16(C x, bool fwd)
{
    if constexpr (E % 16 == 0) {
        return x;
    } else if constexpr (E % 16 == 4) {
        return rot(x, fwd);
    } else if constexpr (std::is_same_v<C, CFF> && E % 2 == 0) {
        // an odd multiple of 45 degrees: W = c (+-1 +- i), c = sqrt(2)/2 as a float-float pair, so the
        // product is two float-float sums and two products by c (24 flops instead of a complex product's 32)
        constexpr double cd = 0.70710678118654752440;
        const float ch = (float)cd, cl = (float)(cd - (double)ch);
        auto byc = [&](acc::Pair<float> a) {
            const float p = a.h * ch;
            float e = __builtin_fmaf(a.h, ch, -p);
            e = __builtin_fmaf(a.h, cl, e);
            e = __builtin_fmaf(a.c, ch, e);
            return acc::Pair<float>{p, e};
        };
        // forward W16^t = c ((1 - i), (-1 - i), (-1 + i), (1 + i)) for t = 2, 6, 10, 14; the inverse is the
        // forward form of 16 - t
        constexpr int tf = E % 16, ti = (16 - E % 16) % 16;
        const int t = fwd ? tf : ti;
        const auto sum = ffadd(x.rh, x.rl, x.ih, x.il); // xr + xi
        const auto dif = ffsub(x.ih, x.il, x.rh, x.rl); // xi - xr
        acc::Pair<float> re, im;
        if (t == 2) { re = sum; im = dif; }
        else if (t == 6) { re = dif; im = acc::Pair<float>{-sum.h, -sum.c}; }
        else if (t == 10) { re = acc::Pair<float>{-sum.h, -sum.c}; im = acc::Pair<float>{-dif.h, -dif.c}; }
        else { re = acc::Pair<float>{-dif.h, -dif.c}; im = sum; }
        const auto r = byc(re), i = byc(im);
        return CFF{r.h, r.c, i.h, i.c};
    } else if constexpr (std::is_same_v<C, CFF>) {
        // the constant as a float-float pair (hi, lo) from its double value
        const double cd = kCos16[E % 16], sd = kCos16[(E + 12) % 16];
        const float ch = (float)cd, cl = (float)(cd - (double)ch), sh = (float)sd, sl = (float)(sd - (double)sh);
        return cmul(x, fwd ? CFF{ch, cl, -sh, -sl} : CFF{ch, cl, sh, sl});
    } else {
        using R = decltype(x.re);
        const R c = (R)kCos16[E % 16], s = (R)kCos16[(E + 12) % 16]; // sin(t) = cos(t - pi/2)
        return cmul(x, C{c, fwd ? -s : s});
    }
}

template<class C>
__device__ __forceinline__ void dft4(C& a, C& b, C& c, C& d, bool fwd)
{
    const C s02 = add(a, c), d02 = sub(a, c);
    const C s13 = add(b, d), t13 = sub(b, d);
    a = add(s02, s13);
    b = addrot(d02, t13, fwd);
    c = sub(s02, s13);
    d = subrot(d02, t13, fwd);
}

// In-place DFT of R points, natural order in and out. N = A B: y[c + B d] = DFT_A over a of
// W_N^(a c) DFT_B over b of x[a + A b].
template<int R, class C>
__device__ __forceinline__ void dft(C* x, bool fwd)
{
    if constexpr (R == 2) {
        const C t = x[0];
        x[0] = add(t, x[1]);
        x[1] = sub(t, x[1]);
    } else if constexpr (R == 4) {
        dft4(x[0], x[1], x[2], x[3], fwd);
    } else if constexpr (R == 8) {
        C z0[4] = {x[0], x[2], x[4], x[6]}, z1[4] = {x[1], x[3], x[5], x[7]};
        dft4(z0[0], z0[1], z0[2], z0[3], fwd);
        dft4(z1[0], z1[1], z1[2], z1[3], fwd);
        z1[1] = w16<2>(z1[1], fwd);
        z1[2] = w16<4>(z1[2], fwd);
        z1[3] = w16<6>(z1[3], fwd);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            x[c] = add(z0[c], z1[c]);
            x[c + 4] = sub(z0[c], z1[c]);
        }
    } else {
        static_assert(R == 16, "radix 2, 4, 8 or 16");
#pragma unroll
        for (int a = 0; a < 4; ++a) dft4(x[a], x[a + 4], x[a + 8], x[a + 12], fwd);
        // z[a][c] sits in x[a + 4c]; times W16^(a c)
        x[5] = w16<1>(x[5], fwd);
        x[9] = w16<2>(x[9], fwd);
        x[13] = w16<3>(x[13], fwd);
        x[6] = w16<2>(x[6], fwd);
        x[10] = w16<4>(x[10], fwd);
        x[14] = w16<6>(x[14], fwd);
        x[7] = w16<3>(x[7], fwd);
        x[11] = w16<6>(x[11], fwd);
        x[15] = w16<9>(x[15], fwd);
        // y[c + 4 d] = DFT4 over a of z[a][c]: the four points of column c are x[4c .. 4c+3]
        C y[16];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            C p0 = x[4 * c], p1 = x[4 * c + 1], p2 = x[4 * c + 2], p3 = x[4 * c + 3];
            dft4(p0, p1, p2, p3, fwd);
            y[c] = p0;
            y[c + 4] = p1;
            y[c + 8] = p2;
            y[c + 12] = p3;
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = y[r];
    }
}

// The twiddle source: the device table W^m = exp(-2 pi i m / 8192) (conjugated for the inverse).
struct TwGlobal {
    TwTables t;
    template<class C>
    __device__ __forceinline__ C at(uint32_t m, bool fwd) const
    {
        return twiddle(t, m, fwd, (C*)nullptr);
    }
};

// The base twiddles of a radix-R butterfly: W^m, W^2m, W^4m, W^8m (as many as R needs)
template<int R, class C>
struct TwBase {
    static constexpr int n = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
    C w[n];
};

template<int R, class C, class TW>
__device__ __forceinline__ TwBase<R, C> twiddle_fetch(uint32_t m, const TW& q, bool fwd)
{
    TwBase<R, C> t;
#pragma unroll
    for (int e = 0; e < TwBase<R, C>::n; ++e) t.w[e] = q.template at<C>(m << e, fwd);
    return t;
}

// x[r] *= W^(r m) for r = 1..R-1, from the base twiddles; each power is formed where it is used
template<int R, class C>
__device__ __forceinline__ void twiddle_apply(C* x, const TwBase<R, C>& t)
{
    const C w1 = t.w[0];
    x[1] = cmul(x[1], w1);
    if constexpr (R > 2) {
        const C w2 = t.w[1];
        x[2] = cmul(x[2], w2);
        const C w3 = cmul(w1, w2);
        x[3] = cmul(x[3], w3);
        if constexpr (R > 4) {
            const C w4 = t.w[2];
            x[4] = cmul(x[4], w4);
            const C w5 = cmul(w1, w4), w6 = cmul(w2, w4), w7 = cmul(w3, w4);
            x[5] = cmul(x[5], w5);
            x[6] = cmul(x[6], w6);
            x[7] = cmul(x[7], w7);
            if constexpr (R > 8) {
                const C w8 = t.w[3];
                x[8] = cmul(x[8], w8);
                x[9] = cmul(x[9], cmul(w1, w8));
                x[10] = cmul(x[10], cmul(w2, w8));
                x[11] = cmul(x[11], cmul(w3, w8));
                x[12] = cmul(x[12], cmul(w4, w8));
                x[13] = cmul(x[13], cmul(w5, w8));
                x[14] = cmul(x[14], cmul(w6, w8));
                x[15] = cmul(x[15], cmul(w7, w8));
            }
        }
    }
}

// The compensated complex64 type runs on the vector ALU's full issue rate, so its products of base twiddles
// (11 float-float complex products a radix-16 butterfly) cost more than loading every power from the table
// (which is also exact to the table's rounding): it reads W^(r m) for r = 1..R-1 directly.
template<class C>
constexpr bool kDirectTw = std::is_same_v<C, CFF>;

template<int R, class C, class TW>
__device__ __forceinline__ void twiddle_direct(C* x, uint32_t m, const TW& q, bool fwd)
{
#pragma unroll
    for (int r = 1; r < R; ++r) x[r] = cmul(x[r], q.template at<C>(r * m, fwd));
}

template<int R, class C, class TW>
__device__ __forceinline__ void twiddle_run(C* x, uint32_t m, const TW& q, bool fwd)
{
    if constexpr (kDirectTw<C>) twiddle_direct<R>(x, m, q, fwd);
    else twiddle_apply<R>(x, twiddle_fetch<R, C>(m, q, fwd));
}

// One pass of radix R (stride Ns = product of the earlier radices) over the thread's 16 / R butterflies
// b = threadIdx.x + u * kRT, b < nb; butterfly u holds v[u R .. u R + R - 1]; 2^l2nr butterflies a signal.
template<int T, int R, class C, class TW>
__device__ __forceinline__ void rpass(C (&v)[kRP], uint32_t l2nr, uint32_t nb, uint32_t Ns, const TW& q, bool fwd)
{
    constexpr int U = kRP / R;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t b = threadIdx.x + u * T;
        if (b < nb) {
            if (Ns > 1) {
                const uint32_t j = b & ((1u << l2nr) - 1u);
                const uint32_t k = j & (Ns - 1u);
                twiddle_run<R>(&v[u * R], k * (kMaxN / (R * Ns)), q, fwd);
            }
            dft<R>(&v[u * R], fwd);
        }
    }
}

// The twiddles of this thread's radix-16 butterfly (one a thread) in the pass of stride Ns, fetched ahead of
// the exchange before that pass so the loads' latency passes while the exchange runs.
template<class C, class TW>
__device__ __forceinline__ TwBase<16, C> rfetch16(uint32_t l2nr, uint32_t nb, uint32_t Ns, const TW& q, bool fwd)
{
    static_assert(kRP == 16, "one radix-16 butterfly a thread");
    TwBase<16, C> t{};
    const uint32_t b = threadIdx.x;
    if (b < nb) {
        const uint32_t k = (b & ((1u << l2nr) - 1u)) & (Ns - 1u);
        t = twiddle_fetch<16, C>(k * (kMaxN / (16 * Ns)), q, fwd);
    }
    return t;
}

template<class C>
__device__ __forceinline__ void rpass16(C (&v)[kRP], uint32_t nb, const TwBase<16, C>& t, bool fwd)
{
    if (threadIdx.x < nb) {
        twiddle_apply<16>(v, t);
        dft<16>(v, fwd);
    }
}

// The first pass's inputs: butterfly (s, j) reads x[j + r N / R] of signal s, for r in [R0, R1).
template<int T, int R, class C, int R0 = 0, int R1 = R, int NV>
__device__ __forceinline__ void rload(C (&v)[NV], const Sig* __restrict__ sigs, const void* __restrict__ in,
                                      uint32_t l2nr, uint32_t nb)
{
    constexpr int U = kRP / R;
    static_assert(U == 1 || (R0 == 0 && R1 == R), "a point range is for radix-16 loads");
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t b = threadIdx.x + u * T;
        if (b < nb) {
            const uint32_t s = b >> l2nr, j = b & ((1u << l2nr) - 1u);
            const uint64_t base = sigs[s].in_off + j;
#pragma unroll
            for (int r = R0; r < R1; ++r) v[u * R + r - R0] = gload(in, base + ((uint64_t)r << l2nr), (C*)nullptr);
        }
    }
}

// The last pass's outputs (Ns = N / R): butterfly (s, j) writes X[j + r Ns] of signal s.
template<int T, int R, class C>
__device__ __forceinline__ void rstore(const C (&v)[kRP], const Sig* __restrict__ sigs, void* __restrict__ out,
                                       uint32_t l2nr, uint32_t nb, double scale)
{
    constexpr int U = kRP / R;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t b = threadIdx.x + u * T;
        if (b < nb) {
            const uint32_t s = b >> l2nr, j = b & ((1u << l2nr) - 1u);
            const uint64_t base = sigs[s].out_off + j;
#pragma unroll
            for (int r = 0; r < R; ++r) gstore(out, base + ((uint64_t)r << l2nr), v[u * R + r], scale);
        }
    }
}

// LDS place of exchange element e: e with its low bits XORed by the bits above them, so that the writes of
// the first pass (stride 16 across lanes) spread over all banks while consecutive elements (every read,
// and the writes of the later passes) stay a permutation inside aligned groups. Measured with ds_*_b128
// on this GPU: a stride of 16 words conflicts on every access and a plain padding (e + e/16) conflicts on
// reads of consecutive words, while this placement conflicts on neither. 8-byte words swizzle four bits.
template<class W>
__device__ __forceinline__ uint32_t xpos(uint32_t e)
{
    return e ^ ((e >> 4) & (sizeof(W) == 16 ? 7u : 15u));
}

// Pass (R, Ns) outputs to LDS at their Stockham places, then the next pass's (R2) inputs back, one word
// of every point at a time.
template<int T, int R, int R2, class C>
__device__ __forceinline__ void rexchange(C (&v)[kRP], typename Words<C>::W* lds, uint32_t log2n, uint32_t P,
                                          uint32_t Ns)
{
    using Word = typename Words<C>::W;
    constexpr int U = kRP / R, U2 = kRP / R2;
    const uint32_t N = 1u << log2n;
    const uint32_t nr = N / R, nr2 = N / R2, nb = P / R, nb2 = P / R2;
#pragma unroll
    for (int w = 0; w < Words<C>::n; ++w) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t b = threadIdx.x + u * T;
            if (b < nb) {
                const uint32_t s = b / nr, j = b - s * nr;
                const uint32_t k = j & (Ns - 1u);
                const uint32_t idx = s * N + (j - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r) lds[xpos<Word>(idx + r * Ns)] = Words<C>::get(v[u * R + r], w);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U2; ++u) {
            const uint32_t b = threadIdx.x + u * T;
            if (b < nb2) {
                const uint32_t s = b / nr2, j = b - s * nr2;
                const uint32_t base = s * N + j;
#pragma unroll
                for (int r = 0; r < R2; ++r) Words<C>::put(v[u * R2 + r], w, lds[xpos<Word>(base + r * nr2)]);
            }
        }
        __syncthreads();
    }
}

template<int T, int R, class C, class TW>
__device__ __forceinline__ void rsingle(C (&v)[kRP], const Sig* sigs, const void* in, void* out, uint32_t P,
                                        double scale, const TW& q, bool fwd)
{
    rload<T, R>(v, sigs, in, 0, P / R);
    rpass<T, R>(v, 0, P / R, 1, q, fwd);
    rstore<T, R>(v, sigs, out, 0, P / R, scale);
}

// The tail pass of radix R (stride Ns = N / R) after the passes of 16. Butterfly u of a thread sits at
// j = (threadIdx.x + u kRT) mod Ns. When Ns <= kRT every butterfly of the thread has the same j, so one set of
// base twiddles serves all of them. When R Ns = kMaxN (one whole 8192-point signal) j = threadIdx.x + u kRT
// and W^(2^e j) = W^(2^e threadIdx.x) W16^(2^e u), a product by a constant. The base set (of j =
// threadIdx.x mod Ns) is fetched by the caller, ahead of the exchange before this pass.
template<int R, class C, class TW>
__device__ __forceinline__ TwBase<R, C> rfetch_tail(uint32_t Ns, const TW& q, bool fwd)
{
    if constexpr (kDirectTw<C>) return TwBase<R, C>{}; // read in the pass (rtail)
    else return twiddle_fetch<R, C>((threadIdx.x & (Ns - 1u)) * (kMaxN / (R * Ns)), q, fwd);
}

template<int R, class C, int u>
__device__ __forceinline__ TwBase<R, C> tw_shift(const TwBase<R, C>& t0, bool fwd)
{
    TwBase<R, C> t;
    t.w[0] = w16<u>(t0.w[0], fwd);
    if constexpr (TwBase<R, C>::n > 1) t.w[1] = w16<2 * u>(t0.w[1], fwd);
    if constexpr (TwBase<R, C>::n > 2) t.w[2] = w16<4 * u>(t0.w[2], fwd);
    if constexpr (TwBase<R, C>::n > 3) t.w[3] = w16<8 * u>(t0.w[3], fwd);
    return t;
}

template<int R, class C, int u = 0>
__device__ __forceinline__ void rtail_whole(C (&v)[kRP], const TwBase<R, C>& t0, bool fwd)
{
    if constexpr (u < kRP / R) {
        twiddle_apply<R>(&v[u * R], tw_shift<R, C, u>(t0, fwd));
        dft<R>(&v[u * R], fwd);
        rtail_whole<R, C, u + 1>(v, t0, fwd);
    }
}

template<int T, int R, class C, class TW>
__device__ __forceinline__ void rtail(C (&v)[kRP], const Sig* sigs, void* out, uint32_t log2n, uint32_t P,
                                      uint32_t Ns, double scale, const TwBase<R, C>& t0, const TW& q, bool fwd)
{
    constexpr uint32_t lr = R == 2 ? 1 : R == 4 ? 2 : 3;
    constexpr int U = kRP / R;
    const uint32_t nb = P / R;
    if constexpr (kDirectTw<C>) {
        // butterfly u sits at j = (threadIdx.x + u T) mod Ns (T is a multiple of Ns, or Ns = 4096 > T)
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t b = threadIdx.x + u * T;
            if (b < nb) {
                twiddle_direct<R>(&v[u * R], (b & (Ns - 1u)) * (kMaxN / (R * Ns)), q, fwd);
                dft<R>(&v[u * R], fwd);
            }
        }
    } else if (Ns <= (uint32_t)T) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (threadIdx.x + u * T < nb) {
                twiddle_apply<R>(&v[u * R], t0);
                dft<R>(&v[u * R], fwd);
            }
        }
    } else if constexpr (T == kRT) {
        // Ns > T happens only for log2n = 13 in the 512-thread form (R = 2, Ns = 4096): R Ns = kMaxN and
        // nb = U T
        rtail_whole<R>(v, t0, fwd);
    }
    rstore<T, R>(v, sigs, out, log2n - lr, nb, scale);
}

// Transform one chunk of signals of 2^L points: read, passes, write. Called with a constant L (rchunk), every
// shift, mask, pass count and stride folds and the passes unroll, so a pass's outputs stay in the registers
// the next pass reads (no moves between passes; half the registers of the form with L read at run time).
// LC < 0: the length is read from the descriptor at run time.
template<int T, int LC, class C, class TW>
__device__ __forceinline__ void rchunk_n(C (&v)[kRP], const Desc d, const Sig* s, const void* in, void* out,
                                         typename Words<C>::W* xch, const TW& q, bool fwd, int norm,
                                         double fixed_scale)
{
    const uint32_t L = LC >= 0 ? (uint32_t)LC : (uint32_t)d.log2n;
    const uint32_t N = 1u << L;
    const uint32_t P = (uint32_t)d.count << L;
    // numpy norms (fft/Fft.h:73-76); norm 3: a caller-fixed scale
    const double scale = norm == 3 ? fixed_scale
                         : norm == 2 ? 1.0 / __builtin_sqrt((double)N)
                                     : ((fwd == (norm == 1)) ? 1.0 / (double)N : 1.0);
    if (L == 0) {
        for (uint32_t e = threadIdx.x; e < P; e += T) gstore(out, s[e].out_off, gload(in, s[e].in_off, (C*)nullptr), scale);
        return;
    }
    const uint32_t nq = L / 4, rem = L % 4;
    if (nq == 0) {
        if (rem == 1) rsingle<T, 2>(v, s, in, out, P, scale, q, fwd);
        else if (rem == 2) rsingle<T, 4>(v, s, in, out, P, scale, q, fwd);
        else rsingle<T, 8>(v, s, in, out, P, scale, q, fwd);
        return;
    }
    rload<T, 16>(v, s, in, L - 4, P / 16);
    uint32_t Ns = 1;
    TwBase<16, C> next{};
    // complex128 fetches each radix-16 pass's twiddles ahead of the exchange before it; complex64 keeps its
    // registers for two workgroups a CU and fetches them in the pass
    constexpr bool ahead = sizeof(C) == 16 && !kDirectTw<C>;
    // at most three radix-16 passes (kMaxLog2 = 13): a constant trip count unrolls for every length
#pragma unroll
    for (uint32_t p = 0; p < (uint32_t)kMaxLog2 / 4; ++p) {
        if (p >= nq) break;
        if (p == 0 || !ahead) rpass<T, 16>(v, L - 4, P / 16, Ns, q, fwd);
        else rpass16(v, P / 16, next, fwd);
        if (p + 1 == nq) break;
        if constexpr (ahead) next = rfetch16<C>(L - 4, P / 16, Ns * 16, q, fwd);
        rexchange<T, 16, 16>(v, xch, L, P, Ns);
        Ns *= 16;
    }
    if (rem == 0) {
        rstore<T, 16>(v, s, out, L - 4, P / 16, scale);
    } else if (rem == 1) {
        const TwBase<2, C> t0 = rfetch_tail<2, C>(Ns * 16, q, fwd);
        rexchange<T, 16, 2>(v, xch, L, P, Ns);
        rtail<T, 2>(v, s, out, L, P, Ns * 16, scale, t0, q, fwd);
    } else if (rem == 2) {
        const TwBase<4, C> t0 = rfetch_tail<4, C>(Ns * 16, q, fwd);
        rexchange<T, 16, 4>(v, xch, L, P, Ns);
        rtail<T, 4>(v, s, out, L, P, Ns * 16, scale, t0, q, fwd);
    } else {
        const TwBase<8, C> t0 = rfetch_tail<8, C>(Ns * 16, q, fwd);
        rexchange<T, 16, 8>(v, xch, L, P, Ns);
        rtail<T, 8>(v, s, out, L, P, Ns * 16, scale, t0, q, fwd);
    }
}

template<int T, class C, class TW>
__device__ __forceinline__ void rchunk(C (&v)[kRP], const Desc d, const Sig* s, const void* in, void* out,
                                       typename Words<C>::W* xch, const TW& q, bool fwd, int norm, double fixed_scale)
{
    switch (d.log2n) {
#define AETH_FFT_LEN(l)                                                                                               \
    case l: rchunk_n<T, l>(v, d, s, in, out, xch, q, fwd, norm, fixed_scale); break;
        AETH_FFT_LEN(0) AETH_FFT_LEN(1) AETH_FFT_LEN(2) AETH_FFT_LEN(3) AETH_FFT_LEN(4) AETH_FFT_LEN(5)
        AETH_FFT_LEN(6) AETH_FFT_LEN(7) AETH_FFT_LEN(8) AETH_FFT_LEN(9) AETH_FFT_LEN(10) AETH_FFT_LEN(11)
        AETH_FFT_LEN(12)
#undef AETH_FFT_LEN
    case 13:
        if constexpr (T == kRT) rchunk_n<T, 13>(v, d, s, in, out, xch, q, fwd, norm, fixed_scale);
        break;
    default: break;
    }
}

// T = 512 threads hold a chunk of up to 8192 points of one length. With the length a constant every form fits
// 128 registers and 64 KiB of LDS (8-byte exchange words), so two workgroups (16 waves) share a CU.
// The direction is a template parameter, so the conjugations of the inverse fold into the arithmetic's sign
// modifiers instead of selects.
template<class C, int T, bool FWD>
__global__ __launch_bounds__(T, 4) void fft_reg_kernel(const Desc* __restrict__ descs, const Sig* __restrict__ sigs,
                                                         const void* __restrict__ in, void* __restrict__ out,
                                                         TwTables tw, int norm, double fixed_scale)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    auto* xch = reinterpret_cast<typename Words<C>::W*>(lds_raw);
    const Desc d = descs[blockIdx.x];
    C v[kRP];
    rchunk<T>(v, d, sigs + d.first_sig, in, out, xch, TwGlobal{tw}, FWD, norm, fixed_scale);
}

// The persistent form: a fixed grid walks the descriptors, their number read from device memory, so one
// launch (or one node of a captured graph) serves any table that fits the buffers it points at.
template<class C>
__global__ __launch_bounds__(kRT, sizeof(C) == 8 ? 4 : 2) void fft_reg_persist_kernel(
    const Desc* __restrict__ descs, const Sig* __restrict__ sigs, const uint32_t* __restrict__ ndesc,
    const void* __restrict__ in, void* __restrict__ out, TwTables tw, int fwd_flag, int norm, double fixed_scale)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    auto* xch = reinterpret_cast<typename Words<C>::W*>(lds_raw);
    const uint32_t count = *ndesc;
    for (uint32_t k = blockIdx.x; k < count; k += gridDim.x) {
        __syncthreads(); // the previous chunk's exchanges are done with the LDS
        const Desc d = descs[k];
        C v[kRP];
        // the length read at run time: inside the persistent loop the per-length forms' loop-invariant
        // values would all be hoisted and held at once
        rchunk_n<kRT, -1>(v, d, sigs + d.first_sig, in, out, xch, TwGlobal{tw}, fwd_flag != 0, norm, fixed_scale);
    }
}

__global__ void twiddle_kernel(double2* d, float2* f, float4* ff)
{
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (uint32_t)kMaxN) return;
    // angle -2 pi m / 8192; m/8192 is exact, sincos of the product is within an ulp
    double s, c;
    dm::sincos(-2.0 * 3.141592653589793238 * ((double)m / kMaxN), &s, &c);
    // snap the exact cases (quarter turns) so symmetric twiddles are exact
    if ((m & (kMaxN / 4 - 1)) == 0) {
        const uint32_t q = m / (kMaxN / 4);
        c = q == 0 ? 1.0 : q == 2 ? -1.0 : 0.0;
        s = q == 1 ? -1.0 : q == 3 ? 1.0 : 0.0;
    }
    d[m] = make_double2(c, s);
    f[m] = make_float2((float)c, (float)s);
    const float ch = (float)c, sh = (float)s;
    ff[m] = make_float4(ch, (float)(c - (double)ch), sh, (float)(s - (double)sh));
}

template<class C>
void set_reg_lds()
{
    for (auto fn : {(const void*)fft_reg_kernel<C, kRT, true>, (const void*)fft_reg_kernel<C, kRT, false>})
        check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(kRT * kRP * sizeof(typename Words<C>::W))),
              "hipFuncSetAttribute(LDS)");
}

// Per-device twiddle tables, created once.
struct DeviceTables {
    double2* d = nullptr;
    float2* f = nullptr;
    float4* ff = nullptr;
};
std::mutex g_tw_mutex;
std::map<int, DeviceTables> g_tw;

TwTables tables(hipStream_t st)
{
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    auto it = g_tw.find(dev);
    if (it == g_tw.end()) {
        DeviceTables t;
        check(hipMalloc(&t.d, sizeof(double2) * kMaxN), "hipMalloc(twiddles)", sizeof(double2) * kMaxN);
        check(hipMalloc(&t.f, sizeof(float2) * kMaxN), "hipMalloc(twiddles)", sizeof(float2) * kMaxN);
        check(hipMalloc(&t.ff, sizeof(float4) * kMaxN), "hipMalloc(twiddles)", sizeof(float4) * kMaxN);
        hipLaunchKernelGGL(twiddle_kernel, dim3(kMaxN / 256), dim3(256), 0, st, t.d, t.f, t.ff);
        check(hipGetLastError(), "twiddle_kernel launch");
        check(hipStreamSynchronize(st), "hipStreamSynchronize(twiddles)");
        for (auto fn : {(const void*)fft_kernel<C128>, (const void*)fft_kernel<C64>}) {
            check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsElems * 16),
                  "hipFuncSetAttribute(LDS)");
        }
        set_reg_lds<C128>();
        set_reg_lds<C64>();
        set_reg_lds<CFF>();
        set_reg_lds<CW>();
        check(hipFuncSetAttribute((const void*)fft_reg_persist_kernel<C128>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kMaxN * sizeof(Words<C128>::W)),
              "hipFuncSetAttribute(LDS)");
        check(hipFuncSetAttribute((const void*)fft_reg_persist_kernel<CFF>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kMaxN * sizeof(Words<CFF>::W)),
              "hipFuncSetAttribute(LDS)");

        it = g_tw.emplace(dev, t).first;
    }
    return TwTables{it->second.d, it->second.f, it->second.ff};
}


uint32_t log2_exact(uint64_t n)
{
    if (n < 1 || (n & (n - 1)) != 0) return 0xFFFFFFFFu;
    uint32_t l = 0;
    while ((1ull << l) < n) ++l;
    return l;
}

size_t lds_bytes(int prec) { return (size_t)kLdsElems * (prec == 1 ? sizeof(C64) : 16); }

void build_tables(const uint32_t* log2n, const uint64_t* in_off, const uint64_t* out_off, uint64_t count,
                  std::vector<Desc>& descs, std::vector<Sig>& sigs, uint32_t stride)
{
    // group by length, largest first; a chunk holds up to 8192 points of one length
    std::vector<uint64_t> order(count);
    for (uint64_t i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return log2n[a] > log2n[b]; });
    sigs.resize(count);
    uint64_t i = 0;
    while (i < count) {
        const uint32_t L = log2n[order[i]];
        const uint32_t per = (uint32_t)kMaxN >> L;
        Desc d;
        d.first_sig = (uint32_t)i;
        d.log2n = (uint16_t)L;
        d.stride = stride;
        uint32_t c = 0;
        while (i < count && c < per && log2n[order[i]] == L) {
            sigs[i] = Sig{in_off[order[i]], out_off[order[i]]};
            ++i;
            ++c;
        }
        d.count = (uint16_t)c;
        descs.push_back(d);
    }
}

namespace {
template<class C>
int persist_grid(size_t lds)
{
    int dev = 0, per_cu = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fft_reg_persist_kernel<C>, kRT, lds),
          "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    return std::max(1, per_cu) * device_cus();
}
} // namespace

void prepare(hipStream_t st) { (void)tables(st); }

int persistent_grid(int prec)
{
    if (prec == 2) return persist_grid<CFF>(kMaxN * sizeof(Words<CFF>::W));
    if (prec == 0) return persist_grid<C128>(kMaxN * sizeof(Words<C128>::W));
    fail("fft persistent_grid", "precision not served");
    return 0;
}

void launch_persistent(int prec, int grid, const Desc* descs, const Sig* sigs, const uint32_t* ndesc, const void* in,
                       void* out, int forward, int norm, hipStream_t st)
{
    const TwTables tw = tables(st);
    if (prec == 2)
        hipLaunchKernelGGL(fft_reg_persist_kernel<CFF>, dim3(grid), dim3(kRT), kMaxN * sizeof(Words<CFF>::W), st, descs,
                           sigs, ndesc, in, out, tw, forward, norm, 1.0);
    else if (prec == 0)
        hipLaunchKernelGGL(fft_reg_persist_kernel<C128>, dim3(grid), dim3(kRT), kMaxN * sizeof(Words<C128>::W), st,
                           descs, sigs, ndesc, in, out, tw, forward, norm, 1.0);
    else
        fail("fft launch_persistent", "precision not served");
    check(hipGetLastError(), "fft_reg_persist_kernel launch");
}

namespace {
// Contiguous signals: one launch over every descriptor.
template<class C>
void launch_reg(const BatchPlan& p, const void* in, void* out, const TwTables& tw, int forward, int norm,
                double fixed_scale, hipStream_t st)
{
    const size_t lds = kRT * kRP * sizeof(typename Words<C>::W);
    if (forward)
        hipLaunchKernelGGL((fft_reg_kernel<C, kRT, true>), dim3(p.ndesc), dim3(kRT), lds, st, p.descs, p.sigs, in,
                           out, tw, norm, fixed_scale);
    else
        hipLaunchKernelGGL((fft_reg_kernel<C, kRT, false>), dim3(p.ndesc), dim3(kRT), lds, st, p.descs, p.sigs, in,
                           out, tw, norm, fixed_scale);
    check(hipGetLastError(), "fft_reg_kernel launch");
}
} // namespace

void launch(const BatchPlan& p, const void* in, void* out, int forward, int norm, hipStream_t st,
            double fixed_scale)
{
    const TwTables tw = tables(st);
    const size_t lds = lds_bytes(p.prec);
    if (p.ndesc == 0) return;
    if (p.stride == 1) {
        if (p.prec == 2) launch_reg<CFF>(p, in, out, tw, forward, norm, fixed_scale, st);
        else if (p.prec == 3) launch_reg<CW>(p, in, out, tw, forward, norm, fixed_scale, st);
        else if (p.prec == 0) launch_reg<C128>(p, in, out, tw, forward, norm, fixed_scale, st);
        else launch_reg<C64>(p, in, out, tw, forward, norm, fixed_scale, st);
        return;
    }
    switch (p.prec) {
    case 0:
        hipLaunchKernelGGL(fft_kernel<C128>, dim3(p.ndesc), dim3(kThreads), lds, st, p.descs, p.sigs, in, out, tw,
                           forward, norm, fixed_scale);
        break;
    case 1:
        hipLaunchKernelGGL(fft_kernel<C64>, dim3(p.ndesc), dim3(kThreads), lds, st, p.descs, p.sigs, in, out, tw,
                           forward, norm, fixed_scale);
        break;
    default:
        // strided lines come from the seam's plans, which are complex128 or complex64
        fail("fft launch", "strided lines in a compensated precision");
    }
    check(hipGetLastError(), "fft_kernel launch");
}

} // namespace fft
} // namespace aeth

using namespace aeth;
using namespace aeth::fft;


extern "C" int aeth_fft_batch_plan(int prec, uint64_t count, const uint64_t* lengths, const uint64_t* in_offsets,
                                   const uint64_t* out_offsets, void** plan)
{
    return guarded([&] {
        if (!plan) fail("aeth_fft_batch_plan", "plan is NULL");
        *plan = nullptr;
        if (prec < 0 || prec > 3) fail("aeth_fft_batch_plan", "unknown precision");
        std::vector<uint32_t> l2(count);
        for (uint64_t i = 0; i < count; ++i) {
            const uint32_t l = log2_exact(lengths[i]);
            if (l == 0xFFFFFFFFu || l > (uint32_t)kMaxLog2)
                fail("aeth_fft_batch_plan", "lengths must be powers of two in [1, 8192]", 0, 2);
            l2[i] = l;
        }
        std::vector<Desc> descs;
        std::vector<Sig> sigs;
        build_tables(l2.data(), in_offsets, out_offsets, count, descs, sigs);
        auto* p = new BatchPlan;
        p->prec = prec;
        check(hipGetDevice(&p->dev), "hipGetDevice");
        p->ndesc = (uint32_t)descs.size();
        p->nsig = sigs.size();
        try {
            if (!descs.empty()) {
                check(hipMalloc(&p->descs, sizeof(Desc) * descs.size()), "hipMalloc(descriptors)");
                check(hipMalloc(&p->sigs, sizeof(Sig) * sigs.size()), "hipMalloc(signals)");
                check(hipMemcpy(p->descs, descs.data(), sizeof(Desc) * descs.size(), hipMemcpyHostToDevice),
                      "hipMemcpy(descriptors)");
                check(hipMemcpy(p->sigs, sigs.data(), sizeof(Sig) * sigs.size(), hipMemcpyHostToDevice),
                      "hipMemcpy(signals)");
            }
            (void)tables(nullptr);
        } catch (...) {
            (void)hipFree(p->descs);
            (void)hipFree(p->sigs);
            delete p;
            throw;
        }
        *plan = p;
    });
}

extern "C" int aeth_fft_batch_exec(void* plan, const void* in, void* out, int forward, int norm, void* stream)
{
    return guarded([&] {
        if (!plan) fail("aeth_fft_batch_exec", "plan is NULL");
        auto* p = static_cast<BatchPlan*>(plan);
        if (norm < 0 || norm > 2) fail("aeth_fft_batch_exec", "unknown norm");
        int dev = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        if (dev != p->dev) fail("aeth_fft_batch_exec", "plan belongs to another device");
        launch(*p, in, out, forward, norm, as_stream(stream));
    });
}

extern "C" int aeth_fft_batch_destroy(void* plan)
{
    return guarded([&] {
        auto* p = static_cast<BatchPlan*>(plan);
        if (!p) return;
        // safe while a capture is in progress: the tables are freed once no capture stands in the way
        free_device_later({p->descs, p->sigs});
        delete p;
    });
}

// ---------------------------------------------------------------------------
// Plans for the memory seam's fft_plan/fft_exec/fft_destroy
// (dsc/src/seam.h:41-50, dsc/src/cufft_plan.h:125-222): c2c over any set of
// axes of a dense row-major array, one strided-line pass per axis. Real
// transforms (cufft_plan.h:9-12: `shape` is the real side's, the last axis
// is halved to n/2 + 1 on the complex side) run the halved axis as a complex
// transform of the whole line in a plan-owned work array: r2c widens the
// reals, transforms, keeps bins 0..n/2, then runs the other axes in place on
// the output; c2r runs the other axes first (into a plan-owned copy, the
// input is left unchanged), extends each line by Hermitian symmetry
// (X[n-k] = conj X[k]), transforms and keeps the real parts -- which ignores
// the imaginary parts of the DC and Nyquist bins, as a real inverse does.
namespace aeth {
namespace fft {

namespace {

constexpr int kAuxThreads = 256;

inline dim3 aux_grid(uint64_t items)
{
    const uint64_t b = (items + kAuxThreads - 1) / kAuxThreads;
    return dim3((unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 65536)));
}

// reals -> complex (x, 0)
template<class T>
__global__ __launch_bounds__(kAuxThreads) void widen_kernel(const T* __restrict__ in, T* __restrict__ out,
                                                            uint64_t count)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * kAuxThreads) {
        out[2 * i] = in[i];
        out[2 * i + 1] = T(0);
    }
}

// bins 0..n/2 of every line along the halved axis: full (outer, n, inner) -> half (outer, n/2+1, inner)
template<class T>
__global__ __launch_bounds__(kAuxThreads) void keep_half_kernel(const T* __restrict__ full, T* __restrict__ half,
                                                                uint64_t outer, uint64_t n, uint64_t inner)
{
    const uint64_t h = n / 2 + 1, count = outer * h * inner;
    for (uint64_t t = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; t < count;
         t += (uint64_t)gridDim.x * kAuxThreads) {
        const uint64_t j = t % inner, r = t / inner, k = r % h, o = r / h;
        const uint64_t src = (o * n + k) * inner + j;
        half[2 * t] = full[2 * src];
        half[2 * t + 1] = full[2 * src + 1];
    }
}

// Hermitian extension: half (outer, n/2+1, inner) -> full (outer, n, inner), X[k] = conj X[n-k] for k > n/2
template<class T>
__global__ __launch_bounds__(kAuxThreads) void hermitian_kernel(const T* __restrict__ half, T* __restrict__ full,
                                                                uint64_t outer, uint64_t n, uint64_t inner)
{
    const uint64_t h = n / 2 + 1, count = outer * n * inner;
    for (uint64_t t = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; t < count;
         t += (uint64_t)gridDim.x * kAuxThreads) {
        const uint64_t j = t % inner, r = t / inner, k = r % n, o = r / n;
        const bool mirror = k >= h;
        const uint64_t src = (o * h + (mirror ? n - k : k)) * inner + j;
        full[2 * t] = half[2 * src];
        full[2 * t + 1] = mirror ? -half[2 * src + 1] : half[2 * src + 1];
    }
}

template<class T>
__global__ __launch_bounds__(kAuxThreads) void real_part_kernel(const T* __restrict__ full, T* __restrict__ out,
                                                                uint64_t count)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * kAuxThreads)
        out[i] = full[2 * i];
}

// ---- Bluestein (chirp-z) passes for lengths that are not powers of two.
// X_k = c_k * sum_j (x_j c_j) conj(c_{k-j}) with c_j = exp(-+ i pi j^2 / n) (forward / inverse):
// the sum is a circular convolution of length m (a power of two >= 2n - 1) done with two of the
// batched LDS transforms and a pointwise product with the transform of the chirp, precomputed per
// plan with the 1/m of the inverse folded in. The chirp angle is pi (j^2 mod 2n) / n, reduced
// exactly in integers before the one rounding. The convolution runs in complex128 for both widths
// (a complex64 plan reads and writes float, rounding once at the end), so a float transform of any
// length is as accurate as the double one rounded to float.
// ---- four-step passes for powers of two above 8192 (up to 2^34 points per line): a line of
// n = n1 n2 points is an n1 x n2 matrix (element (j1, j2) at n2 j1 + j2); transform its columns
// (n1 points, stride n2), multiply by W_n^(j2 k1), transform its rows (n2 points), and read the result
// out transposed: X[k1 + n1 k2]. All in complex128 in a plan-owned copy of the lines; W_n^m is the
// product of two table entries, W_n^(8192 h) W_n^l with m = 8192 h + l. Up to 2^26, n1 and n2 are at
// most 8192 (one batched launch each); above, n1 = 8192 and the rows (n2 > 8192 points) are themselves
// a four-step pass over the dense rows, in place in `work`.
struct FourStepPass {
    uint64_t n = 0, n1 = 0, n2 = 0, lines = 0, inner = 1;
    int prec = 0;                     // of the data
    BatchPlan* cols = nullptr;        // lines x n2 signals of n1 points, stride n2, in `work`
    BatchPlan* rows = nullptr;        // lines x n1 signals of n2 points, contiguous, in `work` (n2 <= 8192)
    FourStepPass* rows_four = nullptr; // the same rows when n2 > 8192
    double* work = nullptr;           // lines x n complex128
    double* thi = nullptr;            // W_n^(8192 h), h < n / 8192
    double* tlo = nullptr;            // W_n^l, l < 8192
};

struct BluesteinPass;

// One axis of a plan: a Stockham pass (powers of two up to 8192), a four-step pass (powers of two up
// to 2^34) or a Bluestein pass (other lengths below 2^32).
struct Pass {
    BatchPlan* stockham = nullptr;
    FourStepPass* four = nullptr;
    BluesteinPass* blue = nullptr;
};

struct BluesteinPass {
    uint64_t n = 0, m = 0, lines = 0, inner = 1;
    int prec = 0;           // of the data: 0 complex128, 1 complex64
    Pass conv;              // lines x m contiguous complex128 signals
    double* work = nullptr; // lines x m complex128
    double* chirp = nullptr; // n complex: exp(i pi j^2 / n) (sign applied per direction)
    double* bhat = nullptr;  // m complex: FFT_m(b) / m for the plan's direction
};

constexpr uint32_t kTwLo = 8192;

// W_n^m for m < n (n a power of two): angle -2 pi m / n, m / n exact; quarter turns exact
__device__ __forceinline__ void unit_root(uint64_t m, uint64_t n, double* re, double* im)
{
    double s, c;
    dm::sincos(-2.0 * 3.141592653589793238 * ((double)m / (double)n), &s, &c);
    if (n >= 4 && (m % (n / 4)) == 0) {
        const uint64_t q = m / (n / 4);
        c = q == 0 ? 1.0 : q == 2 ? -1.0 : 0.0;
        s = q == 1 ? -1.0 : q == 3 ? 1.0 : 0.0;
    }
    *re = c;
    *im = s;
}

__global__ __launch_bounds__(kAuxThreads) void fs_table_kernel(double* __restrict__ thi, double* __restrict__ tlo,
                                                               uint64_t n)
{
    const uint64_t nhi = n / kTwLo;
    for (uint64_t t = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; t < kTwLo + nhi;
         t += (uint64_t)gridDim.x * kAuxThreads) {
        if (t < kTwLo) unit_root(t, n, &tlo[2 * t], &tlo[2 * t + 1]);
        else unit_root((t - kTwLo) * kTwLo, n, &thi[2 * (t - kTwLo)], &thi[2 * (t - kTwLo) + 1]);
    }
}

// lines of `in` (n points, stride inner) into dense complex128 lines
template<class T>
__global__ __launch_bounds__(kAuxThreads) void fs_gather_kernel(const T* __restrict__ in, double* __restrict__ work,
                                                                uint64_t lines, uint64_t n, uint64_t inner)
{
    const uint64_t count = lines * n;
    for (uint64_t t = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; t < count; t += (uint64_t)gridDim.x * kAuxThreads) {
        const uint64_t li = t / n, e = t - li * n;
        const uint64_t o = li / inner, q = li - o * inner;
        const uint64_t src = (o * n + e) * inner + q;
        work[2 * t] = (double)in[2 * src];
        work[2 * t + 1] = (double)in[2 * src + 1];
    }
}

// after the column transforms: element (k1, j2) of every line times W_n^(j2 k1) (conjugated for the inverse)
__global__ __launch_bounds__(kAuxThreads) void fs_twiddle_kernel(double* __restrict__ work,
                                                                 const double* __restrict__ thi,
                                                                 const double* __restrict__ tlo, uint64_t lines,
                                                                 uint64_t n2, uint64_t n, double sign)
{
    const uint64_t count = lines * n;
    for (uint64_t t = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; t < count; t += (uint64_t)gridDim.x * kAuxThreads) {
        const uint64_t e = t % n, k1 = e / n2, j2 = e - k1 * n2;
        const uint64_t m = k1 * j2; // < n
        const uint64_t h = m / kTwLo, l = m - h * kTwLo;
        const double ar = thi[2 * h], ai = thi[2 * h + 1], br = tlo[2 * l], bi = tlo[2 * l + 1];
        const double wr = ar * br - ai * bi, wi = sign * (ar * bi + ai * br);
        const double xr = work[2 * t], xi = work[2 * t + 1];
        work[2 * t] = xr * wr - xi * wi;
        work[2 * t + 1] = xr * wi + xi * wr;
    }
}

// read out transposed: X[k1 + n1 k2] = work[k1 n2 + k2], scaled, along the strided output line
template<class T>
__global__ __launch_bounds__(kAuxThreads) void fs_scatter_kernel(const double* __restrict__ work, T* __restrict__ out,
                                                                 uint64_t lines, uint64_t n1, uint64_t n2,
                                                                 uint64_t inner, double scale)
{
    const uint64_t n = n1 * n2, count = lines * n;
    for (uint64_t t = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; t < count; t += (uint64_t)gridDim.x * kAuxThreads) {
        const uint64_t li = t / n, e = t - li * n;
        const uint64_t k1 = e / n2, k2 = e - k1 * n2;
        const uint64_t o = li / inner, q = li - o * inner;
        const uint64_t dst = (o * n + k1 + n1 * k2) * inner + q;
        out[2 * dst] = (T)(work[2 * t] * scale);
        out[2 * dst + 1] = (T)(work[2 * t + 1] * scale);
    }
}

__device__ __forceinline__ void chirp_at(uint64_t j, uint64_t n, double* re, double* im)
{
    const uint64_t r = (j * j) % (2 * n);
    dm::sincos(3.141592653589793238 * ((double)r / (double)n), im, re);
}

__global__ __launch_bounds__(kAuxThreads) void chirp_kernel(double* __restrict__ chirp, uint64_t n)
{
    for (uint64_t j = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; j < n; j += (uint64_t)gridDim.x * kAuxThreads)
        chirp_at(j, n, &chirp[2 * j], &chirp[2 * j + 1]);
}

// b_l = conj^dir(c_l) for l < n and b_{m-l} = b_l, zero between; sign = -1 makes it conj
__global__ __launch_bounds__(kAuxThreads) void chirp_b_kernel(const double* __restrict__ chirp, double* __restrict__ b,
                                                              uint64_t n, uint64_t m, double sign)
{
    for (uint64_t l = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; l < m; l += (uint64_t)gridDim.x * kAuxThreads) {
        const uint64_t k = l < n ? l : (m - l < n ? m - l : m);
        b[2 * l] = k < m ? chirp[2 * k] : 0.0;
        b[2 * l + 1] = k < m ? sign * chirp[2 * k + 1] : 0.0;
    }
}

__global__ __launch_bounds__(kAuxThreads) void scale_kernel(double* __restrict__ x, uint64_t count, double s)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; i < count; i += (uint64_t)gridDim.x * kAuxThreads)
        x[i] *= s;
}

// a_j = x_j * conj^dir(c_j), zero-padded to m; x read along a strided line
template<class T>
__global__ __launch_bounds__(kAuxThreads) void blue_pre_kernel(const T* __restrict__ in, double* __restrict__ work,
                                                               const double* __restrict__ chirp, uint64_t lines,
                                                               uint64_t n, uint64_t m, uint64_t inner, double sign)
{
    const uint64_t count = lines * m;
    for (uint64_t t = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; t < count; t += (uint64_t)gridDim.x * kAuxThreads) {
        const uint64_t li = t / m, j = t - li * m;
        double re = 0.0, im = 0.0;
        if (j < n) {
            const uint64_t o = li / inner, q = li - o * inner;
            const uint64_t e = (o * n + j) * inner + q;
            const double xr = in[2 * e], xi = in[2 * e + 1];
            const double cr = chirp[2 * j], ci = sign * chirp[2 * j + 1];
            re = xr * cr - xi * ci;
            im = xr * ci + xi * cr;
        }
        work[2 * t] = re;
        work[2 * t + 1] = im;
    }
}

__global__ __launch_bounds__(kAuxThreads) void blue_mul_kernel(double* __restrict__ work,
                                                               const double* __restrict__ bhat, uint64_t lines,
                                                               uint64_t m)
{
    const uint64_t count = lines * m;
    for (uint64_t t = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; t < count; t += (uint64_t)gridDim.x * kAuxThreads) {
        const uint64_t j = t % m;
        const double ar = work[2 * t], ai = work[2 * t + 1];
        const double br = bhat[2 * j], bi = bhat[2 * j + 1];
        work[2 * t] = ar * br - ai * bi;
        work[2 * t + 1] = ar * bi + ai * br;
    }
}

// X_k = conj^dir(c_k) * conv_k * scale, written along the strided line
template<class T>
__global__ __launch_bounds__(kAuxThreads) void blue_post_kernel(const double* __restrict__ work, T* __restrict__ out,
                                                                const double* __restrict__ chirp, uint64_t lines,
                                                                uint64_t n, uint64_t m, uint64_t inner, double sign,
                                                                double scale)
{
    const uint64_t count = lines * n;
    for (uint64_t t = (uint64_t)blockIdx.x * kAuxThreads + threadIdx.x; t < count; t += (uint64_t)gridDim.x * kAuxThreads) {
        const uint64_t li = t / n, k = t - li * n;
        const uint64_t w = li * m + k;
        const double vr = work[2 * w], vi = work[2 * w + 1];
        const double cr = chirp[2 * k], ci = sign * chirp[2 * k + 1];
        const uint64_t o = li / inner, q = li - o * inner;
        const uint64_t e = (o * n + k) * inner + q;
        out[2 * e] = (T)((vr * cr - vi * ci) * scale);
        out[2 * e + 1] = (T)((vr * ci + vi * cr) * scale);
    }
}

} // namespace

struct SeamPlan {
    int dev = 0;
    int kind = AETHER_FFT_C2C;
    int forward = 1;
    bool dbl = true;
    std::vector<Pass> passes; // complex passes on the complex-side shape
    std::vector<double> scales;
    Pass line;                // real transforms: the halved axis over the full complex shape
    double line_scale = 1.0;
    void* full = nullptr;            // real transforms: (outer, n, inner) complex work
    void* half = nullptr;            // c2r with other axes: (outer, n/2+1, inner) complex work
    uint64_t outer = 1, n = 1, inner = 1;
    std::mutex mu; // cufft_plan.h:187 -- one execution at a time per plan
    // Plan-owned device memory goes through free_device_later: a plan destroyed while a stream
    // capture is in progress (dsc/src/seam.h:50, "safe during a capture") neither breaks the capture
    // nor leaks; its buffers are freed once no capture stands in the way.
    static void free_pass(BatchPlan* p, std::vector<void*>& bufs)
    {
        if (!p) return;
        bufs.push_back(p->descs);
        bufs.push_back(p->sigs);
        delete p;
    }
    static void free_four(FourStepPass* f, std::vector<void*>& bufs)
    {
        if (!f) return;
        free_pass(f->cols, bufs);
        free_pass(f->rows, bufs);
        free_four(f->rows_four, bufs);
        bufs.push_back(f->work);
        bufs.push_back(f->thi);
        bufs.push_back(f->tlo);
        delete f;
    }
    static void free_pass(Pass& p, std::vector<void*>& bufs)
    {
        free_pass(p.stockham, bufs);
        free_four(p.four, bufs);
        if (p.blue) {
            free_pass(p.blue->conv, bufs);
            bufs.push_back(p.blue->work);
            bufs.push_back(p.blue->chirp);
            bufs.push_back(p.blue->bhat);
            delete p.blue;
        }
        p = Pass{};
    }
    static void free_pass(BatchPlan* p)
    {
        std::vector<void*> bufs;
        free_pass(p, bufs);
        free_device_later(std::move(bufs));
    }
    static void free_pass(Pass& p)
    {
        std::vector<void*> bufs;
        free_pass(p, bufs);
        free_device_later(std::move(bufs));
    }
    ~SeamPlan()
    {
        std::vector<void*> bufs;
        for (auto& p : passes) free_pass(p, bufs);
        free_pass(line, bufs);
        bufs.push_back(full);
        bufs.push_back(half);
        free_device_later(std::move(bufs));
    }
};

namespace {

// The descriptor tables of `lines` signals of 2^l2 points at `off[li]`, `stride` apart, on the device.
BatchPlan* make_batch(const std::vector<uint64_t>& off, uint32_t l2, uint64_t stride, int dev, bool dbl)
{
    const uint64_t lines = off.size();
    std::vector<uint32_t> lg(lines, l2);
    std::vector<Desc> descs;
    std::vector<Sig> sigs;
    build_tables(lg.data(), off.data(), off.data(), lines, descs, sigs, (uint32_t)stride);
    auto bp = std::make_unique<BatchPlan>();
    bp->prec = dbl ? 0 : 1;
    bp->dev = dev;
    bp->stride = (uint32_t)stride;
    bp->ndesc = (uint32_t)descs.size();
    bp->nsig = sigs.size();
    if (!descs.empty()) {
        check(hipMalloc(&bp->descs, sizeof(Desc) * descs.size()), "hipMalloc(descriptors)");
        if (hipMalloc(&bp->sigs, sizeof(Sig) * sigs.size()) != hipSuccess) {
            (void)hipFree(bp->descs);
            fail("fft_plan", "hipMalloc(signals) failed", sizeof(Sig) * sigs.size());
        }
        const bool ok =
            hipMemcpy(bp->descs, descs.data(), sizeof(Desc) * descs.size(), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(bp->sigs, sigs.data(), sizeof(Sig) * sigs.size(), hipMemcpyHostToDevice) == hipSuccess;
        if (!ok) {
            SeamPlan::free_pass(bp.release());
            fail("fft_plan", "hipMemcpy(descriptor tables) failed");
        }
    }
    return bp.release();
}

constexpr uint32_t kMaxFourLog2 = 34; // four-step: up to 2^34 points (two levels of 8192-point factors)
// Bluestein: 2n - 1 <= 2^34, and n < 2^32 so the chirp's j^2 stays in 64 bits
constexpr uint64_t kMaxBluestein = 0xFFFFFFFFull;

FourStepPass* make_four(uint64_t n, uint64_t lines, uint64_t inner, int dev, bool dbl, hipStream_t st)
{
    auto f = std::make_unique<FourStepPass>();
    struct Guard {
        std::unique_ptr<FourStepPass>& p;
        ~Guard()
        {
            if (!p) return;
            std::vector<void*> bufs;
            SeamPlan::free_four(p.release(), bufs);
            for (void* b : bufs) (void)hipFree(b);
        }
    } guard{f};
    const uint32_t L = log2_exact(n);
    // up to 2^26: both factors <= 8192; above: columns of 8192 points, rows of 2^(L - 13) (a four-step)
    const uint32_t L1 = L <= 2 * (uint32_t)kMaxLog2 ? L / 2 : (uint32_t)kMaxLog2, L2 = L - L1;
    f->n = n;
    f->n1 = 1ull << L1;
    f->n2 = 1ull << L2;
    f->lines = lines;
    f->inner = inner;
    f->prec = dbl ? 0 : 1;
    check(hipMalloc(&f->work, 2 * sizeof(double) * lines * n), "hipMalloc(fft work)", 2 * sizeof(double) * lines * n);
    check(hipMalloc(&f->thi, 2 * sizeof(double) * (n / kTwLo)), "hipMalloc(four-step twiddles)");
    check(hipMalloc(&f->tlo, 2 * sizeof(double) * kTwLo), "hipMalloc(four-step twiddles)");
    std::vector<uint64_t> off;
    off.reserve(lines * f->n2);
    for (uint64_t li = 0; li < lines; ++li)
        for (uint64_t j2 = 0; j2 < f->n2; ++j2) off.push_back(li * n + j2);
    f->cols = make_batch(off, L1, f->n2, dev, true);
    off.clear();
    if (L2 <= (uint32_t)kMaxLog2) {
        for (uint64_t li = 0; li < lines; ++li)
            for (uint64_t k1 = 0; k1 < f->n1; ++k1) off.push_back(li * n + k1 * f->n2);
        f->rows = make_batch(off, L2, 1, dev, true);
    } else {
        f->rows_four = make_four(f->n2, lines * f->n1, 1, dev, true, st);
    }
    hipLaunchKernelGGL(fs_table_kernel, aux_grid(kTwLo + n / kTwLo), dim3(kAuxThreads), 0, st, f->thi, f->tlo, n);
    check(hipGetLastError(), "four-step table kernel");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    return f.release();
}

template<class T>
void run_four(const FourStepPass& f, const void* in, void* out, int forward, double scale, hipStream_t st)
{
    const uint64_t count = f.lines * f.n;
    hipLaunchKernelGGL(fs_gather_kernel<T>, aux_grid(count), dim3(kAuxThreads), 0, st, static_cast<const T*>(in), f.work,
                       f.lines, f.n, f.inner);
    launch(*f.cols, f.work, f.work, forward, 3, st, 1.0);
    hipLaunchKernelGGL(fs_twiddle_kernel, aux_grid(count), dim3(kAuxThreads), 0, st, f.work, (const double*)f.thi,
                       (const double*)f.tlo, f.lines, f.n2, f.n, forward ? 1.0 : -1.0);
    if (f.rows_four) run_four<double>(*f.rows_four, f.work, f.work, forward, 1.0, st);
    else launch(*f.rows, f.work, f.work, forward, 3, st, 1.0);
    hipLaunchKernelGGL(fs_scatter_kernel<T>, aux_grid(count), dim3(kAuxThreads), 0, st, (const double*)f.work,
                       static_cast<T*>(out), f.lines, f.n1, f.n2, f.inner, scale);
    check(hipGetLastError(), "four-step kernels");
}

// A power-of-two transform over `lines` dense complex128 lines of m points (Bluestein's convolution).
Pass make_dense_pow2(uint64_t m, uint64_t lines, int dev, hipStream_t st)
{
    Pass p;
    const uint32_t l2 = log2_exact(m);
    if (l2 <= (uint32_t)kMaxLog2) {
        std::vector<uint64_t> off(lines);
        for (uint64_t li = 0; li < lines; ++li) off[li] = li * m;
        p.stockham = make_batch(off, l2, 1, dev, true);
    } else {
        p.four = make_four(m, lines, 1, dev, true, st);
    }
    return p;
}

void run_pass(const Pass& p, const void* in, void* out, int forward, double scale, hipStream_t st);

BluesteinPass* make_bluestein(uint64_t n, uint64_t lines, uint64_t inner, int dev, bool dbl, bool forward,
                              hipStream_t st)
{
    auto bp = std::make_unique<BluesteinPass>();
    struct Guard {
        std::unique_ptr<BluesteinPass>& p;
        ~Guard()
        {
            if (!p) return;
            SeamPlan::free_pass(p->conv);
            (void)hipFree(p->work);
            (void)hipFree(p->chirp);
            (void)hipFree(p->bhat);
        }
    } guard{bp};
    bp->n = n;
    bp->lines = lines;
    bp->inner = inner;
    bp->prec = dbl ? 0 : 1;
    uint64_t m = 1;
    uint32_t l2 = 0;
    while (m < 2 * n - 1) {
        m <<= 1;
        ++l2;
    }
    bp->m = m;
    (void)l2;
    check(hipMalloc(&bp->work, 2 * sizeof(double) * lines * m), "hipMalloc(fft work)", 2 * sizeof(double) * lines * m);
    check(hipMalloc(&bp->chirp, 2 * sizeof(double) * n), "hipMalloc(chirp)");
    check(hipMalloc(&bp->bhat, 2 * sizeof(double) * m), "hipMalloc(chirp transform)");
    bp->conv = make_dense_pow2(m, lines, dev, st);
    // the chirp, then b and its transform in complex128 (one signal of m points), scaled by 1/m
    hipLaunchKernelGGL(chirp_kernel, aux_grid(n), dim3(kAuxThreads), 0, st, bp->chirp, n);
    hipLaunchKernelGGL(chirp_b_kernel, aux_grid(m), dim3(kAuxThreads), 0, st, (const double*)bp->chirp, bp->bhat, n, m,
                       forward ? 1.0 : -1.0);
    Pass one = make_dense_pow2(m, 1, dev, st);
    run_pass(one, bp->bhat, bp->bhat, 1, 1.0, st);
    hipLaunchKernelGGL(scale_kernel, aux_grid(2 * m), dim3(kAuxThreads), 0, st, bp->bhat, 2 * m, 1.0 / (double)m);
    check(hipGetLastError(), "chirp kernels");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    SeamPlan::free_pass(one);
    return bp.release();
}

// One strided-line pass along `ax` of a dense row-major complex `shape`: Stockham for powers of two up
// to 8192, four-step for powers of two up to 2^34, Bluestein for other lengths below 2^32; anything else
// is unsupported (2).
Pass make_pass(const std::vector<int64_t>& shape, int ax, int dev, bool dbl, bool forward, hipStream_t st)
{
    const int ndim = (int)shape.size();
    uint64_t total = 1;
    for (int d = 0; d < ndim; ++d) total *= (uint64_t)shape[d];
    const uint64_t n = (uint64_t)shape[ax];
    uint64_t inner = 1;
    for (int d = ax + 1; d < ndim; ++d) inner *= (uint64_t)shape[d];
    if (inner > 0xFFFFFFFFull) fail("fft_plan", "inner extent too large");
    const uint64_t lines = n ? total / n : 0;
    Pass p;
    const uint32_t l2 = log2_exact(n);
    if (l2 != 0xFFFFFFFFu && l2 <= (uint32_t)kMaxLog2) {
        std::vector<uint64_t> off(lines);
        for (uint64_t li = 0; li < lines; ++li) {
            const uint64_t o = li / inner, in = li % inner;
            off[li] = o * n * inner + in;
        }
        p.stockham = make_batch(off, l2, inner, dev, dbl);
        return p;
    }
    if (l2 != 0xFFFFFFFFu && l2 <= kMaxFourLog2) {
        if (lines) p.four = make_four(n, lines, inner, dev, dbl, st);
        return p;
    }
    if (n == 0 || n > kMaxBluestein)
        fail("fft_plan", "axis lengths must be powers of two up to 2^34 or other lengths below 2^32", 0, 2);
    if (lines) p.blue = make_bluestein(n, lines, inner, dev, dbl, forward, st);
    return p;
}

template<class T>
void run_bluestein(const BluesteinPass& b, const void* in, void* out, hipStream_t st, double scale, bool forward)
{
    const double sign = forward ? -1.0 : 1.0; // conj(c) forward
    double* work = b.work;
    hipLaunchKernelGGL(blue_pre_kernel<T>, aux_grid(b.lines * b.m), dim3(kAuxThreads), 0, st,
                       static_cast<const T*>(in), work, (const double*)b.chirp, b.lines, b.n, b.m, b.inner, sign);
    run_pass(b.conv, work, work, 1, 1.0, st);
    hipLaunchKernelGGL(blue_mul_kernel, aux_grid(b.lines * b.m), dim3(kAuxThreads), 0, st, work,
                       (const double*)b.bhat, b.lines, b.m);
    run_pass(b.conv, work, work, 0, 1.0, st);
    hipLaunchKernelGGL(blue_post_kernel<T>, aux_grid(b.lines * b.n), dim3(kAuxThreads), 0, st, (const double*)work,
                       static_cast<T*>(out), (const double*)b.chirp, b.lines, b.n, b.m, b.inner, sign, scale);
    check(hipGetLastError(), "bluestein kernels");
}

// One pass from `in` to `out` (in == out allowed) with the transform's direction and scale.
void run_pass(const Pass& p, const void* in, void* out, int forward, double scale, hipStream_t st)
{
    if (p.stockham) {
        launch(*p.stockham, in, out, forward, 3, st, scale);
    } else if (p.four) {
        // the convolution passes of Bluestein run on complex128 work whatever the data width
        if (p.four->prec == 0) run_four<double>(*p.four, in, out, forward, scale, st);
        else run_four<float>(*p.four, in, out, forward, scale, st);
    } else if (p.blue) {
        if (p.blue->prec == 0) run_bluestein<double>(*p.blue, in, out, st, scale, forward != 0);
        else run_bluestein<float>(*p.blue, in, out, st, scale, forward != 0);
    }
}

} // namespace

int seam_plan(int dev, int kind, int is_double, int norm, int forward, int ndim, const int64_t* shape, int naxes,
              const int32_t* axes, void* stream, void** plan)
{
    return guarded([&] {
        if (!plan) fail("fft_plan", "plan is NULL");
        *plan = nullptr;
        if (kind < AETHER_FFT_C2C || kind > AETHER_FFT_C2R) fail("fft_plan", "unknown transform kind");
        if (ndim < 1 || naxes < 1) fail("fft_plan", "no axes to transform");
        std::vector<bool> seen(ndim, false);
        for (int a = 0; a < naxes; ++a) {
            if (axes[a] < 0 || axes[a] >= ndim) fail("fft_plan", "axis out of range");
            if (seen[axes[a]]) fail("fft_plan", "repeated axis");
            seen[axes[a]] = true;
        }
        for (int d = 0; d < ndim; ++d)
            if (shape[d] < 0) fail("fft_plan", "negative extent");
        if (norm < 0 || norm > 2) fail("fft_plan", "unknown norm");
        // dsc/src/cufft_plan.h:129-131: a plan allocates, so it is built before a capture begins
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        check(hipStreamIsCapturing(as_stream(stream), &cs), "hipStreamIsCapturing");
        if (cs != hipStreamCaptureStatusNone)
            fail("fft_plan", "fft plan creation is not allowed inside a stream capture: build the plan before the "
                             "capture begins");
        int prev = 0;
        check(hipGetDevice(&prev), "hipGetDevice");
        if (prev != dev) check(hipSetDevice(dev), "hipSetDevice");
        struct Restore {
            int prev, dev;
            ~Restore()
            {
                if (prev != dev) (void)hipSetDevice(prev);
            }
        } restore{prev, dev};
        auto sp = std::make_unique<SeamPlan>();
        sp->dev = dev;
        sp->kind = kind;
        // cufft_plan.h:127 -- r2c is forward and c2r inverse whatever `forward` says
        sp->forward = kind == AETHER_FFT_R2C ? 1 : kind == AETHER_FFT_C2R ? 0 : forward;
        sp->dbl = is_double != 0;
        const std::vector<int64_t> real_shape(shape, shape + ndim);
        uint64_t points = 1;
        for (int a = 0; a < naxes; ++a) points *= (uint64_t)shape[axes[a]];
        // fft/Fft.h:73-76 scale of the whole transform, applied by the last complex step
        double sc = 1.0;
        if (norm == 2) sc = 1.0 / std::sqrt((double)points);
        else if ((sp->forward != 0) == (norm == 1)) sc = 1.0 / (double)points;
        if (kind == AETHER_FFT_C2C) {
            for (int a = 0; a < naxes; ++a) {
                sp->passes.push_back(make_pass(real_shape, axes[a], dev, sp->dbl, sp->forward != 0, as_stream(stream)));
                sp->scales.push_back(a + 1 == naxes ? sc : 1.0);
            }
        } else {
            const int ax = axes[naxes - 1];
            std::vector<int64_t> half = real_shape;
            half[ax] = real_shape[ax] / 2 + 1;
            sp->n = (uint64_t)real_shape[ax];
            for (int d = 0; d < ax; ++d) sp->outer *= (uint64_t)real_shape[d];
            for (int d = ax + 1; d < ndim; ++d) sp->inner *= (uint64_t)real_shape[d];
            // r2c: halved axis first, then the others (the last carries the scale);
            // c2r: the others first, then the halved axis with the scale
            sp->line = make_pass(real_shape, ax, dev, sp->dbl, sp->forward != 0, as_stream(stream));
            for (int a = 0; a + 1 < naxes; ++a) {
                sp->passes.push_back(make_pass(half, axes[a], dev, sp->dbl, sp->forward != 0, as_stream(stream)));
                sp->scales.push_back(kind == AETHER_FFT_R2C && a + 2 == naxes ? sc : 1.0);
            }
            sp->line_scale = (kind == AETHER_FFT_C2R || naxes == 1) ? sc : 1.0;
            const size_t elem = sp->dbl ? sizeof(double) : sizeof(float);
            const uint64_t total_full = sp->outer * sp->n * sp->inner;
            const uint64_t total_half = sp->outer * (sp->n / 2 + 1) * sp->inner;
            if (total_full) check(hipMalloc(&sp->full, 2 * elem * total_full), "hipMalloc(fft work)");
            if (kind == AETHER_FFT_C2R && naxes > 1 && total_half)
                check(hipMalloc(&sp->half, 2 * elem * total_half), "hipMalloc(fft work)");
        }
        (void)tables(as_stream(stream));
        check(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
        *plan = sp.release();
    });
}

namespace {

template<class T>
void run_real(SeamPlan& sp, const void* in, void* out, hipStream_t st)
{
    const uint64_t h = sp.n / 2 + 1;
    const uint64_t total_full = sp.outer * sp.n * sp.inner, total_half = sp.outer * h * sp.inner;
    if (total_full == 0) return;
    T* full = static_cast<T*>(sp.full);
    if (sp.kind == AETHER_FFT_R2C) {
        hipLaunchKernelGGL(widen_kernel<T>, aux_grid(total_full), dim3(kAuxThreads), 0, st,
                           static_cast<const T*>(in), full, total_full);
        check(hipGetLastError(), "widen_kernel launch");
        run_pass(sp.line, full, full, 1, sp.line_scale, st);
        hipLaunchKernelGGL(keep_half_kernel<T>, aux_grid(total_half), dim3(kAuxThreads), 0, st, full,
                           static_cast<T*>(out), sp.outer, sp.n, sp.inner);
        check(hipGetLastError(), "keep_half_kernel launch");
        for (size_t i = 0; i < sp.passes.size(); ++i) run_pass(sp.passes[i], out, out, 1, sp.scales[i], st);
        return;
    }
    const void* src = in;
    for (size_t i = 0; i < sp.passes.size(); ++i) {
        run_pass(sp.passes[i], i == 0 ? in : sp.half, sp.half, 0, sp.scales[i], st);
        src = sp.half;
    }
    hipLaunchKernelGGL(hermitian_kernel<T>, aux_grid(total_full), dim3(kAuxThreads), 0, st,
                       static_cast<const T*>(src), full, sp.outer, sp.n, sp.inner);
    check(hipGetLastError(), "hermitian_kernel launch");
    run_pass(sp.line, full, full, 0, sp.line_scale, st);
    hipLaunchKernelGGL(real_part_kernel<T>, aux_grid(total_full), dim3(kAuxThreads), 0, st, full,
                       static_cast<T*>(out), total_full);
    check(hipGetLastError(), "real_part_kernel launch");
}

} // namespace

int seam_exec(void* plan, const void* in, void* out, void* stream)
{
    return guarded([&] {
        auto* sp = static_cast<SeamPlan*>(plan);
        if (!sp) fail("fft_exec", "plan is NULL");
        std::lock_guard<std::mutex> lock(sp->mu);
        const hipStream_t st = as_stream(stream);
        if (sp->kind == AETHER_FFT_C2C) {
            for (size_t i = 0; i < sp->passes.size(); ++i)
                run_pass(sp->passes[i], i == 0 ? in : out, out, sp->forward, sp->scales[i], st);
            return;
        }
        if (in == out) fail("fft_exec", "in == out is allowed for c2c only");
        if (sp->dbl) run_real<double>(*sp, in, out, st);
        else run_real<float>(*sp, in, out, st);
    });
}

void seam_destroy(void* plan) { delete static_cast<SeamPlan*>(plan); }

} // namespace fft
} // namespace aeth
