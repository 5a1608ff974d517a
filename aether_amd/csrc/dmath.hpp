// dmath.hpp -- device transcendental functions for gfx950.
//
// The reference routes device FP64 sin/cos to its own Cody-Waite/Payne-Hanek
// routine and everything else to the vendor library (math/math.h:228-287,
// math/detail/BoundedTrig.h:262-598, math/detail/MathDispatch.h:54-112).
// These are written from the mathematics instead, so the CPU restatement can
// run the identical operation sequence and the two agree bit for bit:
//   sin/cos : n = rint(x 2/pi); Cody-Waite in three parts of pi/2 with an
//             exact two-product/two-sum tail for |x| < 2^20; Payne-Hanek with
//             a 1280-bit table of 2/pi above (out of line, wave-uniform skip);
//             Taylor polynomials on [-pi/4, pi/4] (sin to r^17, cos to r^16),
//             first-order correction for the reduced argument's tail.
//   exp     : n = rint(x log2 e), two-part ln 2 reduction with tail, Taylor to
//             r^13, scale by ldexp.
//   log     : x = 2^k m, m in [sqrt(2)/2, sqrt(2)); s = f/(2+f);
//             log(1+f) = f - (hfsq - s(hfsq + R(s^2))), R from 2 atanh(s).
//   float   : float arithmetic with float-float (unevaluated sum) steps for the
//             reduced argument and the leading terms; double only in the
//             reduction of trig arguments of magnitude 2^17 and up.
// Every multiply-add that is fused is an explicit __builtin_fma; the library
// is compiled with -ffp-contract=off so nothing else is fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aeth {
namespace dm {

__device__ __forceinline__ uint64_t bits_of(double x) { return __double_as_longlong(x); }
__device__ __forceinline__ double from_bits(uint64_t u) { return __longlong_as_double((long long)u); }

constexpr double kPio2_1 = 0x1.921fb54442d18p+0;
constexpr double kPio2_2 = 0x1.1a62633145c07p-54;
constexpr double kPio2_3 = -0x1.f1976b7ed8fbcp-110;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;

// Fraction bits of 2/pi, most significant first (computed with mpmath).
__device__ __constant__ static const uint64_t kTwoOverPiBits[20] = {
    0xA2F9836E4E441529ULL, 0xFC2757D1F534DDC0ULL, 0xDB6295993C439041ULL, 0xFE5163ABDEBBC561ULL,
    0xB7246E3A424DD2E0ULL, 0x06492EEA09D1921CULL, 0xFE1DEB1CB129A73EULL, 0xE88235F52EBB4484ULL,
    0xE99C7026B45F7E41ULL, 0x3991D639835339F4ULL, 0x9C845F8BBDF9283BULL, 0x1FF897FFDE05980FULL,
    0xEF2F118B5A0A6D1FULL, 0x6D367ECF27CB09B7ULL, 0x4F463F669E5FEA2DULL, 0x7527BAC7EBE5F17BULL,
    0x3D0739F78A5292EAULL, 0x6BFB5FB11F8D5D08ULL, 0x56033046FC7B6BABULL, 0xF0CFBC209AF4361DULL,
};

__device__ __forceinline__ uint64_t twopi_window(int j)
{
    if (j <= -64) return 0;
    if (j < 0) return kTwoOverPiBits[0] >> (-j);
    const int wi = j >> 6, sh = j & 63;
    uint64_t v = kTwoOverPiBits[wi] << sh;
    if (sh) v |= kTwoOverPiBits[wi + 1] >> (64 - sh);
    return v;
}

// |x| >= 2^20, finite: the quadrant q of |x| and (rh, rl) ~ |x| - q pi/2. Out of line (waves rarely take it)
// and returned by value, so the result comes back in registers and the callers keep no stack.
struct LargeReduction {
    double rh, rl;
    int q;
};
__device__ __noinline__ LargeReduction reduce_large(double ax)
{
    const uint64_t u = bits_of(ax);
    const int E = (int)(u >> 52) - 1075;
    const uint64_t M = (u & 0xFFFFFFFFFFFFFULL) | 0x10000000000000ULL;
    const int j = E - 2;
    const uint64_t w0 = twopi_window(j), w1 = twopi_window(j + 64), w2 = twopi_window(j + 128);
    // P = M (w0 2^128 + w1 2^64 + w2) mod 2^192, in three 64-bit words
    const uint64_t p2lo = M * w2, p2hi = __umul64hi(M, w2);
    const uint64_t p1lo = M * w1, p1hi = __umul64hi(M, w1);
    const uint64_t p0 = M * w0;
    uint64_t P0 = p2lo;
    uint64_t P1 = p2hi + p1lo;
    const uint64_t carry = P1 < p2hi ? 1u : 0u;
    uint64_t P2 = p1hi + p0 + carry;
    int q = (int)(P2 >> 62);
    P2 &= 0x3FFFFFFFFFFFFFFFULL;
    const bool neg = (P2 >> 61) != 0;
    if (neg) {
        q += 1;
        P0 = ~P0;
        P1 = ~P1;
        P2 = (~P2) & 0x3FFFFFFFFFFFFFFFULL;
        P0 += 1;
        if (P0 == 0) {
            P1 += 1;
            if (P1 == 0) P2 += 1;
        }
    }
    uint64_t G2 = (P2 << 2) | (P1 >> 62);
    uint64_t G1 = (P1 << 2) | (P0 >> 62);
    uint64_t G0 = P0 << 2;
    int shift = 0;
    if (G2 == 0) {
        G2 = G1;
        G1 = G0;
        G0 = 0;
        shift += 64;
    }
    const int lz = G2 ? __clzll((long long)G2) : 0;
    if (lz) {
        G2 = (G2 << lz) | (G1 >> (64 - lz));
        G1 = (G1 << lz) | (G0 >> (64 - lz));
    }
    shift += lz;
    const double fh = (double)(G2 & ~0x7FFULL);
    const double fl = (double)(G2 & 0x7FFULL) + (double)G1 * 0x1p-64;
    const double sh = __builtin_ldexp(fh, -64 - shift), sl = __builtin_ldexp(fl, -64 - shift);
    const double h = sh * kPio2_1;
    const double l = __builtin_fma(sh, kPio2_1, -h) + __builtin_fma(sh, kPio2_2, sl * kPio2_1);
    const double r = h + l;
    const double t = l - (r - h);
    return LargeReduction{neg ? -r : r, neg ? -t : t, q & 3};
}

__device__ __forceinline__ int reduce_pio2(double x, double* rh, double* rl)
{
    const double n = __builtin_rint(x * kTwoOverPi);
    const double r1 = __builtin_fma(-n, kPio2_1, x); // exact
    const double ph = -n * kPio2_2;
    const double pl = __builtin_fma(-n, kPio2_2, -ph); // exact
    const double s = r1 + ph;
    const double bb = s - r1;
    const double e = (r1 - (s - bb)) + (ph - bb);
    double t = e + pl;
    t = __builtin_fma(-n, kPio2_3, t);
    const double r = s + t;
    *rl = t - (r - s);
    *rh = r;
    int q = ((int)(long long)n) & 3;
    const double ax = __builtin_fabs(x);
    if (ax >= 0x1p20 && ax <= 0x1.fffffffffffffp+1023) {
        const LargeReduction lr = reduce_large(ax);
        double lh = lr.rh, ll = lr.rl;
        int ql = lr.q;
        if (x < 0) {
            lh = -lh;
            ll = -ll;
            ql = (-ql) & 3;
        }
        *rh = lh;
        *rl = ll;
        q = ql;
    }
    return q;
}

__device__ __forceinline__ double ksin(double r, double rl)
{
    const double z = r * r;
    double ps = 0x1.952c77030ad4ap-49;
    ps = __builtin_fma(ps, z, -0x1.ae7f3e733b81fp-41);
    ps = __builtin_fma(ps, z, 0x1.6124613a86d09p-33);
    ps = __builtin_fma(ps, z, -0x1.ae64567f544e4p-26);
    ps = __builtin_fma(ps, z, 0x1.71de3a556c734p-19);
    ps = __builtin_fma(ps, z, -0x1.a01a01a01a01ap-13);
    ps = __builtin_fma(ps, z, 0x1.1111111111111p-7);
    ps = __builtin_fma(ps, z, -0x1.5555555555555p-3);
    const double tail = __builtin_fma(-0.5 * z, rl, rl);
    return r + __builtin_fma(r * z, ps, tail);
}

__device__ __forceinline__ double kcos(double r, double rl)
{
    const double z = r * r;
    double pc = 0x1.ae7f3e733b81fp-45;
    pc = __builtin_fma(pc, z, -0x1.93974a8c07c9dp-37);
    pc = __builtin_fma(pc, z, 0x1.1eed8eff8d898p-29);
    pc = __builtin_fma(pc, z, -0x1.27e4fb7789f5cp-22);
    pc = __builtin_fma(pc, z, 0x1.a01a01a01a01ap-16);
    pc = __builtin_fma(pc, z, -0x1.6c16c16c16c17p-10);
    pc = __builtin_fma(pc, z, 0x1.5555555555555p-5);
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    const double corr = ((1.0 - w) - hz) + __builtin_fma(z * z, pc, -(r * rl));
    return w + corr;
}

__device__ __forceinline__ void sincos(double x, double* s, double* c)
{
    double rh, rl;
    const int q = reduce_pio2(x, &rh, &rl);
    const double sn = ksin(rh, rl), cs = kcos(rh, rl);
    double so = (q & 1) ? cs : sn;
    double co = (q & 1) ? sn : cs;
    so = (q & 2) ? -so : so;
    co = ((q + 1) & 2) ? -co : co;
    const double ax = __builtin_fabs(x);
    const bool tiny = ax < 0x1p-27;
    const bool bad = !(ax <= 0x1.fffffffffffffp+1023);
    so = tiny ? x : so;
    co = tiny ? 1.0 : co;
    so = bad ? x - x : so;
    co = bad ? x - x : co;
    *s = so;
    *c = co;
}

__device__ __forceinline__ double sin(double x)
{
    double rh, rl;
    const int q = reduce_pio2(x, &rh, &rl);
    // sin needs ksin in even quadrants and kcos in odd ones; both are cheap
    // next to a divergent branch across a wave.
    const double sn = ksin(rh, rl), cs = kcos(rh, rl);
    double so = (q & 1) ? cs : sn;
    so = (q & 2) ? -so : so;
    const double ax = __builtin_fabs(x);
    so = ax < 0x1p-27 ? x : so;
    so = !(ax <= 0x1.fffffffffffffp+1023) ? x - x : so;
    return so;
}

__device__ __forceinline__ double cos(double x)
{
    double rh, rl;
    const int q = reduce_pio2(x, &rh, &rl);
    const double sn = ksin(rh, rl), cs = kcos(rh, rl);
    double co = (q & 1) ? sn : cs;
    co = ((q + 1) & 2) ? -co : co;
    const double ax = __builtin_fabs(x);
    co = ax < 0x1p-27 ? 1.0 : co;
    co = !(ax <= 0x1.fffffffffffffp+1023) ? x - x : co;
    return co;
}

constexpr double kLog2e = 0x1.71547652b82fep+0;
constexpr double kLn2Hi = 0x1.62e42fefa39efp-1;
constexpr double kLn2Lo = 0x1.abc9e3b39803fp-56;
constexpr double kLn2Hi42 = 0x1.62e42fee00000p-1;
constexpr double kLn2Lo42 = 0x1.a39ef35793c76p-33;

__device__ __forceinline__ double exp(double x)
{
    const double xc = __builtin_fmin(__builtin_fmax(x, -746.0), 710.0);
    const double n = __builtin_rint(xc * kLog2e);
    const double r1 = __builtin_fma(-n, kLn2Hi, xc);
    const double r = __builtin_fma(-n, kLn2Lo, r1);
    const double rl = __builtin_fma(-n, kLn2Lo, r1 - r);
    double q = 0x1.6124613a86d09p-33;
    q = __builtin_fma(q, r, 0x1.1eed8eff8d898p-29);
    q = __builtin_fma(q, r, 0x1.ae64567f544e4p-26);
    q = __builtin_fma(q, r, 0x1.27e4fb7789f5cp-22);
    q = __builtin_fma(q, r, 0x1.71de3a556c734p-19);
    q = __builtin_fma(q, r, 0x1.a01a01a01a01ap-16);
    q = __builtin_fma(q, r, 0x1.a01a01a01a01ap-13);
    q = __builtin_fma(q, r, 0x1.6c16c16c16c17p-10);
    q = __builtin_fma(q, r, 0x1.1111111111111p-7);
    q = __builtin_fma(q, r, 0x1.5555555555555p-5);
    q = __builtin_fma(q, r, 0x1.5555555555555p-3);
    q = __builtin_fma(q, r, 0.5);
    const double u = __builtin_fma(r * r, q, r);
    const double p = 1.0 + (u + rl);
    double y = __builtin_ldexp(p, (int)n);
    y = x > 710.0 ? 0x1p1023 * 2.0 : y;
    y = x < -746.0 ? 0.0 : y;
    y = (x != x) ? x + x : y;
    return y;
}

__device__ __forceinline__ double log(double x)
{
    const bool sub = x < 0x1p-1022;
    const double xs = sub ? x * 0x1p54 : x;
    const uint64_t u = bits_of(xs);
    int k = (int)(u >> 52) - 1023 - (sub ? 54 : 0);
    uint64_t mb = (u & 0xFFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL;
    const bool big = mb > 0x3FF6A09E667F3BCCULL;
    mb = big ? mb - 0x0010000000000000ULL : mb;
    k += big ? 1 : 0;
    const double f = from_bits(mb) - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 0x1.642c8590b2164p-4;
    R = __builtin_fma(R, z, 0x1.8618618618618p-4);
    R = __builtin_fma(R, z, 0x1.af286bca1af28p-4);
    R = __builtin_fma(R, z, 0x1.e1e1e1e1e1e1ep-4);
    R = __builtin_fma(R, z, 0x1.1111111111111p-3);
    R = __builtin_fma(R, z, 0x1.3b13b13b13b14p-3);
    R = __builtin_fma(R, z, 0x1.745d1745d1746p-3);
    R = __builtin_fma(R, z, 0x1.c71c71c71c71cp-3);
    R = __builtin_fma(R, z, 0x1.2492492492492p-2);
    R = __builtin_fma(R, z, 0x1.999999999999ap-2);
    R = __builtin_fma(R, z, 0x1.5555555555555p-1);
    R = R * z;
    const double hfsq = 0.5 * f * f;
    const double dk = (double)k;
    double y = dk * kLn2Hi42 - ((hfsq - (s * (hfsq + R) + dk * kLn2Lo42)) - f);
    y = (x == 0.0) ? -__builtin_inf() : y;
    y = (x < 0.0) ? __builtin_nan("") : y;
    y = (x == __builtin_inf()) ? x : y;
    y = (x != x) ? x + x : y;
    return y;
}

// the double reduction of float arguments from 2^17: below 2^20 r = x - n P1 - n P2 (two FMAs; P1 + P2
// is pi/2 to 2^-107, so r is within 2^-80 absolute); above, the Payne-Hanek path, out of line.
__device__ __forceinline__ int reduce_pio2_f(double x, double* r)
{
    const double n = __builtin_rint(x * kTwoOverPi);
    double rr = __builtin_fma(-n, kPio2_1, x);
    rr = __builtin_fma(-n, kPio2_2, rr);
    int q = ((int)(long long)n) & 3;
    const double ax = __builtin_fabs(x);
    if (ax >= 0x1p20 && ax <= 0x1.fffffffffffffp+1023) {
        const LargeReduction lr = reduce_large(ax);
        double lh = lr.rh, ll = lr.rl;
        int ql = lr.q;
        if (x < 0) {
            lh = -lh;
            ll = -ll;
            ql = (-ql) & 3;
        }
        rr = lh + ll;
        q = ql;
    }
    *r = rr;
    return q;
}

// ---- float routines in float arithmetic. Products of two floats are split exactly with an FMA and
// float-float (unevaluated sum) steps carry the reduced argument and the leading terms, so each result is
// one rounding of a value good to about 2^-29: over every float argument at most 0.67 ulp for sin/cos,
// 0.68 for exp (0.78 where the result is subnormal and rounds twice) and 0.643 for log, more than 99.8%
// correctly rounded. Double appears only in the reduction of trig arguments of magnitude 2^17 and up.

__device__ __forceinline__ void f_fast2(float a, float b, float& s, float& e)
{ // |a| >= |b| or a == 0: s + e == a + b exactly
    const float ss = a + b;
    e = b - (ss - a);
    s = ss;
}

__device__ __forceinline__ void f_two2(float a, float b, float& s, float& e)
{ // any order: s + e == a + b exactly
    const float ss = a + b;
    const float bb = ss - a;
    e = (a - (ss - bb)) + (b - bb);
    s = ss;
}

// |x| < 2^17: n = rint(x 2/pi), r = x - n (C1 + C2 + C3), C1 + C2 + C3 = pi/2 to 2^-73 relative. x - n C1
// is exact (FMA) and n C2 is split exactly, so rh + rl is within 2^-55 of x - n pi/2.
__device__ __forceinline__ int reduce_ff(float x, float& rh, float& rl)
{
    const float n = __builtin_rintf(x * 0x1.45f306p-1f);
    const float r1 = __builtin_fmaf(-n, 0x1.921fb6p+0f, x);
    const float p = n * -0x1.777a5cp-25f;
    const float pe = __builtin_fmaf(n, -0x1.777a5cp-25f, -p);
    float h, e;
    f_two2(r1, -p, h, e);
    const float l = __builtin_fmaf(-n, -0x1.ee59dap-50f, e - pe);
    f_fast2(h, l, rh, rl);
    return ((int)n) & 3;
}

// 2^17 and above (and non-finite): the double reduction, its result split into two floats. Inline: the
// Payne-Hanek part it may reach is the one out-of-line call, a leaf returning in registers, so kernels keep
// no stack.
struct BigReductionF {
    float rh, rl;
    int q;
};
__device__ __forceinline__ BigReductionF reduce_big_f(float x)
{
    double r;
    const int q = reduce_pio2_f((double)x, &r);
    const float rh = (float)r;
    return BigReductionF{rh, (float)(r - (double)rh), q};
}

__device__ __forceinline__ int reduce_f32(float x, float& rh, float& rl)
{
    if (__builtin_fabsf(x) < 0x1p17f) return reduce_ff(x, rh, rl);
    const BigReductionF b = reduce_big_f(x);
    rh = b.rh;
    rl = b.rl;
    return b.q;
}

// sin(rh + rl) on |rh| <= pi/4: rh + (rh^3 P(rh^2) + rl (1 - rh^2/2)), Taylor to r^11; the leading
// -rh^3/6 also takes the rounding errors of rh^2 and rh^3. Returns the rounded value and, in lo, what it
// left out (the pair feeds tan's quotient).
__device__ __forceinline__ float ksinf(float rh, float rl, float& lo)
{
    const float z = rh * rh;
    float P = -0x1.ae6456p-26f;
    P = __builtin_fmaf(P, z, 0x1.71de3ap-19f);
    P = __builtin_fmaf(P, z, -0x1.a01a02p-13f);
    P = __builtin_fmaf(P, z, 0x1.111112p-7f);
    P = __builtin_fmaf(P, z, -0x1.555556p-3f);
    const float zl = __builtin_fmaf(rh, rh, -z);
    const float c = __builtin_fmaf(-0.5f * z, rl, rl);
    const float r3 = rh * z;
    const float r3l = __builtin_fmaf(rh, z, -r3);
    const float t = __builtin_fmaf(__builtin_fmaf(rh, zl, r3l), -0x1.555556p-3f, c);
    float h;
    f_fast2(rh, __builtin_fmaf(r3, P, t), h, lo);
    return h;
}

// cos(rh + rl) on |rh| <= pi/4: (1 - z/2) split exactly, + z^2 Q(z) - rh rl, Taylor to r^12; a pair as
// above
__device__ __forceinline__ float kcosf(float rh, float rl, float& lo)
{
    const float z = rh * rh;
    const float zl = __builtin_fmaf(rh, rh, -z);
    float Q = 0x1.1eed8ep-29f;
    Q = __builtin_fmaf(Q, z, -0x1.27e4fcp-22f);
    Q = __builtin_fmaf(Q, z, 0x1.a01a02p-16f);
    Q = __builtin_fmaf(Q, z, -0x1.6c16c2p-10f);
    Q = __builtin_fmaf(Q, z, 0x1.555556p-5f);
    const float hz = 0.5f * z;
    const float w = 1.0f - hz;
    const float we = (1.0f - w) - hz;
    float h;
    f_fast2(w, __builtin_fmaf(z * z, Q, __builtin_fmaf(-rh, rl, __builtin_fmaf(-0.5f, zl, we))), h, lo);
    return h;
}

__device__ __forceinline__ float sinf(float x)
{
    float rh, rl, lo;
    const int q = reduce_f32(x, rh, rl);
    float v = (q & 1) ? kcosf(rh, rl, lo) : ksinf(rh, rl, lo);
    v = (q & 2) ? -v : v;
    const float ax = __builtin_fabsf(x);
    v = ax < 0x1p-12f ? x : v; // sin x rounds to x; keeps -0
    v = !(ax <= 0x1.fffffep+127f) ? x - x : v;
    return v;
}

__device__ __forceinline__ float cosf(float x)
{
    float rh, rl, lo;
    const int q = reduce_f32(x, rh, rl);
    float v = (q & 1) ? ksinf(rh, rl, lo) : kcosf(rh, rl, lo);
    v = ((q + 1) & 2) ? -v : v;
    v = !(__builtin_fabsf(x) <= 0x1.fffffep+127f) ? x - x : v;
    return v;
}

// both from one reduction; each equals sinf / cosf
__device__ __forceinline__ void sincosf(float x, float* sp, float* cp)
{
    float rh, rl, lo;
    const int q = reduce_f32(x, rh, rl);
    const float sn = ksinf(rh, rl, lo), cs = kcosf(rh, rl, lo);
    float s = (q & 1) ? cs : sn;
    float c = (q & 1) ? sn : cs;
    s = (q & 2) ? -s : s;
    c = ((q + 1) & 2) ? -c : c;
    const float ax = __builtin_fabsf(x);
    s = ax < 0x1p-12f ? x : s;
    const bool bad = !(ax <= 0x1.fffffep+127f);
    *sp = bad ? x - x : s;
    *cp = bad ? x - x : c;
}

// e^x: clamp to [-104, 89] (beyond: 0 and +inf after rounding), n = rint(x log2 e), r = x - n ln2 as
// rh + rl (x - n L1 exact by FMA), e^r = (1 + rh) [split exactly] + rh^2 P(rh) + rl (1 + rh), Taylor to
// r^7, then the scaling by 2^n
__device__ __forceinline__ float expf(float x)
{
    const float xc = __builtin_fminf(__builtin_fmaxf(x, -104.0f), 89.0f);
    const float n = __builtin_rintf(xc * 0x1.715476p+0f);
    const float r1 = __builtin_fmaf(-n, 0x1.62e430p-1f, xc);
    const float rh = __builtin_fmaf(-n, -0x1.05c610p-29f, r1);
    const float rl = __builtin_fmaf(-n, -0x1.05c610p-29f, r1 - rh);
    float P = 0x1.a01a02p-13f;
    P = __builtin_fmaf(P, rh, 0x1.6c16c2p-10f);
    P = __builtin_fmaf(P, rh, 0x1.111112p-7f);
    P = __builtin_fmaf(P, rh, 0x1.555556p-5f);
    P = __builtin_fmaf(P, rh, 0x1.555556p-3f);
    P = __builtin_fmaf(P, rh, 0.5f);
    const float s = 1.0f + rh;
    const float se = (1.0f - s) + rh;
    const float tail = __builtin_fmaf(rh * rh, P, __builtin_fmaf(rl, rh, se + rl));
    float y = __builtin_ldexpf(s + tail, (int)n);
    y = (x != x) ? x + x : y;
    return y;
}

// log x = k ln2 + log(1+f), f = m - 1 with m in [sqrt2/2, sqrt2); log(1+f) = f - f^2/2 + s (f^2/2 + R),
// s = f/(2+f), R to s^10; f - f^2/2 and k K1 + (f - f^2/2) are exact float-float sums (K1 = ln2 to 16
// bits, so k K1 is exact)
__device__ __forceinline__ float logf(float x)
{
    const bool sub = __builtin_fabsf(x) < 0x1p-126f;
    const float xs = sub ? x * 0x1p25f : x;
    const uint32_t u = __builtin_bit_cast(uint32_t, xs);
    int k = (int)(u >> 23) - 127 - (sub ? 25 : 0);
    uint32_t mb = (u & 0x7FFFFFu) | 0x3F800000u;
    const bool big = mb > 0x3FB504F3u;
    mb = big ? mb - 0x00800000u : mb;
    k += big ? 1 : 0;
    const float f = __builtin_bit_cast(float, mb) - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float R = 0x1.745d18p-3f;
    R = __builtin_fmaf(R, z, 0x1.c71c72p-3f);
    R = __builtin_fmaf(R, z, 0x1.24924ap-2f);
    R = __builtin_fmaf(R, z, 0x1.99999ap-2f);
    R = __builtin_fmaf(R, z, 0x1.555556p-1f);
    R = R * z;
    const float ffh = f * f;
    const float ffl = __builtin_fmaf(f, f, -ffh);
    const float hf = 0.5f * ffh;
    const float d = f - hf;
    const float de = (f - d) - hf;
    const float fk = (float)k;
    float sh, se;
    f_two2(fk * 0x1.62e400p-1f, d, sh, se);
    float y = sh + __builtin_fmaf(s, hf + R,
                                  __builtin_fmaf(fk, 0x1.7f7d1cp-20f, __builtin_fmaf(-0.5f, ffl, de) + se));
    y = (x == 0.0f) ? -__builtin_inff() : y;
    y = (x < 0.0f) ? __builtin_nanf("") : y;
    y = (x == __builtin_inff()) ? x : y;
    y = (x != x) ? x + x : y;
    return y;
}

// ---- the rest of the reference's elementwise set (expr/nodes/Elementwise.h:200-851): tan, atan,
// asin, acos, tanh, pow, and the exact rsqrt / sign / fmin / fmax / ordered max. Double-double
// (unevaluated sum) steps wherever a plain evaluation would lose more than half an ulp; constants from
// gen_math_constants.py. Each is <= 1 ulp from the correctly rounded value on the reference's MPFR rows.

__device__ __forceinline__ void two_sum(double a, double b, double* s, double* e)
{
    const double ss = a + b;
    const double bb = ss - a;
    *e = (a - (ss - bb)) + (b - bb);
    *s = ss;
}

// |a| >= |b| (or a == 0)
__device__ __forceinline__ void fast_two_sum(double a, double b, double* s, double* e)
{
    const double ss = a + b;
    *e = b - (ss - a);
    *s = ss;
}

// atan(k/16) and pi/2 - atan(k/16), k = 0..16, as (hi, lo)
__device__ __constant__ static const double kAtanK[17][2] = {
    {0x0.0p+0, 0x0.0p+0},
    {0x1.ff55bb72cfdeap-5, -0x1.c934d86d23f1dp-60},
    {0x1.fd5ba9aac2f6ep-4, -0x1.cd37686760c17p-59},
    {0x1.7b97b4bce5b02p-3, 0x1.347b0b4f881cap-58},
    {0x1.f5b75f92c80ddp-3, 0x1.8ab6e3cf7afbdp-57},
    {0x1.362773707ebccp-2, -0x1.963a544b672d8p-57},
    {0x1.6f61941e4def1p-2, -0x1.c63aae6f6e918p-56},
    {0x1.a64eec3cc23fdp-2, -0x1.24dec1b50b7ffp-56},
    {0x1.dac670561bb4fp-2, 0x1.a2b7f222f65e2p-56},
    {0x1.0657e94db30d0p-1, -0x1.d5b495f6349e6p-56},
    {0x1.1e00babdefeb4p-1, -0x1.928df287a668fp-58},
    {0x1.345f01cce37bbp-1, 0x1.1021137c71102p-55},
    {0x1.4978fa3269ee1p-1, 0x1.2419a87f2a458p-56},
    {0x1.5d58987169b18p-1, 0x1.0028e4bc5e7cap-57},
    {0x1.700a7c5784634p-1, -0x1.8c34d25aadef6p-56},
    {0x1.819d0b7158a4dp-1, -0x1.bf76229d3b917p-56},
    {0x1.921fb54442d18p-1, 0x1.1a62633145c07p-55},
};
__device__ __constant__ static const double kAtanCoK[17][2] = {
    {0x1.921fb54442d18p+0, 0x1.1a62633145c07p-54},
    {0x1.82250768ac529p+0, -0x1.e78c96d05afcbp-58},
    {0x1.7249faa996a21p+0, 0x1.a8cc1e7480c68p-54},
    {0x1.62acbeaca61b8p+0, 0x1.c6ac9f134fa91p-60},
    {0x1.5368c951e9cfdp+0, -0x1.96f47948a99f1p-54},
    {0x1.4495d86823225p+0, 0x1.4d29adbab2a62p-54},
    {0x1.3647503caf55cp+0, 0x1.17e21d9a42c9ap-55},
    {0x1.288bfa3512419p+0, 0x1.8e684e7a2281bp-56},
    {0x1.1b6e192ebbe44p+0, 0x1.b1b466a88828ep-54},
    {0x1.0ef3c09d694b0p+0, 0x1.8fcf88aed2e80p-54},
    {0x1.031f57e54adbep+0, 0x1.338b4259c0270p-54},
    {0x1.efe068bba2275p-1, 0x1.24a3b2e61a70bp-55},
    {0x1.dac670561bb4fp-1, 0x1.a2b7f222f65e2p-55},
    {0x1.c6e6d2171bf18p-1, 0x1.f4ba8d3373e1bp-55},
    {0x1.b434ee31013fdp-1, -0x1.0520d0701d877p-55},
    {0x1.a2a25f172cfe4p-1, -0x1.d700509dad6cep-56},
    {0x1.921fb54442d18p-1, 0x1.1a62633145c07p-55},
};
constexpr double kPio2H = 0x1.921fb54442d18p+0, kPio2L = 0x1.1a62633145c07p-54;
constexpr double kPiH = 0x1.921fb54442d18p+1, kPiL = 0x1.1a62633145c07p-53;

// tan: sin and cos of the reduced argument as unevaluated sums (ksin / kcos before their last
// addition), then one double-double quotient; odd quadrants give -cos/sin.
__device__ __forceinline__ double tan(double x)
{
    const double ax = __builtin_fabs(x);
    if (!(ax <= 0x1.fffffffffffffp+1023)) return x - x;
    if (ax < 0x1p-27) return x;
    double rh, rl;
    const int q = reduce_pio2(x, &rh, &rl);
    const double z = rh * rh;
    double ps = 0x1.952c77030ad4ap-49;
    ps = __builtin_fma(ps, z, -0x1.ae7f3e733b81fp-41);
    ps = __builtin_fma(ps, z, 0x1.6124613a86d09p-33);
    ps = __builtin_fma(ps, z, -0x1.ae64567f544e4p-26);
    ps = __builtin_fma(ps, z, 0x1.71de3a556c734p-19);
    ps = __builtin_fma(ps, z, -0x1.a01a01a01a01ap-13);
    ps = __builtin_fma(ps, z, 0x1.1111111111111p-7);
    ps = __builtin_fma(ps, z, -0x1.5555555555555p-3);
    const double stail = __builtin_fma(-0.5 * z, rl, rl);
    double sh, sl;
    fast_two_sum(rh, __builtin_fma(rh * z, ps, stail), &sh, &sl);
    double pc = 0x1.ae7f3e733b81fp-45;
    pc = __builtin_fma(pc, z, -0x1.93974a8c07c9dp-37);
    pc = __builtin_fma(pc, z, 0x1.1eed8eff8d898p-29);
    pc = __builtin_fma(pc, z, -0x1.27e4fb7789f5cp-22);
    pc = __builtin_fma(pc, z, 0x1.a01a01a01a01ap-16);
    pc = __builtin_fma(pc, z, -0x1.6c16c16c16c17p-10);
    pc = __builtin_fma(pc, z, 0x1.5555555555555p-5);
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    double ch, cl;
    fast_two_sum(w, ((1.0 - w) - hz) + __builtin_fma(z * z, pc, -(rh * rl)), &ch, &cl);
    const bool odd = (q & 1) != 0;
    const double nh = odd ? -ch : sh, nl = odd ? -cl : sl;
    const double dh = odd ? sh : ch, dl = odd ? sl : cl;
    const double th = nh / dh;
    const double tl = (__builtin_fma(-th, dh, nh) + nl - th * dl) / dh;
    return th + tl;
}

// atan(zh + zl) for 0 <= zh <= 1 (zl a tail) as hi + lo, or pi/2 minus it (co): atan(c) from the table
// plus atan(t), t = (z - c) / (1 + z c), c = k/16 the nearest sixteenth, |t| <= 1/32; atan(t) to t^13.
__device__ __forceinline__ void atan_core(double zh, double zl, bool co, double* hi, double* lo)
{
    double kd = __builtin_rint(zh * 16.0);
    kd = (kd >= 0.0 && kd <= 16.0) ? kd : 0.0;
    const int k = (int)kd;
    const double c = kd * 0.0625;
    const double num = zh - c; // exact
    double nh, nl;
    two_sum(num, zl, &nh, &nl);
    const double ph = zh * c;
    const double pe = __builtin_fma(zh, c, -ph);
    double dh, dl;
    fast_two_sum(1.0, ph, &dh, &dl);
    dl += pe;
    const double th = nh / dh;
    const double tl = (__builtin_fma(-th, dh, nh) + nl - th * dl) / dh;
    const double w = th * th;
    double P = 1.0 / 13.0;
    P = __builtin_fma(P, w, -1.0 / 11.0);
    P = __builtin_fma(P, w, 1.0 / 9.0);
    P = __builtin_fma(P, w, -1.0 / 7.0);
    P = __builtin_fma(P, w, 1.0 / 5.0);
    P = __builtin_fma(P, w, -1.0 / 3.0);
    const double tail = __builtin_fma(th * w, P, tl); // atan(t) - th
    const double bh = co ? kAtanCoK[k][0] : kAtanK[k][0];
    const double bl = co ? kAtanCoK[k][1] : kAtanK[k][1];
    double s, e;
    two_sum(bh, co ? -th : th, &s, &e);
    fast_two_sum(s, e + (bl + (co ? -tail : tail)), hi, lo);
}

__device__ __forceinline__ double atan(double x)
{
    if (x != x) return x + x;
    const double a = __builtin_fabs(x);
    const bool big = a > 1.0;
    double zh = a, zl = 0.0;
    if (big) {
        zh = 1.0 / a;
        zl = (a <= 0x1.fffffffffffffp+1023) ? __builtin_fma(-zh, a, 1.0) / a : 0.0;
    }
    double h, l;
    atan_core(zh, zl, big, &h, &l);
    const double r = h + l;
    return copysign(r, x);
}

// asin(a) for 0 <= a <= 1/2 as hi + lo: atan(a / sqrt(1 - a^2)), every step an unevaluated sum
__device__ __forceinline__ void asin_small(double a, double* hi, double* lo)
{
    const double p = a * a;
    const double pe = __builtin_fma(a, a, -p);
    double wh, wl;
    fast_two_sum(1.0, -p, &wh, &wl);
    wl -= pe;
    const double sh = __builtin_sqrt(wh);
    const double sl = (__builtin_fma(-sh, sh, wh) + wl) / (2.0 * sh);
    const double zh = a / sh;
    const double zl = (__builtin_fma(-zh, sh, a) - zh * sl) / sh;
    atan_core(zh, zl, false, hi, lo);
}

// asin(s) for s = sqrt(v), 0 <= v <= 1/4 (v exact), as hi + lo
__device__ __forceinline__ void asin_of_root(double v, double* hi, double* lo)
{
    const double sh = __builtin_sqrt(v);
    const double sl = sh > 0.0 ? __builtin_fma(-sh, sh, v) / (2.0 * sh) : 0.0;
    double wh, wl;
    fast_two_sum(1.0, -v, &wh, &wl); // 1 - s^2 = 1 - v
    const double qh = __builtin_sqrt(wh);
    const double ql = (__builtin_fma(-qh, qh, wh) + wl) / (2.0 * qh);
    const double zh = sh / qh;
    const double zl = (__builtin_fma(-zh, qh, sh) + sl - zh * ql) / qh;
    atan_core(zh, zl, false, hi, lo);
}

__device__ __forceinline__ double asin(double x)
{
    const double a = __builtin_fabs(x);
    if (!(a <= 1.0)) return (x - x) / (x - x);
    double h, l, r;
    if (a <= 0.5) {
        asin_small(a, &h, &l);
        r = h + l;
    } else {
        // asin(a) = pi/2 - 2 asin(sqrt((1 - a) / 2))
        asin_of_root((1.0 - a) * 0.5, &h, &l);
        double s, e;
        two_sum(kPio2H, -2.0 * h, &s, &e);
        r = s + (e + (kPio2L - 2.0 * l));
    }
    return copysign(r, x);
}

__device__ __forceinline__ double acos(double x)
{
    const double a = __builtin_fabs(x);
    if (!(a <= 1.0)) return (x - x) / (x - x);
    double h, l, s, e;
    if (a <= 0.5) {
        asin_small(a, &h, &l); // pi/2 - asin(x)
        if (x < 0) {
            h = -h;
            l = -l;
        }
        two_sum(kPio2H, -h, &s, &e);
        return s + (e + (kPio2L - l));
    }
    asin_of_root((1.0 - a) * 0.5, &h, &l);
    if (x > 0) return 2.0 * h + 2.0 * l; // 2 asin(sqrt((1 - x) / 2))
    two_sum(kPiH, -2.0 * h, &s, &e);     // pi - 2 asin(sqrt((1 + x) / 2))
    return s + (e + (kPiL - 2.0 * l));
}

// tanh: |x| < 0.625 its Taylor series to x^43; above 1 - 2 / (exp(2|x|) + 1)
__device__ __forceinline__ double tanh(double x)
{
    const double a = __builtin_fabs(x);
    double r;
    if (a < 0.625) {
        const double w = a * a;
        double P = -0x1.42ba1a349b490p-28;
        P = __builtin_fma(P, w, 0x1.8e25ff9327e2cp-27);
        P = __builtin_fma(P, w, -0x1.eb3229047434cp-26);
        P = __builtin_fma(P, w, 0x1.2efe8db3aff1fp-24);
        P = __builtin_fma(P, w, -0x1.75cde6563fed9p-23);
        P = __builtin_fma(P, w, 0x1.cd299de4ae6bbp-22);
        P = __builtin_fma(P, w, -0x1.1c77df95c1c0dp-20);
        P = __builtin_fma(P, w, 0x1.5ef2da474e5b7p-19);
        P = __builtin_fma(P, w, -0x1.b0f72d3ee24e9p-18);
        P = __builtin_fma(P, w, 0x1.0b132d39a6050p-16);
        P = __builtin_fma(P, w, -0x1.497d8eea25259p-15);
        P = __builtin_fma(P, w, 0x1.967e18afcafadp-14);
        P = __builtin_fma(P, w, -0x1.f57d7734d1664p-13);
        P = __builtin_fma(P, w, 0x1.3558248036744p-11);
        P = __builtin_fma(P, w, -0x1.7da36452b75e3p-10);
        P = __builtin_fma(P, w, 0x1.d6d3d0e157de0p-9);
        P = __builtin_fma(P, w, -0x1.226e355e6c23dp-7);
        P = __builtin_fma(P, w, 0x1.664f4882c10fap-6);
        P = __builtin_fma(P, w, -0x1.ba1ba1ba1ba1cp-5);
        P = __builtin_fma(P, w, 0x1.1111111111111p-3);
        P = __builtin_fma(P, w, -0x1.5555555555555p-2);
        r = __builtin_fma(a * w, P, a);
    } else {
        const double e = exp(2.0 * a);
        r = 1.0 - 2.0 / (e + 1.0);
    }
    return copysign(r, x);
}

// log(a) for finite a > 0 as hi + lo to about 2^-66 relative: a = 2^k m, m in [sqrt(2)/2, sqrt(2)),
// s = f / (2 + f) as an unevaluated sum, log m = 2s + 2s^3/3 + 2s^5 Q(s^2) with the cubic term as an
// unevaluated product, k ln 2 in two parts.
__device__ __forceinline__ void log_dd(double a, double* hi, double* lo)
{
    const bool sub = a < 0x1p-1022;
    const double as = sub ? a * 0x1p54 : a;
    const uint64_t u = bits_of(as);
    int k = (int)(u >> 52) - 1023 - (sub ? 54 : 0);
    uint64_t mb = (u & 0xFFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL;
    const bool half = mb > 0x3FF6A09E667F3BCCULL;
    mb = half ? mb - 0x0010000000000000ULL : mb;
    k += half ? 1 : 0;
    const double f = from_bits(mb) - 1.0; // exact
    double dh, dl;
    fast_two_sum(2.0, f, &dh, &dl);
    const double sh = f / dh;
    const double sl = (__builtin_fma(-sh, dh, f) - sh * dl) / dh;
    const double w = sh * sh;
    const double wl = __builtin_fma(sh, sh, -w);
    const double c3 = w * sh;
    const double c3l = __builtin_fma(w, sh, -c3) + wl * sh;
    const double m3 = c3 * 0x1.5555555555555p-1; // s^3 2/3
    const double m3l = __builtin_fma(c3, 0x1.5555555555555p-1, -m3) +
                       (c3l * 0x1.5555555555555p-1 + c3 * 0x1.5555555555555p-55);
    double Q = 2.0 / 25.0;
    Q = __builtin_fma(Q, w, 2.0 / 23.0);
    Q = __builtin_fma(Q, w, 2.0 / 21.0);
    Q = __builtin_fma(Q, w, 2.0 / 19.0);
    Q = __builtin_fma(Q, w, 2.0 / 17.0);
    Q = __builtin_fma(Q, w, 2.0 / 15.0);
    Q = __builtin_fma(Q, w, 2.0 / 13.0);
    Q = __builtin_fma(Q, w, 2.0 / 11.0);
    Q = __builtin_fma(Q, w, 2.0 / 9.0);
    Q = __builtin_fma(Q, w, 2.0 / 7.0);
    Q = __builtin_fma(Q, w, 2.0 / 5.0);
    const double tail = c3 * w * Q; // 2 s^5 / 5 + ...
    double g, ge;
    two_sum(2.0 * sh, m3, &g, &ge);
    // 2 sl (1 + s^2): the tail of s through the linear and cubic terms
    const double low = ge + (__builtin_fma(2.0 * w, sl, 2.0 * sl) + (m3l + tail));
    const double dk = (double)k;
    double s1, e1;
    two_sum(dk * kLn2Hi42, g, &s1, &e1);
    fast_two_sum(s1, e1 + (low + dk * kLn2Lo42), hi, lo);
}

// exp(xh + xl), |xl| tiny next to xh, xh in [-746, 710]: exp's reduction with the tail added
__device__ __forceinline__ double exp_dd(double xh, double xl)
{
    const double n = __builtin_rint(xh * kLog2e);
    const double r1 = __builtin_fma(-n, kLn2Hi, xh);
    const double r = __builtin_fma(-n, kLn2Lo, r1);
    const double rl = __builtin_fma(-n, kLn2Lo, r1 - r) + xl;
    double q = 0x1.6124613a86d09p-33;
    q = __builtin_fma(q, r, 0x1.1eed8eff8d898p-29);
    q = __builtin_fma(q, r, 0x1.ae64567f544e4p-26);
    q = __builtin_fma(q, r, 0x1.27e4fb7789f5cp-22);
    q = __builtin_fma(q, r, 0x1.71de3a556c734p-19);
    q = __builtin_fma(q, r, 0x1.a01a01a01a01ap-16);
    q = __builtin_fma(q, r, 0x1.a01a01a01a01ap-13);
    q = __builtin_fma(q, r, 0x1.6c16c16c16c17p-10);
    q = __builtin_fma(q, r, 0x1.1111111111111p-7);
    q = __builtin_fma(q, r, 0x1.5555555555555p-5);
    q = __builtin_fma(q, r, 0x1.5555555555555p-3);
    q = __builtin_fma(q, r, 0.5);
    const double uu = __builtin_fma(r * r, q, r);
    // exp(r + rl) = (1 + uu)(1 + rl) to first order in rl: the tail carries the argument's low part (up to
    // 2^-53 |xh|), so its product with uu is kept
    const double p = 1.0 + (uu + __builtin_fma(rl, uu, rl));
    return __builtin_ldexp(p, (int)n);
}

// 0: not an integer (or not finite); 1: an even integer; 2: an odd integer
__device__ __forceinline__ int pow_int_kind(double y)
{
    if (!(__builtin_fabs(y) <= 0x1.fffffffffffffp+1023)) return 0;
    if (__builtin_rint(y) != y) return 0;
    if (__builtin_fabs(y) >= 0x1p53) return 1;
    const double h = y * 0.5;
    return __builtin_rint(h) != h ? 2 : 1;
}

// pow with C99 Annex F special cases; |x|^y = exp(y log|x|) with both steps as unevaluated sums
__device__ __forceinline__ double pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (x == 1.0) return 1.0;
    if (x != x || y != y) return x + y;
    const int ik = pow_int_kind(y);
    const double ax = __builtin_fabs(x), ay = __builtin_fabs(y);
    if (ay > 0x1.fffffffffffffp+1023) {
        if (ax == 1.0) return 1.0;
        return ((ax < 1.0) == (y < 0)) ? __builtin_inf() : 0.0;
    }
    const bool neg = x < 0 || (x == 0.0 && __builtin_signbit(x));
    const double sgn = (neg && ik == 2) ? -1.0 : 1.0;
    if (x == 0.0) return y < 0 ? sgn / 0.0 : sgn * 0.0;
    if (ax > 0x1.fffffffffffffp+1023) return y < 0 ? sgn * 0.0 : sgn / 0.0;
    if (x < 0 && ik == 0) return (x - x) / (x - x);
    double lh, ll;
    log_dd(ax, &lh, &ll);
    const double ph = y * lh;
    const double pl = __builtin_fma(y, lh, -ph) + y * ll;
    if (ph > 710.0) return sgn / 0.0;
    if (ph < -746.0) return sgn * 0.0;
    return sgn * exp_dd(ph, pl);
}

// 1 / sqrt(x), two correctly rounded operations (the reference's host rsqrt, math/root.h:74)
__device__ __forceinline__ double rsqrt(double x) { return 1.0 / __builtin_sqrt(x); }
__device__ __forceinline__ float rsqrtf(float x) { return 1.0f / __builtin_sqrtf(x); }
// math/math.h:161-171: 1, -1, or 0 (zeros and NaN)
template<class T>
__device__ __forceinline__ T sign(T x)
{
    return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0));
}
// fmin / fmax as glibc's generic form: the other operand when one is NaN; on equal values (+-0) the second
template<class T>
__device__ __forceinline__ T fmin(T a, T b)
{
    return (a < b || b != b) ? a : b;
}
template<class T>
__device__ __forceinline__ T fmax(T a, T b)
{
    return (a > b || b != b) ? a : b;
}
// the reductions' ordered max (expr/Reduce.h MaxOp): a > b ? a : b
template<class T>
__device__ __forceinline__ T selgt(T a, T b)
{
    return a > b ? a : b;
}

// ---- float forms of tan, atan, asin, acos, tanh in float arithmetic, on the float routines' pieces:
// pair quotients with an exact FMA residual and a table of atan(k/16) as float pairs. Over every float
// argument (|x| < 2^17 for tan): at most 0.783 ulp for tan, 0.501 for atan/asin/acos, 0.640 for tanh.

// (nh + nl) / (dh + dl) as a pair: q = nh/dh, its residual nh - q dh exact by FMA
__device__ __forceinline__ void f_div2(float nh, float nl, float dh, float dl, float& h, float& l)
{
    const float q = nh / dh;
    const float r = __builtin_fmaf(-q, dl, __builtin_fmaf(-q, dh, nh) + nl);
    f_fast2(q, r / dh, h, l);
}

__device__ __forceinline__ float tanf(float x)
{
    float rh, rl, sl, cl, h, l;
    const int q = reduce_f32(x, rh, rl);
    const float sh = ksinf(rh, rl, sl), ch = kcosf(rh, rl, cl);
    if (q & 1) f_div2(-ch, -cl, sh, sl, h, l);
    else f_div2(sh, sl, ch, cl, h, l);
    float v = h + l;
    const float ax = __builtin_fabsf(x);
    v = ax < 0x1p-12f ? x : v;
    v = !(ax <= 0x1.fffffep+127f) ? x - x : v;
    return v;
}

// atan(k/16), k = 0..16, as float pairs (mpmath)
__device__ __constant__ static const float kAtan16[17][2] = {
    {0x0p+0f, 0x0p+0f}, {0x1.ff55bcp-5f, -0x1.1a6042p-30f}, {0x1.fd5baap-4f, -0x1.54f424p-30f},
    {0x1.7b97b4p-3f, 0x1.79cb60p-28f}, {0x1.f5b760p-3f, -0x1.b4dfc8p-29f}, {0x1.362774p-2f, -0x1.1f0286p-27f},
    {0x1.6f6194p-2f, 0x1.e4def0p-30f}, {0x1.a64eecp-2f, 0x1.e611fep-29f}, {0x1.dac670p-2f, 0x1.586ed4p-28f},
    {0x1.0657eap-1f, -0x1.6499e6p-26f}, {0x1.1e00bap-1f, 0x1.7bdfd6p-26f}, {0x1.345f02p-1f, -0x1.98e422p-28f},
    {0x1.4978fap-1f, 0x1.934f70p-28f}, {0x1.5d5898p-1f, 0x1.c5a6c6p-27f}, {0x1.700a7cp-1f, 0x1.5e118cp-27f},
    {0x1.819d0cp-1f, -0x1.1d4eb6p-26f}, {0x1.921fb6p-1f, -0x1.777a5cp-26f}};

// atan(th + tl), 0 <= t <= 1, as a pair: c = k/16 nearest t, u = (t - c)/(1 + t c) (|u| <= 1/32, t - c
// exact), atan(k/16) + u - u^3/3 + u^5/5 - u^7/7. NaN reads entry 0.
__device__ __forceinline__ void atan01_f(float th, float tl, float& yh, float& yl)
{
    const float kf = __builtin_rintf(th * 16.0f);
    const int k = (kf >= 0.0f && kf <= 16.0f) ? (int)kf : 0;
    const float c = (float)k * 0.0625f;
    float nh, nl, uh, ul, s, e;
    f_fast2(th - c, tl, nh, nl);
    const float dh = __builtin_fmaf(th, c, 1.0f);
    const float dl = __builtin_fmaf(tl, c, __builtin_fmaf(th, c, 1.0f - dh));
    f_div2(nh, nl, dh, dl, uh, ul);
    const float u2 = uh * uh;
    float p = __builtin_fmaf(u2, -0x1.24924ap-3f, 0x1.99999ap-3f);
    p = __builtin_fmaf(p, u2, -0x1.555556p-2f);
    const float tail = __builtin_fmaf(uh * u2, p, ul) + kAtan16[k][1];
    f_fast2(kAtan16[k][0], uh, s, e);
    f_fast2(s, e + tail, yh, yl);
}

// c_hi + c_lo - (yh + yl), rounded once (c = pi/2 or pi as a float pair)
__device__ __forceinline__ float f_minus(float chi, float clo, float yh, float yl)
{
    float s, e;
    f_two2(chi, -yh, s, e);
    return s + ((e + clo) - yl);
}

__device__ __forceinline__ float atanf(float x)
{
    const float ax = __builtin_fabsf(x);
    const bool inv = ax > 1.0f;
    const float th = inv ? 1.0f / ax : ax;
    const float tl = (inv && ax <= 0x1.fffffep+127f) ? __builtin_fmaf(-th, ax, 1.0f) * th : 0.0f;
    float yh, yl;
    atan01_f(th, tl, yh, yl);
    float y = inv ? f_minus(0x1.921fb6p+0f, -0x1.777a5cp-25f, yh, yl) : yh + yl;
    y = __builtin_copysignf(y, x);
    y = (x != x) ? x + x : y;
    return y;
}

// atan of the smaller of ax / sqrt(1 - x^2) and its inverse (pairs; 1 - x^2 exact as a pair); far for
// the inverse (ax > sqrt2/2)
__device__ __forceinline__ void asin_core_f(float ax, float& ah, float& al, bool& far)
{
    const float x2h = ax * ax;
    const float x2l = __builtin_fmaf(ax, ax, -x2h);
    float mh, me, ml, th, tl;
    f_two2(1.0f, -x2h, mh, me);
    f_fast2(mh, me - x2l, mh, ml);
    const float sh = __builtin_sqrtf(mh);
    const float sl = mh > 0.0f ? (__builtin_fmaf(-sh, sh, mh) + ml) / (2.0f * sh) : 0.0f;
    far = ax > 0x1.6a09e6p-1f;
    if (far) f_div2(sh, sl, ax, 0.0f, th, tl);
    else f_div2(ax, 0.0f, sh, sl, th, tl);
    atan01_f(th, tl, ah, al);
}

__device__ __forceinline__ float asinf(float x)
{
    const float ax = __builtin_fabsf(x);
    float ah, al;
    bool far;
    asin_core_f(ax, ah, al, far);
    float y = far ? f_minus(0x1.921fb6p+0f, -0x1.777a5cp-25f, ah, al) : ah + al;
    y = __builtin_copysignf(y, x);
    y = (ax > 1.0f) ? __builtin_nanf("") : y;
    y = (x != x) ? x + x : y;
    return y;
}

__device__ __forceinline__ float acosf(float x)
{
    const float ax = __builtin_fabsf(x);
    float ah, al;
    bool far;
    asin_core_f(ax, ah, al, far);
    float y;
    if (far) y = x > 0.0f ? ah + al : f_minus(0x1.921fb6p+1f, -0x1.777a5cp-24f, ah, al);
    else y = x < 0.0f ? f_minus(0x1.921fb6p+0f, -0x1.777a5cp-25f, -ah, -al)
                      : f_minus(0x1.921fb6p+0f, -0x1.777a5cp-25f, ah, al);
    y = (ax > 1.0f) ? __builtin_nanf("") : y;
    y = (x != x) ? x + x : y;
    return y;
}

// e^x as a pair for |x| <= 20: expf's steps without the clamp, the pair scaled by 2^n exactly
__device__ __forceinline__ void expf_pair(float x, float& h, float& l)
{
    const float n = __builtin_rintf(x * 0x1.715476p+0f);
    const float r1 = __builtin_fmaf(-n, 0x1.62e430p-1f, x);
    const float rh = __builtin_fmaf(-n, -0x1.05c610p-29f, r1);
    const float rl = __builtin_fmaf(-n, -0x1.05c610p-29f, r1 - rh);
    float P = 0x1.a01a02p-13f;
    P = __builtin_fmaf(P, rh, 0x1.6c16c2p-10f);
    P = __builtin_fmaf(P, rh, 0x1.111112p-7f);
    P = __builtin_fmaf(P, rh, 0x1.555556p-5f);
    P = __builtin_fmaf(P, rh, 0x1.555556p-3f);
    P = __builtin_fmaf(P, rh, 0.5f);
    const float s = 1.0f + rh;
    const float se = (1.0f - s) + rh;
    float eh, el;
    f_fast2(s, __builtin_fmaf(rh * rh, P, __builtin_fmaf(rl, rh, se + rl)), eh, el);
    h = __builtin_ldexpf(eh, (int)n);
    l = __builtin_ldexpf(el, (int)n);
}

// tanh: below 0.55 the Taylor series to x^21 (x^3/3 with the roundings of x^2 and x^3), above
// 1 - 2/(e^{2|x|} + 1) in pairs; |x| capped at 10 (tanh rounds to 1 from 9.02)
__device__ __forceinline__ float tanhf(float x)
{
    const float ax = __builtin_fabsf(x);
    float y;
    if (ax < 0.55f) {
        const float z = ax * ax;
        const float zl = __builtin_fmaf(ax, ax, -z);
        float P = 0x1.967e18p-14f;
        P = __builtin_fmaf(P, z, -0x1.f57d78p-13f);
        P = __builtin_fmaf(P, z, 0x1.355824p-11f);
        P = __builtin_fmaf(P, z, -0x1.7da364p-10f);
        P = __builtin_fmaf(P, z, 0x1.d6d3d0p-9f);
        P = __builtin_fmaf(P, z, -0x1.226e36p-7f);
        P = __builtin_fmaf(P, z, 0x1.664f48p-6f);
        P = __builtin_fmaf(P, z, -0x1.ba1ba2p-5f);
        P = __builtin_fmaf(P, z, 0x1.111112p-3f);
        P = P * z;
        const float x3 = ax * z;
        const float x3l = __builtin_fmaf(ax, z, -x3);
        const float t = __builtin_fmaf(__builtin_fmaf(ax, zl, x3l), -0x1.555556p-2f, x3 * P);
        y = ax + __builtin_fmaf(x3, -0x1.555556p-2f, t);
    } else {
        float eh, el, dh, de, qh, ql;
        expf_pair(2.0f * __builtin_fminf(ax, 10.0f), eh, el);
        f_two2(eh, 1.0f, dh, de);
        f_fast2(dh, de + el, dh, de);
        f_div2(2.0f, 0.0f, dh, de, qh, ql);
        y = f_minus(1.0f, 0.0f, qh, ql);
    }
    y = __builtin_copysignf(y, x);
    y = (x != x) ? x + x : y;
    return y;
}

// ---- pow's float form in float arithmetic: ln|x| as a float pair good to 2^-35 relative (so y ln|x|
// carries an absolute error below 2^-28 wherever the result is finite and nonzero), its product with y split
// exactly, and e^(ph + pl) with the full e^r factor on the input's low part; at most 0.75 ulp on 1.5*10^8
// random pairs against long-double pow, 0.79 where the result is subnormal.

__device__ __forceinline__ void f_add2(float ah, float al, float bh, float bl, float& h, float& l)
{ // (ah + al) + (bh + bl) as a pair
    float s, e;
    f_two2(ah, bh, s, e);
    f_fast2(s, e + (al + bl), h, l);
}

// ln x for finite x > 0: x = 2^k m, m in [sqrt2/2, sqrt2), s = f/(2+f) as a pair, ln m = 2s + (2/3)s^3 +
// (2/5)s^5 + s^7 (2/7 + ... + 2/13 s^6), the s^3 and s^5 terms as pairs; k ln2 in three parts
__device__ __forceinline__ void logf_pair(float x, float& lh, float& ll)
{
    const bool sub = x < 0x1p-126f;
    const float xs = sub ? x * 0x1p25f : x;
    const uint32_t u = __builtin_bit_cast(uint32_t, xs);
    int k = (int)(u >> 23) - 127 - (sub ? 25 : 0);
    uint32_t mb = (u & 0x7FFFFFu) | 0x3F800000u;
    const bool big = mb > 0x3FB504F3u;
    mb = big ? mb - 0x00800000u : mb;
    k += big ? 1 : 0;
    const float f = __builtin_bit_cast(float, mb) - 1.0f;
    float dh, dl, sh, sl, mh, ml, kh, ke;
    f_two2(2.0f, f, dh, dl);
    f_div2(f, 0.0f, dh, dl, sh, sl);
    const float zh = sh * sh;
    const float zl = __builtin_fmaf(sh, sh, -zh);
    const float c3h = sh * zh;
    const float c3l = __builtin_fmaf(sh, zh, -c3h) + __builtin_fmaf(sh, zl, 3.0f * zh * sl);
    const float t3h = 0x1.555556p-1f * c3h;
    const float t3l = __builtin_fmaf(0x1.555556p-1f, c3h, -t3h) +
                      __builtin_fmaf(0x1.555556p-1f, c3l, -0x1.555556p-26f * c3h);
    const float c5h = c3h * zh;
    const float c5l = __builtin_fmaf(c3h, zh, -c5h) + __builtin_fmaf(c3l, zh, c3h * zl);
    const float t5h = 0x1.99999ap-2f * c5h;
    const float t5l = __builtin_fmaf(0x1.99999ap-2f, c5h, -t5h) +
                      __builtin_fmaf(0x1.99999ap-2f, c5l, -0x1.99999ap-28f * c5h);
    float R = 0x1.3b13b2p-3f;
    R = __builtin_fmaf(R, zh, 0x1.745d18p-3f);
    R = __builtin_fmaf(R, zh, 0x1.c71c72p-3f);
    R = __builtin_fmaf(R, zh, 0x1.24924ap-2f);
    R = t5h + (t5l + R * (zh * c5h));
    f_add2(2.0f * sh, 2.0f * sl, t3h, t3l + R, mh, ml);
    const float fk = (float)k;
    const float k2h = fk * 0x1.7f7d1cp-20f;
    const float k2l = __builtin_fmaf(fk, 0x1.7f7d1cp-20f, -k2h);
    f_two2(fk * 0x1.62e400p-1f, k2h, kh, ke);
    f_add2(kh, ke + __builtin_fmaf(fk, 0x1.ef357ap-45f, k2l), mh, ml, lh, ll);
}

// e^(ph + pl), |ph| <= 104: expf's reduction of ph; the low part pl (up to 2^-17) enters as pl e^r
__device__ __forceinline__ float expf_of_pair(float ph, float pl)
{
    const float n = __builtin_rintf(ph * 0x1.715476p+0f);
    const float r1 = __builtin_fmaf(-n, 0x1.62e430p-1f, ph);
    const float rh = __builtin_fmaf(-n, -0x1.05c610p-29f, r1);
    const float rl = __builtin_fmaf(-n, -0x1.05c610p-29f, r1 - rh) + pl;
    float P = 0x1.a01a02p-13f;
    P = __builtin_fmaf(P, rh, 0x1.6c16c2p-10f);
    P = __builtin_fmaf(P, rh, 0x1.111112p-7f);
    P = __builtin_fmaf(P, rh, 0x1.555556p-5f);
    P = __builtin_fmaf(P, rh, 0x1.555556p-3f);
    P = __builtin_fmaf(P, rh, 0.5f);
    const float s = 1.0f + rh;
    const float t0 = __builtin_fmaf(rh * rh, P, (1.0f - s) + rh);
    return __builtin_ldexpf(s + __builtin_fmaf(rl, s + t0, t0), (int)n);
}

__device__ __forceinline__ int pow_int_kind_f(float y)
{
    if (!(__builtin_fabsf(y) <= 0x1.fffffep+127f)) return 0;
    if (__builtin_rintf(y) != y) return 0;
    if (__builtin_fabsf(y) >= 0x1p24f) return 1;
    const float h = y * 0.5f;
    return __builtin_rintf(h) != h ? 2 : 1;
}

// the special cases of pow (C99 Annex F.9.4.4)
__device__ __forceinline__ float powf(float x, float y)
{
    if (y == 0.0f) return 1.0f;
    if (x == 1.0f) return 1.0f;
    if (x != x || y != y) return x + y;
    const int ik = pow_int_kind_f(y);
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    if (ay > 0x1.fffffep+127f) {
        if (ax == 1.0f) return 1.0f;
        return ((ax < 1.0f) == (y < 0)) ? __builtin_inff() : 0.0f;
    }
    const bool neg = x < 0 || (x == 0.0f && __builtin_signbit(x));
    const float sgn = (neg && ik == 2) ? -1.0f : 1.0f;
    if (x == 0.0f) return y < 0 ? sgn * __builtin_inff() : sgn * 0.0f;
    if (ax > 0x1.fffffep+127f) return y < 0 ? sgn * 0.0f : sgn * __builtin_inff();
    if (x < 0 && ik == 0) return __builtin_nanf("");
    float lh, ll;
    logf_pair(ax, lh, ll);
    const float ph = y * lh;
    const float pl = __builtin_fmaf(y, lh, -ph) + y * ll;
    if (ph > 89.0f) return sgn * __builtin_inff();
    if (ph < -104.0f) return sgn * 0.0f;
    return sgn * expf_of_pair(ph, pl);
}

} // namespace dm
} // namespace aeth
