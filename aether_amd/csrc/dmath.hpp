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
//   float   : the argument widened to double, the same reductions, shorter
//             polynomials (Taylor sin to r^11, cos to r^10, exp to r^8, log's
//             R to s^10), one final rounding.
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

// |x| >= 2^20, finite. Returns the quadrant of |x|; (rh, rl) ~ |x| - q pi/2.
__device__ __noinline__ int reduce_large(double ax, double* rh, double* rl)
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
    *rh = neg ? -r : r;
    *rl = neg ? -t : t;
    return q & 3;
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
        double lh, ll;
        int ql = reduce_large(ax, &lh, &ll);
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

// ---- float routines. The argument is widened to double (exact), reduced
// once, and evaluated with polynomials just long enough for float (error
// below 2^-30 relative before the one rounding to float): a third of the
// FP64 operations of the double routines.

// |x| < 2^20: r = x - n P1 - n P2 (two FMAs; P1 + P2 is pi/2 to 2^-107, so
// r is within 2^-80 absolute); above, the Payne-Hanek path, out of line.
__device__ __forceinline__ int reduce_pio2_f(double x, double* r)
{
    const double n = __builtin_rint(x * kTwoOverPi);
    double rr = __builtin_fma(-n, kPio2_1, x);
    rr = __builtin_fma(-n, kPio2_2, rr);
    int q = ((int)(long long)n) & 3;
    const double ax = __builtin_fabs(x);
    if (ax >= 0x1p20 && ax <= 0x1.fffffffffffffp+1023) {
        double lh, ll;
        int ql = reduce_large(ax, &lh, &ll);
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

// sin r on |r| <= pi/4, Taylor to r^11 (truncation < 2^-37 relative)
__device__ __forceinline__ double ksinf(double r)
{
    const double z = r * r;
    double p = -0x1.ae64567f544e4p-26;
    p = __builtin_fma(p, z, 0x1.71de3a556c734p-19);
    p = __builtin_fma(p, z, -0x1.a01a01a01a01ap-13);
    p = __builtin_fma(p, z, 0x1.1111111111111p-7);
    p = __builtin_fma(p, z, -0x1.5555555555555p-3);
    return __builtin_fma(r * z, p, r);
}

// cos r on |r| <= pi/4, Taylor to r^10 (truncation < 2^-33)
__device__ __forceinline__ double kcosf(double r)
{
    const double z = r * r;
    double p = -0x1.27e4fb7789f5cp-22;
    p = __builtin_fma(p, z, 0x1.a01a01a01a01ap-16);
    p = __builtin_fma(p, z, -0x1.6c16c16c16c17p-10);
    p = __builtin_fma(p, z, 0x1.5555555555555p-5);
    p = __builtin_fma(p, z, -0.5);
    return __builtin_fma(z, p, 1.0);
}

__device__ __forceinline__ float sinf(float x)
{
    double r;
    const int q = reduce_pio2_f((double)x, &r);
    const double sn = ksinf(r), cs = kcosf(r);
    double v = (q & 1) ? cs : sn;
    v = (q & 2) ? -v : v;
    float y = (float)v;
    const float ax = __builtin_fabsf(x);
    y = ax < 0x1p-12f ? x : y; // sin x rounds to x; keeps -0
    y = !(ax <= 0x1.fffffep+127f) ? x - x : y;
    return y;
}

__device__ __forceinline__ float cosf(float x)
{
    double r;
    const int q = reduce_pio2_f((double)x, &r);
    const double sn = ksinf(r), cs = kcosf(r);
    double v = (q & 1) ? sn : cs;
    v = ((q + 1) & 2) ? -v : v;
    float y = (float)v;
    y = !(__builtin_fabsf(x) <= 0x1.fffffep+127f) ? x - x : y;
    return y;
}

// e^x: clamp to [-104, 89] (beyond: 0 and +inf after rounding), n = rint(x log2 e),
// r = x - n ln2 in two parts (|r| <= 0.347), e^r = 1 + r + r^2 p(r), Taylor to r^8
__device__ __forceinline__ float expf(float x)
{
    const double xc = __builtin_fmin(__builtin_fmax((double)x, -104.0), 89.0);
    const double n = __builtin_rint(xc * kLog2e);
    double r = __builtin_fma(-n, kLn2Hi, xc);
    r = __builtin_fma(-n, kLn2Lo, r);
    double p = 0x1.a01a01a01a01ap-16;
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-13);
    p = __builtin_fma(p, r, 0x1.6c16c16c16c17p-10);
    p = __builtin_fma(p, r, 0x1.1111111111111p-7);
    p = __builtin_fma(p, r, 0x1.5555555555555p-5);
    p = __builtin_fma(p, r, 0x1.5555555555555p-3);
    p = __builtin_fma(p, r, 0.5);
    const double u = __builtin_fma(r * r, p, r);
    float y = (float)__builtin_ldexp(1.0 + u, (int)n);
    y = (x != x) ? x + x : y;
    return y;
}

// log x: the double arrangement above with R to s^10 (truncation < 2^-34)
__device__ __forceinline__ float logf(float x)
{
    const uint64_t u = bits_of((double)x);
    int k = (int)(u >> 52) - 1023;
    uint64_t mb = (u & 0xFFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL;
    const bool big = mb > 0x3FF6A09E667F3BCCULL;
    mb = big ? mb - 0x0010000000000000ULL : mb;
    k += big ? 1 : 0;
    const double f = from_bits(mb) - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 0x1.745d1745d1746p-3;
    R = __builtin_fma(R, z, 0x1.c71c71c71c71cp-3);
    R = __builtin_fma(R, z, 0x1.2492492492492p-2);
    R = __builtin_fma(R, z, 0x1.999999999999ap-2);
    R = __builtin_fma(R, z, 0x1.5555555555555p-1);
    R = R * z;
    const double hfsq = 0.5 * f * f;
    const double dk = (double)k;
    float y = (float)(dk * kLn2Hi42 - ((hfsq - (s * (hfsq + R) + dk * kLn2Lo42)) - f));
    y = (x == 0.0f) ? -__builtin_inff() : y;
    y = (x < 0.0f) ? __builtin_nanf("") : y;
    y = (x == __builtin_inff()) ? x : y;
    y = (x != x) ? x + x : y;
    return y;
}

} // namespace dm
} // namespace aeth
