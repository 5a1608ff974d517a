/*
 * amath.c -- oracle copy of the own-math algorithm (sin/cos/exp/log).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference routes host transcendentals to libm and device FP64 sin/cos
 * to its BoundedTrig.h (math/math.h:228-287, math/detail/BoundedTrig.h:
 * 262-598: Cody-Waite below a bound, Payne-Hanek above). The device kernels
 * of this project implement the methods below; this file states the same
 * algorithm in C, operation for operation (explicit fma, no contraction), so
 * a device result can be compared bit for bit, and is itself pinned against
 * the reference's MPFR golden rows.
 *
 * Methods (all from the mathematics; constants computed with mpmath):
 *   sin/cos : Cody-Waite reduction by pi/2 in three parts (FMA, |x| < 2^20),
 *             Payne-Hanek with a 1280-bit table of 2/pi above; Taylor
 *             polynomials on [-pi/4, pi/4] (sin to r^17, cos to r^16) with a
 *             first-order correction for the reduced argument's tail.
 *   exp     : n = rint(x/ln2), two-part ln2 reduction with tail, Taylor to r^13.
 *   log     : x = 2^k m, m in [sqrt(2)/2, sqrt(2)), s = f/(2+f),
 *             log(1+f) = f - (hfsq - s(hfsq + R(s^2))), R from 2atanh(s).
 *   float   : the argument widened to double, the same reductions (two-part
 *             pi/2 below 2^20), Taylor sin to r^11, cos to r^10, exp to r^8,
 *             log's R to s^10, one rounding to float.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static inline uint64_t dbits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double bitsd(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }
static inline uint32_t fbits(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float bitsf(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* pi/2 = P1 + P2 + P3 (each the rounding of what is left) */
#define PIO2_1 0x1.921fb54442d18p+0
#define PIO2_2 0x1.1a62633145c07p-54
#define PIO2_3 -0x1.f1976b7ed8fbcp-110
#define TWO_OVER_PI 0x1.45f306dc9c883p-1

/* fraction bits of 2/pi, most significant first */
static const uint64_t kTwoOverPiBits[20] = {
    0xA2F9836E4E441529ULL, 0xFC2757D1F534DDC0ULL, 0xDB6295993C439041ULL, 0xFE5163ABDEBBC561ULL,
    0xB7246E3A424DD2E0ULL, 0x06492EEA09D1921CULL, 0xFE1DEB1CB129A73EULL, 0xE88235F52EBB4484ULL,
    0xE99C7026B45F7E41ULL, 0x3991D639835339F4ULL, 0x9C845F8BBDF9283BULL, 0x1FF897FFDE05980FULL,
    0xEF2F118B5A0A6D1FULL, 0x6D367ECF27CB09B7ULL, 0x4F463F669E5FEA2DULL, 0x7527BAC7EBE5F17BULL,
    0x3D0739F78A5292EAULL, 0x6BFB5FB11F8D5D08ULL, 0x56033046FC7B6BABULL, 0xF0CFBC209AF4361DULL,
};

/* 64 bits of the 2/pi fraction starting at 0-based bit position j (j may be negative: zeros) */
static uint64_t twopi_bits_at(int j)
{
    if (j <= -64) return 0;
    if (j < 0) return kTwoOverPiBits[0] >> (-j);
    const int wi = j >> 6, sh = j & 63;
    uint64_t v = kTwoOverPiBits[wi] << sh;
    if (sh) v |= kTwoOverPiBits[wi + 1] >> (64 - sh);
    return v;
}

/* Payne-Hanek: |x| = M 2^E (finite, |x| >= 2^20). Returns quadrant, reduced (rh, rl) with rh + rl ~ |x| - q pi/2. */
static int reduce_payne_hanek(double ax, double* rh, double* rl)
{
    const uint64_t u = dbits(ax);
    const int be = (int)(u >> 52);
    const uint64_t M = (u & 0xFFFFFFFFFFFFFULL) | 0x10000000000000ULL;
    const int E = be - 1075;
    const int j = E - 2; /* window starts at fraction bit E-1 (1-based) */
    const uint64_t w0 = twopi_bits_at(j), w1 = twopi_bits_at(j + 64), w2 = twopi_bits_at(j + 128);
    /* P = M * (w0 2^128 + w1 2^64 + w2) mod 2^192 */
    const unsigned __int128 p2 = (unsigned __int128)M * w2;
    const unsigned __int128 p1 = (unsigned __int128)M * w1;
    const uint64_t p0 = M * w0;
    uint64_t P0 = (uint64_t)p2;
    unsigned __int128 mid = (unsigned __int128)(uint64_t)(p2 >> 64) + (uint64_t)p1;
    uint64_t P1 = (uint64_t)mid;
    uint64_t P2 = (uint64_t)(p1 >> 64) + p0 + (uint64_t)(mid >> 64);
    int q = (int)(P2 >> 62);
    P2 &= 0x3FFFFFFFFFFFFFFFULL; /* fraction F = P mod 2^190 */
    int neg = 0;
    if (P2 >> 61) { /* f >= 1/2: f -= 1 */
        neg = 1;
        q += 1;
        /* F' = 2^190 - F */
        P0 = ~P0; P1 = ~P1; P2 = (~P2) & 0x3FFFFFFFFFFFFFFFULL;
        P0 += 1;
        if (P0 == 0) { P1 += 1; if (P1 == 0) P2 += 1; }
    }
    /* G = F' << 2 (192 bits) ; f = G 2^-192 */
    uint64_t G2 = (P2 << 2) | (P1 >> 62);
    uint64_t G1 = (P1 << 2) | (P0 >> 62);
    uint64_t G0 = P0 << 2;
    int shift = 0;
    if (G2 == 0) { G2 = G1; G1 = G0; G0 = 0; shift += 64; }
    const int lz = G2 ? __builtin_clzll(G2) : 0;
    if (lz) {
        G2 = (G2 << lz) | (G1 >> (64 - lz));
        G1 = (G1 << lz) | (G0 >> (64 - lz));
    }
    shift += lz;
    /* f = (G2 + G1 2^-64) 2^(-64-shift) */
    const double fh = (double)(G2 & ~0x7FFULL);
    const double fl = (double)(G2 & 0x7FFULL) + (double)G1 * 0x1p-64;
    const double sh = ldexp(fh, -64 - shift), sl = ldexp(fl, -64 - shift);
    /* r = f pi/2 in double-double */
    const double h = sh * PIO2_1;
    const double l = fma(sh, PIO2_1, -h) + fma(sh, PIO2_2, sl * PIO2_1);
    const double r = h + l;
    double t = l - (r - h);
    *rh = neg ? -r : r;
    *rl = neg ? -t : t;
    return q & 3;
}

/* Cody-Waite for |x| < 2^20: returns quadrant n mod 4 and (rh, rl). */
static int reduce_cody_waite(double x, double* rh, double* rl)
{
    const double n = rint(x * TWO_OVER_PI);
    const double r1 = fma(-n, PIO2_1, x); /* exact */
    const double ph = -n * PIO2_2;
    const double pl = fma(-n, PIO2_2, -ph); /* exact: -n*P2 = ph + pl */
    const double s = r1 + ph;
    const double bb = s - r1; /* Knuth two-sum: s + e = r1 + ph exactly */
    const double e = (r1 - (s - bb)) + (ph - bb);
    double t = e + pl;
    t = fma(-n, PIO2_3, t);
    const double r = s + t;
    *rl = t - (r - s);
    *rh = r;
    return ((int)(int64_t)n) & 3;
}

/* sin(r + rl) for |r| <= pi/4 */
static double ksin(double r, double rl)
{
    const double z = r * r;
    double ps = 0x1.952c77030ad4ap-49;     /* 1/17! */
    ps = fma(ps, z, -0x1.ae7f3e733b81fp-41); /* -1/15! */
    ps = fma(ps, z, 0x1.6124613a86d09p-33);  /* 1/13! */
    ps = fma(ps, z, -0x1.ae64567f544e4p-26); /* -1/11! */
    ps = fma(ps, z, 0x1.71de3a556c734p-19);  /* 1/9! */
    ps = fma(ps, z, -0x1.a01a01a01a01ap-13); /* -1/7! */
    ps = fma(ps, z, 0x1.1111111111111p-7);   /* 1/5! */
    ps = fma(ps, z, -0x1.5555555555555p-3);  /* -1/3! */
    const double tail = fma(-0.5 * z, rl, rl); /* rl (1 - r^2/2) ~ rl cos r */
    return r + fma(r * z, ps, tail);
}

/* cos(r + rl) for |r| <= pi/4 */
static double kcos(double r, double rl)
{
    const double z = r * r;
    double pc = 0x1.ae7f3e733b81fp-45;     /* 1/16! */
    pc = fma(pc, z, -0x1.93974a8c07c9dp-37); /* -1/14! */
    pc = fma(pc, z, 0x1.1eed8eff8d898p-29);  /* 1/12! */
    pc = fma(pc, z, -0x1.27e4fb7789f5cp-22); /* -1/10! */
    pc = fma(pc, z, 0x1.a01a01a01a01ap-16);  /* 1/8! */
    pc = fma(pc, z, -0x1.6c16c16c16c17p-10); /* -1/6! */
    pc = fma(pc, z, 0x1.5555555555555p-5);   /* 1/4! */
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    /* cos = 1 - z/2 + z^2 pc - r rl */
    const double corr = ((1.0 - w) - hz) + fma(z * z, pc, -(r * rl));
    return w + corr;
}

static int reduce(double x, double* rh, double* rl)
{
    const double ax = fabs(x);
    if (ax < 0x1p20) return reduce_cody_waite(x, rh, rl);
    int q = reduce_payne_hanek(ax, rh, rl);
    if (x < 0) { /* x = -|x|: quadrant -q, argument -r */
        *rh = -*rh;
        *rl = -*rl;
        q = (-q) & 3;
    }
    return q;
}

void orc_sincos(double x, double* s, double* c)
{
    const double ax = fabs(x);
    if (!(ax <= 0x1.fffffffffffffp+1023)) { /* inf or nan */
        *s = x - x;
        *c = x - x;
        return;
    }
    if (ax < 0x1p-27) {
        *s = x;
        *c = 1.0;
        return;
    }
    double rh, rl;
    const int q = reduce(x, &rh, &rl);
    const double sn = ksin(rh, rl), cs = kcos(rh, rl);
    switch (q) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}

double orc_sin(double x)
{
    double s, c;
    orc_sincos(x, &s, &c);
    return s;
}

double orc_cos(double x)
{
    double s, c;
    orc_sincos(x, &s, &c);
    return c;
}

#define LOG2E 0x1.71547652b82fep+0
#define LN2_HI 0x1.62e42fefa39efp-1
#define LN2_LO 0x1.abc9e3b39803fp-56
#define LN2_HI42 0x1.62e42fee00000p-1
#define LN2_LO42 0x1.a39ef35793c76p-33

double orc_exp(double x)
{
    if (x != x) return x + x;
    if (x > 710.0) return 0x1p1023 * 2.0;  /* overflow (ldexp below handles the edge) */
    if (x < -746.0) return 0.0;            /* underflow (ldexp below handles the edge) */
    const double n = rint(x * LOG2E);
    const double r1 = fma(-n, LN2_HI, x); /* exact */
    const double r = fma(-n, LN2_LO, r1);
    const double rl = fma(-n, LN2_LO, r1 - r);
    /* Horner for c2 + r c3 + ... + r^11 c13 */
    double q = 0x1.6124613a86d09p-33;    /* 1/13! */
    q = fma(q, r, 0x1.1eed8eff8d898p-29); /* 1/12! */
    q = fma(q, r, 0x1.ae64567f544e4p-26); /* 1/11! */
    q = fma(q, r, 0x1.27e4fb7789f5cp-22); /* 1/10! */
    q = fma(q, r, 0x1.71de3a556c734p-19); /* 1/9! */
    q = fma(q, r, 0x1.a01a01a01a01ap-16); /* 1/8! */
    q = fma(q, r, 0x1.a01a01a01a01ap-13); /* 1/7! */
    q = fma(q, r, 0x1.6c16c16c16c17p-10); /* 1/6! */
    q = fma(q, r, 0x1.1111111111111p-7);  /* 1/5! */
    q = fma(q, r, 0x1.5555555555555p-5);  /* 1/4! */
    q = fma(q, r, 0x1.5555555555555p-3);  /* 1/3! */
    q = fma(q, r, 0.5);                   /* 1/2! */
    const double u = fma(r * r, q, r);    /* exp(r) - 1 */
    const double p = 1.0 + (u + rl);
    return ldexp(p, (int)n);
}

double orc_log(double x)
{
    uint64_t u = dbits(x);
    int k = 0;
    if (x != x) return x + x;
    if (x < 0.0) return (x - x) / 0.0;        /* NaN */
    if (x == 0.0) return -1.0 / 0.0;         /* -inf */
    if (x == 1.0 / 0.0) return x;
    if ((u >> 52) == 0) { /* subnormal */
        x *= 0x1p54;
        u = dbits(x);
        k -= 54;
    }
    k += (int)(u >> 52) - 1023;
    uint64_t mb = (u & 0xFFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL; /* m in [1,2) */
    if (mb > 0x3FF6A09E667F3BCCULL) { /* m > sqrt(2): halve */
        mb -= 0x0010000000000000ULL;
        k += 1;
    }
    const double f = bitsd(mb) - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 0x1.642c8590b2164p-4;     /* 2/23 */
    R = fma(R, z, 0x1.8618618618618p-4);  /* 2/21 */
    R = fma(R, z, 0x1.af286bca1af28p-4);  /* 2/19 */
    R = fma(R, z, 0x1.e1e1e1e1e1e1ep-4);  /* 2/17 */
    R = fma(R, z, 0x1.1111111111111p-3);  /* 2/15 */
    R = fma(R, z, 0x1.3b13b13b13b14p-3);  /* 2/13 */
    R = fma(R, z, 0x1.745d1745d1746p-3);  /* 2/11 */
    R = fma(R, z, 0x1.c71c71c71c71cp-3);  /* 2/9 */
    R = fma(R, z, 0x1.2492492492492p-2);  /* 2/7 */
    R = fma(R, z, 0x1.999999999999ap-2);  /* 2/5 */
    R = fma(R, z, 0x1.5555555555555p-1);  /* 2/3 */
    R = R * z;
    const double hfsq = 0.5 * f * f;
    const double dk = (double)k;
    return dk * LN2_HI42 - ((hfsq - (s * (hfsq + R) + dk * LN2_LO42)) - f);
}

/* ---- float routines in float arithmetic (same operations as the device). The product of two floats
 * is split exactly with an FMA and float-float (unevaluated sum) steps carry the reduced argument and the
 * leading terms, so each result is one rounding of a value accurate to about 2^-29: measured over every
 * float argument, at most 0.67 ulp for sin/cos, 0.68 for exp (0.78 where the result is subnormal, which
 * rounds twice) and 0.643 for log, with more than 99.8% correctly rounded. Double is used only to reduce
 * trig arguments of magnitude 2^17 and above. */

static inline void f_fast2(float a, float b, float* s, float* e)
{   /* |a| >= |b| (or a == 0): s + e == a + b exactly */
    const float ss = a + b;
    *e = b - (ss - a);
    *s = ss;
}

static inline void f_two2(float a, float b, float* s, float* e)
{   /* any order: s + e == a + b exactly */
    const float ss = a + b;
    const float bb = ss - a;
    *e = (a - (ss - bb)) + (b - bb);
    *s = ss;
}

/* |x| < 2^17: n = rint(x 2/pi), r = x - n (C1 + C2 + C3) with C1 + C2 + C3 = pi/2 to 2^-73 relative;
 * x - n C1 is exact (FMA), n C2 is split exactly, so rh + rl is within 2^-55 of x - n pi/2. */
static int reduce_ff(float x, float* rh, float* rl)
{
    const float n = rintf(x * 0x1.45f306p-1f);
    const float r1 = fmaf(-n, 0x1.921fb6p+0f, x);
    const float p = n * -0x1.777a5cp-25f;
    const float pe = fmaf(n, -0x1.777a5cp-25f, -p);
    float h, e;
    f_two2(r1, -p, &h, &e);
    const float l = fmaf(-n, -0x1.ee59dap-50f, e - pe);
    f_fast2(h, l, rh, rl);
    return ((int)n) & 3;
}

/* sin(rh + rl) on |rh| <= pi/4: rh + (rh^3 P(rh^2) + rl (1 - rh^2/2)), Taylor to r^11; the leading
 * -rh^3/6 also takes the rounding errors of rh^2 and rh^3. Returns the rounded value, and in *lo what
 * it left out (the pair feeds tan's quotient). */
static float ksinf_(float rh, float rl, float* lo)
{
    const float z = rh * rh;
    float P = -0x1.ae6456p-26f;              /* -1/11! */
    P = fmaf(P, z, 0x1.71de3ap-19f);         /* 1/9! */
    P = fmaf(P, z, -0x1.a01a02p-13f);        /* -1/7! */
    P = fmaf(P, z, 0x1.111112p-7f);          /* 1/5! */
    P = fmaf(P, z, -0x1.555556p-3f);         /* -1/3! */
    const float zl = fmaf(rh, rh, -z);
    const float c = fmaf(-0.5f * z, rl, rl);
    const float r3 = rh * z;
    const float r3l = fmaf(rh, z, -r3);
    const float t = fmaf(fmaf(rh, zl, r3l), -0x1.555556p-3f, c);
    float h;
    f_fast2(rh, fmaf(r3, P, t), &h, lo);
    return h;
}

/* cos(rh + rl) on |rh| <= pi/4: (1 - z/2) split exactly, + z^2 Q(z) - rh rl, Taylor to r^12; a pair as
 * above */
static float kcosf_(float rh, float rl, float* lo)
{
    const float z = rh * rh;
    const float zl = fmaf(rh, rh, -z);
    float Q = 0x1.1eed8ep-29f;               /* 1/12! */
    Q = fmaf(Q, z, -0x1.27e4fcp-22f);        /* -1/10! */
    Q = fmaf(Q, z, 0x1.a01a02p-16f);         /* 1/8! */
    Q = fmaf(Q, z, -0x1.6c16c2p-10f);        /* -1/6! */
    Q = fmaf(Q, z, 0x1.555556p-5f);          /* 1/4! */
    const float hz = 0.5f * z;
    const float w = 1.0f - hz;
    const float we = (1.0f - w) - hz;
    float h;
    f_fast2(w, fmaf(z * z, Q, fmaf(-rh, rl, fmaf(-0.5f, zl, we))), &h, lo);
    return h;
}

/* the double reduction for float arguments of magnitude 2^17 and above: r = x - n P1 - n P2 below
 * 2^20, Payne-Hanek above */
static int reduce_f(double x, double* r)
{
    const double n = rint(x * TWO_OVER_PI);
    double rr = fma(-n, PIO2_1, x);
    rr = fma(-n, PIO2_2, rr);
    int q = isfinite(n) ? ((int)(long long)n) & 3 : 0;
    const double ax = fabs(x);
    if (ax >= 0x1p20 && ax <= 0x1.fffffffffffffp+1023) {
        double lh, ll;
        int ql = reduce_payne_hanek(ax, &lh, &ll);
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

/* the quadrant and reduced argument: float path below 2^17, the double reduction above */
static int reduce_f32(float x, float* rh, float* rl)
{
    if (fabsf(x) < 0x1p17f) return reduce_ff(x, rh, rl);
    double r;
    const int q = reduce_f((double)x, &r);
    *rh = (float)r;
    *rl = (float)(r - (double)*rh);
    return q;
}

float orc_sinf(float x)
{
    float rh, rl, lo;
    const int q = reduce_f32(x, &rh, &rl);
    float v = (q & 1) ? kcosf_(rh, rl, &lo) : ksinf_(rh, rl, &lo);
    v = (q & 2) ? -v : v;
    const float ax = fabsf(x);
    v = ax < 0x1p-12f ? x : v;
    v = !(ax <= 0x1.fffffep+127f) ? x - x : v;
    return v;
}

float orc_cosf(float x)
{
    float rh, rl, lo;
    const int q = reduce_f32(x, &rh, &rl);
    float v = (q & 1) ? ksinf_(rh, rl, &lo) : kcosf_(rh, rl, &lo);
    v = ((q + 1) & 2) ? -v : v;
    v = !(fabsf(x) <= 0x1.fffffep+127f) ? x - x : v;
    return v;
}

/* e^x on [-104, 89]: n = rint(x log2 e); r = x - n ln2 as rh + rl (x - n L1 exact by FMA);
 * e^r = (1 + rh) [split exactly] + rh^2 P(rh) + rl (1 + rh), Taylor to r^7; then the scaling by 2^n */
float orc_expf(float x)
{
    const float xc = fminf(fmaxf(x, -104.0f), 89.0f);
    const float n = rintf(xc * 0x1.715476p+0f);
    const float r1 = fmaf(-n, 0x1.62e430p-1f, xc);
    const float rh = fmaf(-n, -0x1.05c610p-29f, r1);
    const float rl = fmaf(-n, -0x1.05c610p-29f, r1 - rh);
    float P = 0x1.a01a02p-13f;               /* 1/7! */
    P = fmaf(P, rh, 0x1.6c16c2p-10f);        /* 1/6! */
    P = fmaf(P, rh, 0x1.111112p-7f);         /* 1/5! */
    P = fmaf(P, rh, 0x1.555556p-5f);         /* 1/4! */
    P = fmaf(P, rh, 0x1.555556p-3f);         /* 1/3! */
    P = fmaf(P, rh, 0.5f);
    const float s = 1.0f + rh;
    const float se = (1.0f - s) + rh;
    const float tail = fmaf(rh * rh, P, fmaf(rl, rh, se + rl));
    float y = ldexpf(s + tail, (int)n);
    y = (x != x) ? x + x : y;
    return y;
}

/* log x = k ln2 + log(1+f), f = m - 1 with m in [sqrt2/2, sqrt2); log(1+f) = f - f^2/2 + s (f^2/2 + R),
 * s = f/(2+f), R to s^10; f - f^2/2 and k K1 + (f - f^2/2) are carried as exact float-float sums
 * (K1 = ln2 to 16 bits, so k K1 is exact) */
float orc_logf(float x)
{
    const int sub = fabsf(x) < 0x1p-126f;
    const float xs = sub ? x * 0x1p25f : x;
    const uint32_t u = fbits(xs);
    int k = (int)(u >> 23) - 127 - (sub ? 25 : 0);
    uint32_t mb = (u & 0x7FFFFFu) | 0x3F800000u;
    const int big = mb > 0x3FB504F3u;
    mb = big ? mb - 0x00800000u : mb;
    k += big ? 1 : 0;
    const float f = bitsf(mb) - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float R = 0x1.745d18p-3f;                /* 2/11 */
    R = fmaf(R, z, 0x1.c71c72p-3f);          /* 2/9 */
    R = fmaf(R, z, 0x1.24924ap-2f);          /* 2/7 */
    R = fmaf(R, z, 0x1.99999ap-2f);          /* 2/5 */
    R = fmaf(R, z, 0x1.555556p-1f);          /* 2/3 */
    R = R * z;
    const float ffh = f * f;
    const float ffl = fmaf(f, f, -ffh);
    const float hf = 0.5f * ffh;
    const float d = f - hf;
    const float de = (f - d) - hf;
    const float fk = (float)k;
    float sh, se;
    f_two2(fk * 0x1.62e400p-1f, d, &sh, &se);
    float y = sh + fmaf(s, hf + R, fmaf(fk, 0x1.7f7d1cp-20f, fmaf(-0.5f, ffl, de) + se));
    y = (x == 0.0f) ? -INFINITY : y;
    y = (x < 0.0f) ? NAN : y;
    y = (x == INFINITY) ? x : y;
    y = (x != x) ? x + x : y;
    return y;
}

/* ---- the rest of the reference's elementwise set (expr/nodes/Elementwise.h:200-851): tan, atan,
 * asin, acos, tanh, pow, and the exact rsqrt / sign / fmin / fmax / ordered-max. Same style: every
 * operation as the device states it; double-double (unevaluated sum) corrections where a plain
 * evaluation would lose more than half an ulp. Constants: aether_amd/csrc/gen_math_constants.py. */

static inline void two_sum(double a, double b, double* s, double* e)
{
    const double ss = a + b;
    const double bb = ss - a;
    *e = (a - (ss - bb)) + (b - bb);
    *s = ss;
}

/* |a| >= |b| (or a == 0) */
static inline void fast_two_sum(double a, double b, double* s, double* e)
{
    const double ss = a + b;
    *e = b - (ss - a);
    *s = ss;
}

/* atan(k/16) and pi/2 - atan(k/16), k = 0..16, as (hi, lo) */
static const double kAtanK[17][2] = {
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
static const double kAtanCoK[17][2] = {
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
#define PIO2_H 0x1.921fb54442d18p+0
#define PIO2_L 0x1.1a62633145c07p-54
#define PI_H 0x1.921fb54442d18p+1
#define PI_L 0x1.1a62633145c07p-53

/* tan: sin and cos of the reduced argument as unevaluated sums (the kernels above before their last
 * addition), then one double-double quotient; odd quadrants give -cos/sin. */
double orc_tan(double x)
{
    const double ax = fabs(x);
    if (!(ax <= 0x1.fffffffffffffp+1023)) return x - x;
    if (ax < 0x1p-27) return x;
    double rh, rl;
    const int q = reduce(x, &rh, &rl);
    const double z = rh * rh;
    double ps = 0x1.952c77030ad4ap-49;
    ps = fma(ps, z, -0x1.ae7f3e733b81fp-41);
    ps = fma(ps, z, 0x1.6124613a86d09p-33);
    ps = fma(ps, z, -0x1.ae64567f544e4p-26);
    ps = fma(ps, z, 0x1.71de3a556c734p-19);
    ps = fma(ps, z, -0x1.a01a01a01a01ap-13);
    ps = fma(ps, z, 0x1.1111111111111p-7);
    ps = fma(ps, z, -0x1.5555555555555p-3);
    const double stail = fma(-0.5 * z, rl, rl);
    double sh, sl;
    fast_two_sum(rh, fma(rh * z, ps, stail), &sh, &sl);
    double pc = 0x1.ae7f3e733b81fp-45;
    pc = fma(pc, z, -0x1.93974a8c07c9dp-37);
    pc = fma(pc, z, 0x1.1eed8eff8d898p-29);
    pc = fma(pc, z, -0x1.27e4fb7789f5cp-22);
    pc = fma(pc, z, 0x1.a01a01a01a01ap-16);
    pc = fma(pc, z, -0x1.6c16c16c16c17p-10);
    pc = fma(pc, z, 0x1.5555555555555p-5);
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    double ch, cl;
    fast_two_sum(w, ((1.0 - w) - hz) + fma(z * z, pc, -(rh * rl)), &ch, &cl);
    /* even q: (sh + sl) / (ch + cl); odd q: -(ch + cl) / (sh + sl) */
    const int odd = q & 1;
    const double nh = odd ? -ch : sh, nl = odd ? -cl : sl;
    const double dh = odd ? sh : ch, dl = odd ? sl : cl;
    const double th = nh / dh;
    const double tl = (fma(-th, dh, nh) + nl - th * dl) / dh;
    return th + tl;
}

/* atan(zh + zl) for 0 <= zh <= 1 (zl a tail), as hi + lo: atan(c) from the table plus atan(t),
 * t = (z - c) / (1 + z c), c = k/16 the nearest sixteenth, |t| <= 1/32; atan(t) to t^13. */
static void atan_core(double zh, double zl, int co, double* hi, double* lo)
{
    double kd = rint(zh * 16.0);
    kd = (kd >= 0.0 && kd <= 16.0) ? kd : 0.0;
    const int k = (int)kd;
    const double c = kd * 0.0625;
    const double num = zh - c; /* exact */
    double nh, nl;
    two_sum(num, zl, &nh, &nl);
    const double ph = zh * c;
    const double pe = fma(zh, c, -ph);
    double dh, dl;
    fast_two_sum(1.0, ph, &dh, &dl);
    dl += pe;
    const double th = nh / dh;
    const double tl = (fma(-th, dh, nh) + nl - th * dl) / dh;
    const double w = th * th;
    double P = 1.0 / 13.0;
    P = fma(P, w, -1.0 / 11.0);
    P = fma(P, w, 1.0 / 9.0);
    P = fma(P, w, -1.0 / 7.0);
    P = fma(P, w, 1.0 / 5.0);
    P = fma(P, w, -1.0 / 3.0);
    const double tail = fma(th * w, P, tl); /* atan(t) - th */
    const double bh = co ? kAtanCoK[k][0] : kAtanK[k][0];
    const double bl = co ? kAtanCoK[k][1] : kAtanK[k][1];
    double s, e;
    two_sum(bh, co ? -th : th, &s, &e);
    fast_two_sum(s, e + (bl + (co ? -tail : tail)), hi, lo);
}

double orc_atan(double x)
{
    if (x != x) return x + x;
    const double a = fabs(x);
    const int big = a > 1.0;
    double zh = a, zl = 0.0;
    if (big) {
        zh = 1.0 / a;
        zl = (a <= 0x1.fffffffffffffp+1023) ? fma(-zh, a, 1.0) / a : 0.0;
    }
    double h, l;
    atan_core(zh, zl, big, &h, &l);
    const double r = h + l;
    return copysign(r, x);
}

/* asin(a) for 0 <= a <= 1/2 as hi + lo: atan(a / sqrt(1 - a^2)), every step an unevaluated sum */
static void asin_small(double a, double* hi, double* lo)
{
    const double p = a * a;
    const double pe = fma(a, a, -p);
    double wh, wl;
    fast_two_sum(1.0, -p, &wh, &wl);
    wl -= pe;
    const double sh = sqrt(wh);
    const double sl = (fma(-sh, sh, wh) + wl) / (2.0 * sh);
    const double zh = a / sh;
    const double zl = (fma(-zh, sh, a) - zh * sl) / sh;
    atan_core(zh, zl, 0, hi, lo);
}

/* asin(s) for s = sqrt(v), 0 <= v <= 1/4 (v exact), as hi + lo */
static void asin_of_root(double v, double* hi, double* lo)
{
    const double sh = sqrt(v);
    const double sl = sh > 0.0 ? fma(-sh, sh, v) / (2.0 * sh) : 0.0;
    double wh, wl;
    fast_two_sum(1.0, -v, &wh, &wl); /* 1 - s^2 = 1 - v */
    const double qh = sqrt(wh);
    const double ql = (fma(-qh, qh, wh) + wl) / (2.0 * qh);
    const double zh = sh / qh;
    const double zl = (fma(-zh, qh, sh) + sl - zh * ql) / qh;
    atan_core(zh, zl, 0, hi, lo);
}

double orc_asin(double x)
{
    const double a = fabs(x);
    if (!(a <= 1.0)) return (x - x) / (x - x); /* NaN: |x| > 1 or NaN */
    double h, l, r;
    if (a <= 0.5) {
        asin_small(a, &h, &l);
        r = h + l;
    } else {
        /* asin(a) = pi/2 - 2 asin(sqrt((1 - a) / 2)) */
        asin_of_root((1.0 - a) * 0.5, &h, &l);
        double s, e;
        two_sum(PIO2_H, -2.0 * h, &s, &e);
        r = s + (e + (PIO2_L - 2.0 * l));
    }
    return copysign(r, x);
}

double orc_acos(double x)
{
    const double a = fabs(x);
    if (!(a <= 1.0)) return (x - x) / (x - x);
    double h, l, s, e;
    if (a <= 0.5) {
        /* pi/2 - asin(x) */
        asin_small(a, &h, &l);
        if (x < 0) {
            h = -h;
            l = -l;
        }
        two_sum(PIO2_H, -h, &s, &e);
        return s + (e + (PIO2_L - l));
    }
    asin_of_root((1.0 - a) * 0.5, &h, &l);
    if (x > 0) return 2.0 * h + 2.0 * l; /* 2 asin(sqrt((1 - x) / 2)) */
    two_sum(PI_H, -2.0 * h, &s, &e);     /* pi - 2 asin(sqrt((1 + x) / 2)) */
    return s + (e + (PI_L - 2.0 * l));
}

/* tanh: |x| < 0.625 its Taylor series to x^43; above 1 - 2 / (exp(2|x|) + 1) */
double orc_tanh(double x)
{
    const double a = fabs(x);
    double r;
    if (a < 0.625) {
        const double w = a * a;
        double P = -0x1.42ba1a349b490p-28;
        P = fma(P, w, 0x1.8e25ff9327e2cp-27);
        P = fma(P, w, -0x1.eb3229047434cp-26);
        P = fma(P, w, 0x1.2efe8db3aff1fp-24);
        P = fma(P, w, -0x1.75cde6563fed9p-23);
        P = fma(P, w, 0x1.cd299de4ae6bbp-22);
        P = fma(P, w, -0x1.1c77df95c1c0dp-20);
        P = fma(P, w, 0x1.5ef2da474e5b7p-19);
        P = fma(P, w, -0x1.b0f72d3ee24e9p-18);
        P = fma(P, w, 0x1.0b132d39a6050p-16);
        P = fma(P, w, -0x1.497d8eea25259p-15);
        P = fma(P, w, 0x1.967e18afcafadp-14);
        P = fma(P, w, -0x1.f57d7734d1664p-13);
        P = fma(P, w, 0x1.3558248036744p-11);
        P = fma(P, w, -0x1.7da36452b75e3p-10);
        P = fma(P, w, 0x1.d6d3d0e157de0p-9);
        P = fma(P, w, -0x1.226e355e6c23dp-7);
        P = fma(P, w, 0x1.664f4882c10fap-6);
        P = fma(P, w, -0x1.ba1ba1ba1ba1cp-5);
        P = fma(P, w, 0x1.1111111111111p-3);
        P = fma(P, w, -0x1.5555555555555p-2);
        r = fma(a * w, P, a);
    } else {
        const double e = orc_exp(2.0 * a);
        r = 1.0 - 2.0 / (e + 1.0);
    }
    return copysign(r, x);
}

/* log(a) for finite a > 0 as hi + lo to about 2^-66 relative: a = 2^k m, m in [sqrt(2)/2, sqrt(2)),
 * s = f / (2 + f) as an unevaluated sum, log m = 2s + 2s^3/3 + 2s^5 Q(s^2) with the cubic term as an
 * unevaluated product, k ln 2 in two parts. */
static void log_dd(double a, double* hi, double* lo)
{
    uint64_t u = dbits(a);
    int k = 0;
    if ((u >> 52) == 0) {
        a *= 0x1p54;
        u = dbits(a);
        k -= 54;
    }
    k += (int)(u >> 52) - 1023;
    uint64_t mb = (u & 0xFFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL;
    if (mb > 0x3FF6A09E667F3BCCULL) {
        mb -= 0x0010000000000000ULL;
        k += 1;
    }
    const double f = bitsd(mb) - 1.0; /* exact */
    double dh, dl;
    fast_two_sum(2.0, f, &dh, &dl);
    const double sh = f / dh;
    const double sl = (fma(-sh, dh, f) - sh * dl) / dh;
    const double w = sh * sh;
    const double wl = fma(sh, sh, -w);
    const double c3 = w * sh;
    const double c3l = fma(w, sh, -c3) + wl * sh;
    const double m3 = c3 * 0x1.5555555555555p-1; /* s^3 2/3 */
    const double m3l = fma(c3, 0x1.5555555555555p-1, -m3) + (c3l * 0x1.5555555555555p-1 + c3 * 0x1.5555555555555p-55);
    double Q = 2.0 / 25.0;
    Q = fma(Q, w, 2.0 / 23.0);
    Q = fma(Q, w, 2.0 / 21.0);
    Q = fma(Q, w, 2.0 / 19.0);
    Q = fma(Q, w, 2.0 / 17.0);
    Q = fma(Q, w, 2.0 / 15.0);
    Q = fma(Q, w, 2.0 / 13.0);
    Q = fma(Q, w, 2.0 / 11.0);
    Q = fma(Q, w, 2.0 / 9.0);
    Q = fma(Q, w, 2.0 / 7.0);
    Q = fma(Q, w, 2.0 / 5.0);
    const double tail = c3 * w * Q; /* 2 s^5 / 5 + ... */
    double g, ge;
    two_sum(2.0 * sh, m3, &g, &ge);
    /* 2 sl (1 + s^2): the tail of s through the linear and cubic terms */
    const double low = ge + (fma(2.0 * w, sl, 2.0 * sl) + (m3l + tail));
    const double dk = (double)k;
    double s1, e1;
    two_sum(dk * LN2_HI42, g, &s1, &e1);
    fast_two_sum(s1, e1 + (low + dk * LN2_LO42), hi, lo);
}

/* exp(xh + xl), |xl| tiny next to xh, xh in [-746, 710]: the reduction above with the tail added */
static double exp_dd(double xh, double xl)
{
    const double n = rint(xh * LOG2E);
    const double r1 = fma(-n, LN2_HI, xh);
    const double r = fma(-n, LN2_LO, r1);
    const double rl = fma(-n, LN2_LO, r1 - r) + xl;
    double q = 0x1.6124613a86d09p-33;
    q = fma(q, r, 0x1.1eed8eff8d898p-29);
    q = fma(q, r, 0x1.ae64567f544e4p-26);
    q = fma(q, r, 0x1.27e4fb7789f5cp-22);
    q = fma(q, r, 0x1.71de3a556c734p-19);
    q = fma(q, r, 0x1.a01a01a01a01ap-16);
    q = fma(q, r, 0x1.a01a01a01a01ap-13);
    q = fma(q, r, 0x1.6c16c16c16c17p-10);
    q = fma(q, r, 0x1.1111111111111p-7);
    q = fma(q, r, 0x1.5555555555555p-5);
    q = fma(q, r, 0x1.5555555555555p-3);
    q = fma(q, r, 0.5);
    const double uu = fma(r * r, q, r);
    /* exp(r + rl) = (1 + uu)(1 + rl) to first order in rl: the tail carries the argument's low part
     * (up to 2^-53 |xh|), so its product with uu is kept */
    const double p = 1.0 + (uu + fma(rl, uu, rl));
    return ldexp(p, (int)n);
}

/* y is an integer (finite); odd: also odd */
static int pow_int_kind(double y)
{
    if (!(fabs(y) <= 0x1.fffffffffffffp+1023)) return 0; /* inf/nan: not an integer */
    if (rint(y) != y) return 0;
    if (fabs(y) >= 0x1p53) return 1; /* even */
    const double h = y * 0.5;
    return rint(h) != h ? 2 : 1;      /* 2: odd */
}

/* pow with C99 Annex F special cases; |x|^y = exp(y log|x|) with both steps as unevaluated sums */
double orc_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (x == 1.0) return 1.0;
    if (x != x || y != y) return x + y;
    const int ik = pow_int_kind(y);
    const double ax = fabs(x), ay = fabs(y);
    if (ay > 0x1.fffffffffffffp+1023) { /* y = +-inf */
        if (ax == 1.0) return 1.0;
        return ((ax < 1.0) == (y < 0)) ? 1.0 / 0.0 : 0.0;
    }
    const int neg = x < 0 || (x == 0.0 && signbit(x));
    const double sgn = (neg && ik == 2) ? -1.0 : 1.0;
    if (x == 0.0) return y < 0 ? sgn / 0.0 : sgn * 0.0;
    if (ax > 0x1.fffffffffffffp+1023) return y < 0 ? sgn * 0.0 : sgn / 0.0;
    if (x < 0 && ik == 0) return (x - x) / (x - x);
    double lh, ll;
    log_dd(ax, &lh, &ll);
    const double ph = y * lh;
    const double pl = fma(y, lh, -ph) + y * ll;
    if (ph > 710.0) return sgn / 0.0;
    if (ph < -746.0) return sgn * 0.0;
    return sgn * exp_dd(ph, pl);
}

double orc_rsqrt(double x) { return 1.0 / sqrt(x); }
double orc_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }
/* fmin / fmax as glibc's generic form: the other operand when one is NaN; on equal values (+-0) the
 * second */
double orc_fmin(double a, double b) { return (a < b || b != b) ? a : b; }
double orc_fmax(double a, double b) { return (a > b || b != b) ? a : b; }
/* the reductions' ordered max (expr/Reduce.h MaxOp): a > b ? a : b */
double orc_selgt(double a, double b) { return a > b ? a : b; }

void orc_math_f64(int fn, int mode, uint64_t n, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        const double v = x[i];
        double r;
        if (mode == 0) {
            r = fn == 0 ? sin(v) : fn == 1 ? cos(v) : fn == 2 ? exp(v) : log(v);
        } else {
            switch (fn) {
            case 0: r = orc_sin(v); break;
            case 1: r = orc_cos(v); break;
            case 2: r = orc_exp(v); break;
            case 3: r = orc_log(v); break;
            case 4: r = orc_tan(v); break;
            case 5: r = orc_atan(v); break;
            case 6: r = orc_asin(v); break;
            case 7: r = orc_acos(v); break;
            case 8: r = orc_tanh(v); break;
            case 9: r = orc_rsqrt(v); break;
            default: r = orc_sign(v); break;
            }
        }
        y[i] = r;
    }
}

/* binary: 0 pow, 1 fmin, 2 fmax, 3 ordered max */
void orc_math2_f64(int fn, uint64_t n, const double* x, const double* y, double* z)
{
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        const double a = x[i], b = y[i];
        z[i] = fn == 0 ? orc_pow(a, b) : fn == 1 ? orc_fmin(a, b) : fn == 2 ? orc_fmax(a, b) : orc_selgt(a, b);
    }
}

/* ---- pow's float form in float arithmetic: ln|x| as a float pair good to 2^-35 relative (y ln|x| then
 * carries an absolute error below 2^-28 wherever the result is finite and nonzero), its product with y
 * split exactly, and e^(ph + pl) with the full e^r factor on the input's low part. Measured on 1.5*10^8
 * random pairs (x over the whole float range with y keeping the result in range, x within 10^-4 of 1
 * with |y| to 10^6, and x in [0, 4) with y in [-10, 10]) against long-double powl: at most 0.75 ulp,
 * 0.79 where the result is subnormal (oracle/float_sweep.c). */

static void f_div2(float nh, float nl, float dh, float dl, float* h, float* l);

static void f_add2(float ah, float al, float bh, float bl, float* h, float* l)
{   /* (ah + al) + (bh + bl) as a pair */
    float s, e;
    f_two2(ah, bh, &s, &e);
    f_fast2(s, e + (al + bl), h, l);
}

/* ln x for finite x > 0: x = 2^k m, m in [sqrt2/2, sqrt2), s = f/(2+f) as a pair, ln m = 2s + (2/3)s^3 +
 * (2/5)s^5 + s^7 (2/7 + ... + 2/13 s^6), the s^3 and s^5 terms as pairs; k ln2 in three parts
 * (ln2 to 16 bits, then two floats) */
static void logf_pair(float x, float* lh, float* ll)
{
    const int sub = x < 0x1p-126f;
    const float xs = sub ? x * 0x1p25f : x;
    const uint32_t u = fbits(xs);
    int k = (int)(u >> 23) - 127 - (sub ? 25 : 0);
    uint32_t mb = (u & 0x7FFFFFu) | 0x3F800000u;
    const int big = mb > 0x3FB504F3u;
    mb = big ? mb - 0x00800000u : mb;
    k += big ? 1 : 0;
    const float f = bitsf(mb) - 1.0f;
    float dh, dl, sh, sl, mh, ml, kh, ke;
    f_two2(2.0f, f, &dh, &dl);
    f_div2(f, 0.0f, dh, dl, &sh, &sl);
    const float zh = sh * sh;
    const float zl = fmaf(sh, sh, -zh);
    const float c3h = sh * zh;
    const float c3l = fmaf(sh, zh, -c3h) + fmaf(sh, zl, 3.0f * zh * sl);
    const float t3h = 0x1.555556p-1f * c3h;                 /* 2/3 = 0x1.555556p-1 - 0x1.555556p-26 */
    const float t3l = fmaf(0x1.555556p-1f, c3h, -t3h) + fmaf(0x1.555556p-1f, c3l, -0x1.555556p-26f * c3h);
    const float c5h = c3h * zh;
    const float c5l = fmaf(c3h, zh, -c5h) + fmaf(c3l, zh, c3h * zl);
    const float t5h = 0x1.99999ap-2f * c5h;                 /* 2/5 = 0x1.99999ap-2 - 0x1.99999ap-28 */
    const float t5l = fmaf(0x1.99999ap-2f, c5h, -t5h) + fmaf(0x1.99999ap-2f, c5l, -0x1.99999ap-28f * c5h);
    float R = 0x1.3b13b2p-3f;                               /* 2/13 */
    R = fmaf(R, zh, 0x1.745d18p-3f);                        /* 2/11 */
    R = fmaf(R, zh, 0x1.c71c72p-3f);                        /* 2/9 */
    R = fmaf(R, zh, 0x1.24924ap-2f);                        /* 2/7 */
    R = t5h + (t5l + R * (zh * c5h));
    f_add2(2.0f * sh, 2.0f * sl, t3h, t3l + R, &mh, &ml);
    const float fk = (float)k;
    const float k2h = fk * 0x1.7f7d1cp-20f;
    const float k2l = fmaf(fk, 0x1.7f7d1cp-20f, -k2h);
    f_two2(fk * 0x1.62e400p-1f, k2h, &kh, &ke);
    f_add2(kh, ke + fmaf(fk, 0x1.ef357ap-45f, k2l), mh, ml, lh, ll);
}

/* e^(ph + pl), |ph| <= 104: orc_expf's reduction of ph; the low part pl (up to 2^-17) enters as
 * pl e^r, e^r = 1 + r + r^2 P(r) */
static float expf_of_pair(float ph, float pl)
{
    const float n = rintf(ph * 0x1.715476p+0f);
    const float r1 = fmaf(-n, 0x1.62e430p-1f, ph);
    const float rh = fmaf(-n, -0x1.05c610p-29f, r1);
    const float rl = fmaf(-n, -0x1.05c610p-29f, r1 - rh) + pl;
    float P = 0x1.a01a02p-13f;
    P = fmaf(P, rh, 0x1.6c16c2p-10f);
    P = fmaf(P, rh, 0x1.111112p-7f);
    P = fmaf(P, rh, 0x1.555556p-5f);
    P = fmaf(P, rh, 0x1.555556p-3f);
    P = fmaf(P, rh, 0.5f);
    const float s = 1.0f + rh;
    const float t0 = fmaf(rh * rh, P, (1.0f - s) + rh);
    return ldexpf(s + fmaf(rl, s + t0, t0), (int)n);
}

static int pow_int_kind_f(float y)
{
    if (!(fabsf(y) <= 0x1.fffffep+127f)) return 0;
    if (rintf(y) != y) return 0;
    if (fabsf(y) >= 0x1p24f) return 1;
    const float h = y * 0.5f;
    return rintf(h) != h ? 2 : 1;
}

/* the special cases of orc_pow (C99 Annex F.9.4.4) */
float orc_powf(float x, float y)
{
    if (y == 0.0f) return 1.0f;
    if (x == 1.0f) return 1.0f;
    if (x != x || y != y) return x + y;
    const int ik = pow_int_kind_f(y);
    const float ax = fabsf(x), ay = fabsf(y);
    if (ay > 0x1.fffffep+127f) {
        if (ax == 1.0f) return 1.0f;
        return ((ax < 1.0f) == (y < 0)) ? INFINITY : 0.0f;
    }
    const int neg = x < 0 || (x == 0.0f && signbit(x));
    const float sgn = (neg && ik == 2) ? -1.0f : 1.0f;
    if (x == 0.0f) return y < 0 ? sgn * INFINITY : sgn * 0.0f;
    if (ax > 0x1.fffffep+127f) return y < 0 ? sgn * 0.0f : sgn * INFINITY;
    if (x < 0 && ik == 0) return NAN;
    float lh, ll;
    logf_pair(ax, &lh, &ll);
    const float ph = y * lh;
    const float pl = fmaf(y, lh, -ph) + y * ll;
    if (ph > 89.0f) return sgn * INFINITY;
    if (ph < -104.0f) return sgn * 0.0f;
    return sgn * expf_of_pair(ph, pl);
}

/* the float forms (pow in float arithmetic; the others select or compare, exact in any width) */
void orc_math2_f32(int fn, uint64_t n, const float* x, const float* y, float* z)
{
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        const double a = x[i], b = y[i];
        z[i] = fn == 0 ? orc_powf(x[i], y[i])
                       : (float)(fn == 1 ? orc_fmin(a, b) : fn == 2 ? orc_fmax(a, b) : orc_selgt(a, b));
    }
}

/* ---- float forms of tan, atan, asin, acos, tanh in float arithmetic, on the float routines' pieces:
 * pair quotients with an exact FMA residual, a table of atan(k/16) as float pairs. Measured over every
 * float argument (|x| < 2^17 for tan): at most 0.783 ulp for tan, 0.501 for atan/asin/acos, 0.640
 * for tanh (oracle/float_sweep.c). */

/* (nh + nl) / (dh + dl) as a pair: q = nh/dh, its residual nh - q dh exact by FMA */
static void f_div2(float nh, float nl, float dh, float dl, float* h, float* l)
{
    const float q = nh / dh;
    const float r = fmaf(-q, dl, fmaf(-q, dh, nh) + nl);
    f_fast2(q, r / dh, h, l);
}

float orc_tanf(float x)
{
    float rh, rl, sl, cl, h, l;
    const int q = reduce_f32(x, &rh, &rl);
    const float sh = ksinf_(rh, rl, &sl), ch = kcosf_(rh, rl, &cl);
    if (q & 1) f_div2(-ch, -cl, sh, sl, &h, &l);
    else f_div2(sh, sl, ch, cl, &h, &l);
    float v = h + l;
    const float ax = fabsf(x);
    v = ax < 0x1p-12f ? x : v;
    v = !(ax <= 0x1.fffffep+127f) ? x - x : v;
    return v;
}

/* atan(k/16), k = 0..16, as float pairs (mpmath) */
static const float ATAN16_H[17] = {0x0p+0f, 0x1.ff55bcp-5f, 0x1.fd5baap-4f, 0x1.7b97b4p-3f, 0x1.f5b760p-3f,
    0x1.362774p-2f, 0x1.6f6194p-2f, 0x1.a64eecp-2f, 0x1.dac670p-2f, 0x1.0657eap-1f, 0x1.1e00bap-1f,
    0x1.345f02p-1f, 0x1.4978fap-1f, 0x1.5d5898p-1f, 0x1.700a7cp-1f, 0x1.819d0cp-1f, 0x1.921fb6p-1f};
static const float ATAN16_L[17] = {0x0p+0f, -0x1.1a6042p-30f, -0x1.54f424p-30f, 0x1.79cb60p-28f,
    -0x1.b4dfc8p-29f, -0x1.1f0286p-27f, 0x1.e4def0p-30f, 0x1.e611fep-29f, 0x1.586ed4p-28f, -0x1.6499e6p-26f,
    0x1.7bdfd6p-26f, -0x1.98e422p-28f, 0x1.934f70p-28f, 0x1.c5a6c6p-27f, 0x1.5e118cp-27f, -0x1.1d4eb6p-26f,
    -0x1.777a5cp-26f};

/* atan(th + tl), 0 <= t <= 1, as a pair: c = k/16 nearest t, u = (t - c)/(1 + t c) (|u| <= 1/32, t - c
 * exact), atan(k/16) + u - u^3/3 + u^5/5 - u^7/7 */
static void atan01_f(float th, float tl, float* yh, float* yl)
{
    const float kf = rintf(th * 16.0f);
    const int k = (kf >= 0.0f && kf <= 16.0f) ? (int)kf : 0; /* NaN (and |x| > 1 in asin) read entry 0 */
    const float c = (float)k * 0.0625f;
    float nh, nl, uh, ul, s, e;
    f_fast2(th - c, tl, &nh, &nl);
    const float dh = fmaf(th, c, 1.0f);
    const float dl = fmaf(tl, c, fmaf(th, c, 1.0f - dh));
    f_div2(nh, nl, dh, dl, &uh, &ul);
    const float u2 = uh * uh;
    float p = fmaf(u2, -0x1.24924ap-3f, 0x1.99999ap-3f);   /* -1/7, 1/5 */
    p = fmaf(p, u2, -0x1.555556p-2f);                      /* -1/3 */
    const float tail = fmaf(uh * u2, p, ul) + ATAN16_L[k];
    f_fast2(ATAN16_H[k], uh, &s, &e);
    f_fast2(s, e + tail, yh, yl);
}

/* c_hi + c_lo - (yh + yl), rounded once (c = pi/2 or pi as a float pair) */
static float f_minus(float chi, float clo, float yh, float yl)
{
    float s, e;
    f_two2(chi, -yh, &s, &e);
    return s + ((e + clo) - yl);
}

float orc_atanf(float x)
{
    const float ax = fabsf(x);
    const int inv = ax > 1.0f;
    const float th = inv ? 1.0f / ax : ax;
    const float tl = (inv && ax <= 0x1.fffffep+127f) ? fmaf(-th, ax, 1.0f) * th : 0.0f;
    float yh, yl;
    atan01_f(th, tl, &yh, &yl);
    float y = inv ? f_minus(0x1.921fb6p+0f, -0x1.777a5cp-25f, yh, yl) : yh + yl;
    y = copysignf(y, x);
    y = (x != x) ? x + x : y;
    return y;
}

/* atan of the smaller of ax / sqrt(1 - x^2) and its inverse (both pairs; 1 - x^2 exact as a pair);
 * *far = 1 for the inverse (ax > sqrt2/2) */
static void asin_core_f(float ax, float* ah, float* al, int* far)
{
    const float x2h = ax * ax;
    const float x2l = fmaf(ax, ax, -x2h);
    float mh, me, ml, th, tl;
    f_two2(1.0f, -x2h, &mh, &me);
    f_fast2(mh, me - x2l, &mh, &ml);
    const float sh = sqrtf(mh);
    const float sl = mh > 0.0f ? (fmaf(-sh, sh, mh) + ml) / (2.0f * sh) : 0.0f;
    *far = ax > 0x1.6a09e6p-1f;
    if (*far) f_div2(sh, sl, ax, 0.0f, &th, &tl);
    else f_div2(ax, 0.0f, sh, sl, &th, &tl);
    atan01_f(th, tl, ah, al);
}

float orc_asinf(float x)
{
    const float ax = fabsf(x);
    float ah, al;
    int far;
    asin_core_f(ax, &ah, &al, &far);
    float y = far ? f_minus(0x1.921fb6p+0f, -0x1.777a5cp-25f, ah, al) : ah + al;
    y = copysignf(y, x);
    y = (ax > 1.0f) ? NAN : y;
    y = (x != x) ? x + x : y;
    return y;
}

float orc_acosf(float x)
{
    const float ax = fabsf(x);
    float ah, al;
    int far;
    asin_core_f(ax, &ah, &al, &far);
    float y;
    if (far) y = x > 0.0f ? ah + al : f_minus(0x1.921fb6p+1f, -0x1.777a5cp-24f, ah, al);
    else y = x < 0.0f ? f_minus(0x1.921fb6p+0f, -0x1.777a5cp-25f, -ah, -al)
                      : f_minus(0x1.921fb6p+0f, -0x1.777a5cp-25f, ah, al);
    y = (ax > 1.0f) ? NAN : y;
    y = (x != x) ? x + x : y;
    return y;
}

/* e^x as a pair for |x| <= 20: orc_expf's steps without the clamp, the pair scaled by 2^n exactly */
static void expf_pair(float x, float* h, float* l)
{
    const float n = rintf(x * 0x1.715476p+0f);
    const float r1 = fmaf(-n, 0x1.62e430p-1f, x);
    const float rh = fmaf(-n, -0x1.05c610p-29f, r1);
    const float rl = fmaf(-n, -0x1.05c610p-29f, r1 - rh);
    float P = 0x1.a01a02p-13f;
    P = fmaf(P, rh, 0x1.6c16c2p-10f);
    P = fmaf(P, rh, 0x1.111112p-7f);
    P = fmaf(P, rh, 0x1.555556p-5f);
    P = fmaf(P, rh, 0x1.555556p-3f);
    P = fmaf(P, rh, 0.5f);
    const float s = 1.0f + rh;
    const float se = (1.0f - s) + rh;
    float eh, el;
    f_fast2(s, fmaf(rh * rh, P, fmaf(rl, rh, se + rl)), &eh, &el);
    *h = ldexpf(eh, (int)n);
    *l = ldexpf(el, (int)n);
}

/* tanh: below 0.55 the Taylor series to x^21 (x^3/3 with the roundings of x^2 and x^3), above
 * 1 - 2/(e^{2|x|} + 1) in pairs; |x| is capped at 10 (tanh rounds to 1 from 9.02) */
float orc_tanhf(float x)
{
    const float ax = fabsf(x);
    float y;
    if (ax < 0.55f) {
        const float z = ax * ax;
        const float zl = fmaf(ax, ax, -z);
        float P = 0x1.967e18p-14f;           /* x^21 */
        P = fmaf(P, z, -0x1.f57d78p-13f);    /* x^19 */
        P = fmaf(P, z, 0x1.355824p-11f);     /* x^17 */
        P = fmaf(P, z, -0x1.7da364p-10f);    /* x^15 */
        P = fmaf(P, z, 0x1.d6d3d0p-9f);      /* x^13 */
        P = fmaf(P, z, -0x1.226e36p-7f);     /* x^11 */
        P = fmaf(P, z, 0x1.664f48p-6f);      /* x^9 */
        P = fmaf(P, z, -0x1.ba1ba2p-5f);     /* x^7 */
        P = fmaf(P, z, 0x1.111112p-3f);      /* x^5: 2/15 */
        P = P * z;
        const float x3 = ax * z;
        const float x3l = fmaf(ax, z, -x3);
        const float t = fmaf(fmaf(ax, zl, x3l), -0x1.555556p-2f, x3 * P);
        y = ax + fmaf(x3, -0x1.555556p-2f, t);
    } else {
        float eh, el, dh, de, qh, ql;
        expf_pair(2.0f * fminf(ax, 10.0f), &eh, &el);
        f_two2(eh, 1.0f, &dh, &de);
        f_fast2(dh, de + el, &dh, &de);
        f_div2(2.0f, 0.0f, dh, de, &qh, &ql);
        y = f_minus(1.0f, 0.0f, qh, ql);
    }
    y = copysignf(y, x);
    y = (x != x) ? x + x : y;
    return y;
}

void orc_math_f32(int fn, int mode, uint64_t n, const float* x, float* y)
{
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        const float v = x[i];
        float r;
        if (mode == 0) {
            r = fn == 0 ? sinf(v) : fn == 1 ? cosf(v) : fn == 2 ? expf(v) : logf(v);
        } else {
            switch (fn) {
            case 0: r = orc_sinf(v); break;
            case 1: r = orc_cosf(v); break;
            case 2: r = orc_expf(v); break;
            case 3: r = orc_logf(v); break;
            case 4: r = orc_tanf(v); break;
            case 5: r = orc_atanf(v); break;
            case 6: r = orc_asinf(v); break;
            case 7: r = orc_acosf(v); break;
            case 8: r = orc_tanhf(v); break;
            case 9: r = 1.0f / sqrtf(v); break;
            default: r = (float)orc_sign(v); break;
            }
        }
        y[i] = r;
    }
}
