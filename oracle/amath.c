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

/* ---- float routines: argument widened to double, reduced once, polynomials
 * long enough for float, one rounding (same operations as the device). */

/* |x| < 2^20: r = x - n P1 - n P2; above: Payne-Hanek */
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

/* Taylor sin to r^11 on |r| <= pi/4 */
static double ksinf_(double r)
{
    const double z = r * r;
    double p = -0x1.ae64567f544e4p-26;      /* -1/11! */
    p = fma(p, z, 0x1.71de3a556c734p-19);   /* 1/9! */
    p = fma(p, z, -0x1.a01a01a01a01ap-13);  /* -1/7! */
    p = fma(p, z, 0x1.1111111111111p-7);    /* 1/5! */
    p = fma(p, z, -0x1.5555555555555p-3);   /* -1/3! */
    return fma(r * z, p, r);
}

/* Taylor cos to r^10 on |r| <= pi/4 */
static double kcosf_(double r)
{
    const double z = r * r;
    double p = -0x1.27e4fb7789f5cp-22;      /* -1/10! */
    p = fma(p, z, 0x1.a01a01a01a01ap-16);   /* 1/8! */
    p = fma(p, z, -0x1.6c16c16c16c17p-10);  /* -1/6! */
    p = fma(p, z, 0x1.5555555555555p-5);    /* 1/4! */
    p = fma(p, z, -0.5);
    return fma(z, p, 1.0);
}

float orc_sinf(float x)
{
    double r;
    const int q = reduce_f((double)x, &r);
    const double sn = ksinf_(r), cs = kcosf_(r);
    double v = (q & 1) ? cs : sn;
    v = (q & 2) ? -v : v;
    float y = (float)v;
    const float ax = fabsf(x);
    y = ax < 0x1p-12f ? x : y;
    y = !(ax <= 0x1.fffffep+127f) ? x - x : y;
    return y;
}

float orc_cosf(float x)
{
    double r;
    const int q = reduce_f((double)x, &r);
    const double sn = ksinf_(r), cs = kcosf_(r);
    double v = (q & 1) ? sn : cs;
    v = ((q + 1) & 2) ? -v : v;
    float y = (float)v;
    y = !(fabsf(x) <= 0x1.fffffep+127f) ? x - x : y;
    return y;
}

/* e^x on [-104, 89]: n = rint(x log2 e), two-part ln2, Taylor to r^8 */
float orc_expf(float x)
{
    const double xc = fmin(fmax((double)x, -104.0), 89.0);
    const double n = rint(xc * LOG2E);
    double r = fma(-n, LN2_HI, xc);
    r = fma(-n, LN2_LO, r);
    double p = 0x1.a01a01a01a01ap-16;       /* 1/8! */
    p = fma(p, r, 0x1.a01a01a01a01ap-13);   /* 1/7! */
    p = fma(p, r, 0x1.6c16c16c16c17p-10);   /* 1/6! */
    p = fma(p, r, 0x1.1111111111111p-7);    /* 1/5! */
    p = fma(p, r, 0x1.5555555555555p-5);    /* 1/4! */
    p = fma(p, r, 0x1.5555555555555p-3);    /* 1/3! */
    p = fma(p, r, 0.5);
    const double u = fma(r * r, p, r);
    float y = (float)ldexp(1.0 + u, (int)n);
    y = (x != x) ? x + x : y;
    return y;
}

/* log x: orc_log's arrangement with R to s^10 */
float orc_logf(float x)
{
    const uint64_t u = dbits((double)x);
    int k = (int)(u >> 52) - 1023;
    uint64_t mb = (u & 0xFFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL;
    const int big = mb > 0x3FF6A09E667F3BCCULL;
    mb = big ? mb - 0x0010000000000000ULL : mb;
    k += big ? 1 : 0;
    const double f = bitsd(mb) - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 0x1.745d1745d1746p-3;        /* 2/11 */
    R = fma(R, z, 0x1.c71c71c71c71cp-3);    /* 2/9 */
    R = fma(R, z, 0x1.2492492492492p-2);    /* 2/7 */
    R = fma(R, z, 0x1.999999999999ap-2);    /* 2/5 */
    R = fma(R, z, 0x1.5555555555555p-1);    /* 2/3 */
    R = R * z;
    const double hfsq = 0.5 * f * f;
    const double dk = (double)k;
    float y = (float)(dk * LN2_HI42 - ((hfsq - (s * (hfsq + R) + dk * LN2_LO42)) - f));
    y = (x == 0.0f) ? -INFINITY : y;
    y = (x < 0.0f) ? NAN : y;
    y = (x == INFINITY) ? x : y;
    y = (x != x) ? x + x : y;
    return y;
}

void orc_math_f64(int fn, int mode, uint64_t n, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        const double v = x[i];
        double r;
        if (mode == 0) {
            r = fn == 0 ? sin(v) : fn == 1 ? cos(v) : fn == 2 ? exp(v) : log(v);
        } else {
            r = fn == 0 ? orc_sin(v) : fn == 1 ? orc_cos(v) : fn == 2 ? orc_exp(v) : orc_log(v);
        }
        y[i] = r;
    }
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
            r = fn == 0 ? orc_sinf(v) : fn == 1 ? orc_cosf(v) : fn == 2 ? orc_expf(v) : orc_logf(v);
        }
        y[i] = r;
    }
}
