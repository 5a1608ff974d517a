/* float_sweep.c -- pins the oracle's float math (test infrastructure only): every float argument of each
 * one-argument float routine in oracle/amath.c against long-double libm (64-bit significand, so the
 * reference rounds to float correctly but for ties closer than 2^-40 ulp), and seeded random pairs for
 * pow. Prints, per routine, the largest error in float ulps (of the reference's binade, 2^-149 for
 * subnormal results), its argument and the fraction not correctly rounded.
 *   oracle/float_sweep [stride]     stride 1 (the default) visits every float; a larger stride samples
 * sin/cos/tan: |x| < 2^17 (the float reduction; above it they share the double reduction of the double
 * routines). pow: x over the whole range with y keeping |y ln x| <= 100, x within 1e-4 of 1 with |y| to
 * 1e6, x in [0, 4) with y in [-10, 10]. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static long double ulp_of(long double r)
{
    r = fabsl(r);
    if (r < 0x1p-126L) return 0x1p-149L;
    int e;
    frexpl(r, &e);
    return ldexpl(1.0L, e - 24);
}

typedef struct {
    double max_ulp;
    float worst_x, worst_y;
    long long not_cr, total;
} Tally;

static void score(Tally* t, float got, long double ref, float x, float y)
{
    double e;
    if (isnan(ref)) e = isnan(got) ? 0.0 : 1e9;
    else if (!(fabsl(ref) <= 0x1.fffffep127L)) e = (isinf(got) && !signbit(got) == !signbit(ref)) ? 0.0 : 1e9;
    else e = (double)(fabsl((long double)got - ref) / ulp_of(ref));
    if (got != (float)ref && !(isnan(got) && isnan(ref))) t->not_cr++;
    t->total++;
    if (e > t->max_ulp) {
        t->max_ulp = e;
        t->worst_x = x;
        t->worst_y = y;
    }
}

static void merge(Tally* a, const Tally* b)
{
    if (b->max_ulp > a->max_ulp) {
        a->max_ulp = b->max_ulp;
        a->worst_x = b->worst_x;
        a->worst_y = b->worst_y;
    }
    a->not_cr += b->not_cr;
    a->total += b->total;
}

static long double ref1(int fn, long double x)
{
    switch (fn) {
    case 0: return sinl(x);
    case 1: return cosl(x);
    case 2: return expl(x);
    case 3: return logl(x);
    case 4: return tanl(x);
    case 5: return atanl(x);
    case 6: return asinl(x);
    case 7: return acosl(x);
    default: return tanhl(x);
    }
}

static float got1(int fn, float x)
{
    switch (fn) {
    case 0: return orc_sinf(x);
    case 1: return orc_cosf(x);
    case 2: return orc_expf(x);
    case 3: return orc_logf(x);
    case 4: return orc_tanf(x);
    case 5: return orc_atanf(x);
    case 6: return orc_asinf(x);
    case 7: return orc_acosf(x);
    default: return orc_tanhf(x);
    }
}

static uint64_t mix(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int main(int argc, char** argv)
{
    const long long stride = argc > 1 ? atoll(argv[1]) : 1;
    static const char* names[9] = {"sin", "cos", "exp", "log", "tan", "atan", "asin", "acos", "tanh"};
    /* magnitudes visited (bit patterns of |x|): trig below 2^17, exp to 104, asin/acos to 1, others all */
    static const uint32_t top[9] = {0x48000000u, 0x48000000u, 0x42d00000u, 0x7f800001u, 0x48000000u,
                                    0x7f800001u, 0x3f800001u, 0x3f800001u, 0x7f800001u};
    for (int fn = 0; fn < 9; ++fn) {
        Tally all = {0};
#pragma omp parallel
        {
            Tally t = {0};
#pragma omp for schedule(static, 65536)
            for (long long b = 0; b < (long long)top[fn]; b += stride) {
                for (int sg = 0; sg < 2; ++sg) {
                    if (fn == 3 && sg) continue; /* log of negatives: NaN, covered by the tests */
                    const uint32_t u = (uint32_t)b | (sg ? 0x80000000u : 0u);
                    float x;
                    memcpy(&x, &u, 4);
                    score(&t, got1(fn, x), ref1(fn, (long double)x), x, 0.0f);
                }
            }
#pragma omp critical
            merge(&all, &t);
        }
        printf("%-5s max %.4f ulp at %a  not correctly rounded %lld of %lld (%.5f)\n", names[fn], all.max_ulp,
               all.worst_x, all.not_cr, all.total, (double)all.not_cr / (double)all.total);
        fflush(stdout);
    }
    const long long npow = 50000000 / stride + 1;
    for (int mode = 0; mode < 3; ++mode) {
        Tally all = {0};
#pragma omp parallel
        {
            Tally t = {0};
#pragma omp for schedule(static, 4096)
            for (long long i = 0; i < npow; ++i) {
                uint64_t st = (uint64_t)i * 7919u + (uint64_t)mode * 1000003u;
                const double u1 = (mix(&st) >> 11) * 0x1p-53, u2 = (mix(&st) >> 11) * 0x1p-53,
                             u3 = (mix(&st) >> 11) * 0x1p-53;
                float x, y;
                if (mode == 0) {
                    x = (float)exp2(u1 * 252 - 126);
                    y = (float)((u2 * 2 - 1) * 100.0 / fmax(fabs(log((double)x)), 1e-3));
                } else if (mode == 1) {
                    x = (float)(1.0 + (u1 * 2 - 1) * 1e-4);
                    y = (float)((u2 * 2 - 1) * 1e6 * u3);
                } else {
                    x = (float)(u1 * 4);
                    y = (float)(u2 * 20 - 10);
                }
                score(&t, orc_powf(x, y), powl((long double)x, (long double)y), x, y);
            }
#pragma omp critical
            merge(&all, &t);
        }
        printf("pow%d max %.4f ulp at x=%a y=%a  not correctly rounded %lld of %lld (%.5f)\n", mode, all.max_ulp,
               all.worst_x, all.worst_y, all.not_cr, all.total, (double)all.not_cr / (double)all.total);
        fflush(stdout);
    }
    return 0;
}
