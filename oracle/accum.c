/*
 * accum.c -- oracle restatement of aether's compensated accumulation.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * neumaierAdd_ / pairCombine_ / reduce32Host_: accum/atomic.h:142-188;
 * twoSumErrFinite (host spelling, Knuth): accum/TwoSum.h:87-109;
 * the float-float add: banded/Ff2.h:132-145 with the EFTs of
 * banded/detail/Fp32.h:198-230. The "device shape" functions restate the
 * fixed, size-only reduction tree of this project's kernels.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

static inline int finite_d(double x) { return isfinite(x); }
static inline int finite_f(float x) { return isfinite(x); }

/* TwoSum.h:100-109 host spelling with the +0 rule (positiveZero: -0 -> +0) */
static inline double two_sum_err_d(double a, double b, double s)
{
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return e == 0.0 ? 0.0 : e;
}
static inline float two_sum_err_f(float a, float b, float s)
{
    const float bb = s - a;
    const float e = (a - (s - bb)) + (b - bb);
    return e == 0.0f ? 0.0f : e;
}

void orc_neumaier_add_f64(double* hi, double* comp, double term)
{
    const double s = *hi + term;
    const double e = two_sum_err_d(*hi, term, s);
    *hi = s;
    *comp = finite_d(s) ? *comp + e : 0.0;
}

void orc_pair_combine_f64(double ah, double ac, double bh, double bc, double* oh, double* oc)
{
    const double s = ah + bh;
    const double e = two_sum_err_d(ah, bh, s);
    *oh = s;
    *oc = finite_d(s) ? (ac + bc) + e : 0.0;
}

void orc_neumaier_add_f32(float* hi, float* comp, float term)
{
    const float s = *hi + term;
    const float e = two_sum_err_f(*hi, term, s);
    *hi = s;
    *comp = finite_f(s) ? *comp + e : 0.0f;
}

void orc_pair_combine_f32(float ah, float ac, float bh, float bc, float* oh, float* oc)
{
    const float s = ah + bh;
    const float e = two_sum_err_f(ah, bh, s);
    *oh = s;
    *oc = finite_f(s) ? (ac + bc) + e : 0.0f;
}

/* atomic.h:171-188 */
static void reduce32_d(double h[32], double c[32], double* oh, double* oc)
{
    for (int off = 16; off >= 1; off >>= 1) {
        double nh[32], nc[32];
        for (int l = 0; l < 32; ++l) orc_pair_combine_f64(h[l], c[l], h[l ^ off], c[l ^ off], &nh[l], &nc[l]);
        memcpy(h, nh, sizeof nh);
        memcpy(c, nc, sizeof nc);
    }
    *oh = h[0];
    *oc = c[0];
}

static void reduce32_f(float h[32], float c[32], float* oh, float* oc)
{
    for (int off = 16; off >= 1; off >>= 1) {
        float nh[32], nc[32];
        for (int l = 0; l < 32; ++l) orc_pair_combine_f32(h[l], c[l], h[l ^ off], c[l ^ off], &nh[l], &nc[l]);
        memcpy(h, nh, sizeof nh);
        memcpy(c, nc, sizeof nc);
    }
    *oh = h[0];
    *oc = c[0];
}

/* tests/test_CompensatedSum_common.h:71-80 */
double orc_tree_sum32_f64(const double* x, uint64_t n)
{
    double h[32] = {0}, c[32] = {0}, oh, oc;
    for (uint64_t i = 0; i < n; ++i) orc_neumaier_add_f64(&h[i % 32], &c[i % 32], x[i]);
    reduce32_d(h, c, &oh, &oc);
    return oh + oc;
}

float orc_tree_sum32_f32(const float* x, uint64_t n)
{
    float h[32] = {0}, c[32] = {0}, oh, oc;
    for (uint64_t i = 0; i < n; ++i) orc_neumaier_add_f32(&h[i % 32], &c[i % 32], x[i]);
    reduce32_f(h, c, &oh, &oc);
    return oh + oc;
}

/* The host's parallel compensated sum over samples, as a caller of the reference writes it with its
 * per-block primitives: each OpenMP thread walks one contiguous block (static schedule) with
 * neumaierAdd_ (accum/atomic.h:142-149, the block commit of accum/run.h:345-410), and the block pairs
 * are combined with pairCombine_ (accum/atomic.h:157-164) in block order. The result depends on the
 * thread count (the reference's reduction order is a ULP contract, accum/run.h:34-37). Used to time
 * the CPU baseline; `nthreads` <= 0 takes OpenMP's default. */
void orc_block_sum_f64(const double* x, uint64_t n, int nthreads, double* oh, double* oc)
{
#ifdef _OPENMP
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    const int nt = 1;
    (void)nthreads;
#endif
    double* bh = calloc((size_t)nt, sizeof(double));
    double* bc = calloc((size_t)nt, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        const int t = omp_get_thread_num(), m = omp_get_num_threads();
#else
        const int t = 0, m = 1;
#endif
        const uint64_t lo = n * (uint64_t)t / (uint64_t)m, hi = n * (uint64_t)(t + 1) / (uint64_t)m;
        double h = 0.0, c = 0.0;
        for (uint64_t i = lo; i < hi; ++i) orc_neumaier_add_f64(&h, &c, x[i]);
        bh[t] = h;
        bc[t] = c;
    }
    double h = 0.0, c = 0.0;
    for (int t = 0; t < nt; ++t) orc_pair_combine_f64(h, c, bh[t], bc[t], &h, &c);
    free(bh);
    free(bc);
    *oh = h;
    *oc = c;
}

long double orc_sum_ld(const double* x, uint64_t n)
{
    long double s = 0.0L, c = 0.0L;
    for (uint64_t i = 0; i < n; ++i) {
        const long double t = s + x[i];
        if (fabsl(s) >= fabsl((long double)x[i])) c += (s - t) + x[i];
        else c += ((long double)x[i] - t) + s;
        s = t;
    }
    return s + c;
}

/* ---------------- float-float (Ff2) ---------------- */
typedef struct { float hi, lo; } ff_t;

/* Fp32.h:198-206 Knuth two-sum */
static inline void two_sum_f(float a, float b, float* s, float* e)
{
    *s = a + b;
    const float bp = *s - a;
    const float ap = *s - bp;
    const float eb = b - bp;
    const float ea = a - ap;
    *e = ea + eb;
}

/* Ff2.h:132-145 */
static inline ff_t ff_add(ff_t a, ff_t b)
{
    float sh, sl, th, tl, vh, vl;
    two_sum_f(a.hi, b.hi, &sh, &sl);
    two_sum_f(a.lo, b.lo, &th, &tl);
    const float c = sl + th;
    two_sum_f(sh, c, &vh, &vl);
    const float w = vl + tl;
    ff_t r;
    r.hi = vh + w; /* fast2Sum, Fp32.h:214-220 */
    const float bp = r.hi - vh;
    r.lo = w - bp;
    return r;
}

/* ---------------- the device reduction shape ---------------- */
enum { MODE_NEUMAIER = 0, MODE_PLAIN = 1, MODE_FF = 2 };
enum { TILE = 2048, LANES = 256, WAVE = 64, GROUP = 1024 };

typedef struct { double h, c; } pd_t;
typedef struct { float h, c; } pf_t;

static pd_t combine_d(pd_t a, pd_t b)
{
    pd_t r;
    orc_pair_combine_f64(a.h, a.c, b.h, b.c, &r.h, &r.c);
    return r;
}

static pd_t combine_dm(pd_t a, pd_t b, int mode)
{
    if (mode == MODE_PLAIN) {
        pd_t r = {a.h + b.h, 0.0};
        return r;
    }
    return combine_d(a, b);
}

static pf_t combine_f(pf_t a, pf_t b, int mode)
{
    pf_t r;
    if (mode == MODE_NEUMAIER) {
        orc_pair_combine_f32(a.h, a.c, b.h, b.c, &r.h, &r.c);
    } else if (mode == MODE_PLAIN) {
        r.h = a.h + b.h;
        r.c = 0.0f;
    } else {
        ff_t x = {a.h, a.c}, y = {b.h, b.c}, z = ff_add(x, y);
        r.h = z.hi;
        r.c = z.lo;
    }
    return r;
}

/* In-group halving tree (stride 512, 256, ..., 1), then recursively over group results. */
static pd_t tree_d(pd_t* p, uint64_t T, int mode)
{
    while (T > 1) {
        const uint64_t G = (T + GROUP - 1) / GROUP;
        for (uint64_t g = 0; g < G; ++g) {
            pd_t* q = p + g * GROUP;
            const uint64_t m = (T - g * GROUP) < GROUP ? (T - g * GROUP) : GROUP;
            for (uint64_t stride = GROUP / 2; stride >= 1; stride >>= 1)
                for (uint64_t t = 0; t < stride; ++t)
                    if (t + stride < m) q[t] = combine_dm(q[t], q[t + stride], mode);
            p[g] = q[0];
        }
        T = G;
    }
    return p[0];
}

static pf_t tree_f(pf_t* p, uint64_t T, int mode)
{
    while (T > 1) {
        const uint64_t G = (T + GROUP - 1) / GROUP;
        for (uint64_t g = 0; g < G; ++g) {
            pf_t* q = p + g * GROUP;
            const uint64_t m = (T - g * GROUP) < GROUP ? (T - g * GROUP) : GROUP;
            for (uint64_t stride = GROUP / 2; stride >= 1; stride >>= 1)
                for (uint64_t t = 0; t < stride; ++t)
                    if (t + stride < m) q[t] = combine_f(q[t], q[t + stride], mode);
            p[g] = q[0];
        }
        T = G;
    }
    return p[0];
}

/* lanes -> one pair: butterfly within each 64-lane wave (a balanced tree over
 * adjacent lanes), then ((w0, w1), (w2, w3)). */
static pd_t tile_combine_d(pd_t lane[LANES], int mode)
{
    pd_t w[4];
    for (int wv = 0; wv < 4; ++wv) {
        pd_t v[WAVE];
        memcpy(v, lane + wv * WAVE, sizeof v);
        for (int off = 1; off < WAVE; off <<= 1) {
            pd_t nv[WAVE];
            for (int i = 0; i < WAVE; ++i) nv[i] = combine_dm(v[i], v[i ^ off], mode);
            memcpy(v, nv, sizeof v);
        }
        w[wv] = v[0];
    }
    return combine_dm(combine_dm(w[0], w[1], mode), combine_dm(w[2], w[3], mode), mode);
}

static pf_t tile_combine_f(pf_t lane[LANES], int mode)
{
    pf_t w[4];
    for (int wv = 0; wv < 4; ++wv) {
        pf_t v[WAVE];
        memcpy(v, lane + wv * WAVE, sizeof v);
        for (int off = 1; off < WAVE; off <<= 1) {
            pf_t nv[WAVE];
            for (int i = 0; i < WAVE; ++i) nv[i] = combine_f(v[i], v[i ^ off], mode);
            memcpy(v, nv, sizeof v);
        }
        w[wv] = v[0];
    }
    return combine_f(combine_f(w[0], w[1], mode), combine_f(w[2], w[3], mode), mode);
}

static void device_sum_d(const double* x, uint64_t n, int square, int mode, double* hi, double* comp)
{
    const uint64_t T = n ? (n + TILE - 1) / TILE : 1;
    pd_t* part = (pd_t*)calloc(T, sizeof(pd_t));
#pragma omp parallel for schedule(static)
    for (uint64_t t = 0; t < T; ++t) {
        pd_t lane[LANES];
        for (int l = 0; l < LANES; ++l) {
            double h = 0.0, c = 0.0;
            for (int r = 0; r < TILE / 512; ++r)
                for (int k = 0; k < 2; ++k) {
                    const uint64_t i = t * TILE + (uint64_t)r * 512 + 2 * (uint64_t)l + k;
                    if (i < n) {
                        const double v = x[i];
                        if (mode == MODE_PLAIN) h = h + (square ? v * v : v);
                        else orc_neumaier_add_f64(&h, &c, square ? v * v : v);
                    }
                }
            lane[l].h = h;
            lane[l].c = c;
        }
        part[t] = tile_combine_d(lane, mode);
    }
    const pd_t r = tree_d(part, T, mode);
    *hi = r.h;
    *comp = r.c;
    free(part);
}

static void device_sum_f(const float* x, uint64_t n, int square, int mode, float* hi, float* comp)
{
    const uint64_t T = n ? (n + TILE - 1) / TILE : 1;
    pf_t* part = (pf_t*)calloc(T, sizeof(pf_t));
#pragma omp parallel for schedule(static)
    for (uint64_t t = 0; t < T; ++t) {
        pf_t lane[LANES];
        for (int l = 0; l < LANES; ++l) {
            float h = 0.0f, c = 0.0f;
            /* 16-byte loads: 4 floats per lane per round, 2 rounds of 1024 */
            for (int r = 0; r < TILE / 1024; ++r)
                for (int k = 0; k < 4; ++k) {
                    const uint64_t i = t * TILE + (uint64_t)r * 1024 + 4 * (uint64_t)l + k;
                    if (i >= n) continue;
                    const float v = x[i];
                    if (mode == MODE_NEUMAIER) {
                        orc_neumaier_add_f32(&h, &c, square ? v * v : v);
                    } else if (mode == MODE_PLAIN) {
                        h = h + (square ? v * v : v);
                    } else {
                        ff_t acc = {h, c}, term;
                        if (square) {
                            term.hi = v * v;
                            term.lo = fmaf(v, v, -term.hi); /* twoProd, Fp32.h:224-229 */
                        } else {
                            term.hi = v;
                            term.lo = 0.0f;
                        }
                        acc = ff_add(acc, term);
                        h = acc.hi;
                        c = acc.lo;
                    }
                }
            lane[l].h = h;
            lane[l].c = c;
        }
        part[t] = tile_combine_f(lane, mode);
    }
    const pf_t r = tree_f(part, T, mode);
    *hi = r.h;
    *comp = r.c;
    free(part);
}

void orc_device_sum_f64(const double* x, uint64_t n, double* hi, double* comp)
{
    device_sum_d(x, n, 0, MODE_NEUMAIER, hi, comp);
}
void orc_device_sumsq_f64(const double* x, uint64_t n, double* hi, double* comp)
{
    device_sum_d(x, n, 1, MODE_NEUMAIER, hi, comp);
}
void orc_device_sum_plain_f64(const double* x, uint64_t n, double* hi)
{
    double c;
    device_sum_d(x, n, 0, MODE_PLAIN, hi, &c);
}
void orc_device_sumsq_plain_f64(const double* x, uint64_t n, double* hi)
{
    double c;
    device_sum_d(x, n, 1, MODE_PLAIN, hi, &c);
}
void orc_device_sum_f32(const float* x, uint64_t n, float* hi, float* comp)
{
    device_sum_f(x, n, 0, MODE_NEUMAIER, hi, comp);
}
void orc_device_sumsq_f32(const float* x, uint64_t n, float* hi, float* comp)
{
    device_sum_f(x, n, 1, MODE_NEUMAIER, hi, comp);
}
void orc_device_sum_ff_f32(const float* x, uint64_t n, float* hi, float* lo) { device_sum_f(x, n, 0, MODE_FF, hi, lo); }
void orc_device_sumsq_ff_f32(const float* x, uint64_t n, float* hi, float* lo)
{
    device_sum_f(x, n, 1, MODE_FF, hi, lo);
}
/* plain FP32 in the same shape (exported for the precision comparison) */
void orc_device_sum_plain_f32(const float* x, uint64_t n, float* hi)
{
    float c;
    device_sum_f(x, n, 0, MODE_PLAIN, hi, &c);
}
void orc_device_sumsq_plain_f32(const float* x, uint64_t n, float* hi)
{
    float c;
    device_sum_f(x, n, 1, MODE_PLAIN, hi, &c);
}
