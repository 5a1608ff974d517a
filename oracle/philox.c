/*
 * philox.c -- oracle restatement of aether's counter-based RNG.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows random/detail/Philox.h:76-133 (constants, round, 10-round
 * bijection), random/detail/Backend.h:75-236 (splitmix64, word -> real
 * conversions, philoxBlock counter layout, randomBits64, uniform01,
 * standardNormal) and random/Generator.h:64-81,193 (keying chain).
 * The Philox bijection itself is Salmon et al., SC'11.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

/* Philox.h:78-85 */
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u

void orc_philox4x32_10(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4])
{
    uint32_t x = ctr[0], y = ctr[1], z = ctr[2], w = ctr[3];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { /* key bump between rounds, Philox.h:127-131 */
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        const uint64_t p0 = (uint64_t)PHILOX_M0 * x; /* Philox.h:108-115 */
        const uint64_t p1 = (uint64_t)PHILOX_M1 * z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t nx = hi1 ^ y ^ k0;
        const uint32_t nz = hi0 ^ w ^ k1;
        x = nx;
        y = lo1;
        z = nz;
        w = lo0;
    }
    out[0] = x;
    out[1] = y;
    out[2] = z;
    out[3] = w;
}

/* Backend.h:144-150: counter words {lo(counter), hi(counter), lo(global), hi(global)}, key = seed. */
void orc_philox_block(uint64_t seed, uint32_t global, uint64_t counter, uint32_t out[4])
{
    const uint32_t ctr[4] = {(uint32_t)counter, (uint32_t)(counter >> 32), global, 0u};
    orc_philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32), out);
}

uint64_t orc_random_bits64(uint64_t seed, uint32_t global, uint64_t counter)
{
    uint32_t b[4];
    orc_philox_block(seed, global, counter, b);
    return ((uint64_t)b[1] << 32) | (uint64_t)b[0]; /* Backend.h:170-176 */
}

/* Backend.h:123-126: w*2^-32 + 2^-33, exact */
static double u01_double_from_word(uint32_t w)
{
    return (double)w * 0x1p-32 + 0x1p-33;
}

/* Backend.h:131-134: (2*(w>>9)+1) * 2^-24, exact */
static float u01_float_from_word(uint32_t w)
{
    return (float)(((w >> 9) << 1) | 1u) * 0x1p-24f;
}

/* the word conversions themselves, for checks at chosen words (tests/test_RandomIndependence.cpp:98-119) */
double orc_u01_f64_from_word(uint32_t w) { return u01_double_from_word(w); }
float orc_u01_f32_from_word(uint32_t w) { return u01_float_from_word(w); }

double orc_uniform01_f64(uint64_t seed, uint32_t global, uint64_t counter)
{
    uint32_t b[4];
    orc_philox_block(seed, global, counter, b);
    return u01_double_from_word(b[0]);
}

float orc_uniform01_f32(uint64_t seed, uint32_t global, uint64_t counter)
{
    uint32_t b[4];
    orc_philox_block(seed, global, counter, b);
    return u01_float_from_word(b[0]);
}

/* Backend.h:155-159 */
static double u01_double_hq(uint32_t x, uint32_t y)
{
    const uint64_t z = (uint64_t)x ^ ((uint64_t)y << (53 - 32));
    return (double)z * 0x1p-53 + 0x1p-54;
}

/* Backend.h:209-236, double branch. The mathmode selects libm (the
 * reference's host route, math/math.h:228-287 -> std::) or the own-math
 * algorithm shared with the device kernels. */
double orc_normal_f64(uint64_t seed, uint32_t global, uint64_t counter, int mathmode)
{
    uint32_t b[4];
    orc_philox_block(seed, global, counter, b);
    const double u = u01_double_hq(b[0], b[1]);
    const uint64_t zy = (uint64_t)b[2] ^ ((uint64_t)b[3] << (53 - 32));
    const double v = (double)zy * (0x1p-53 * 2.0) + 0x1p-53;
    const double ang = v * 3.1415926535897932;
    double s, sv, cv;
    if (mathmode == 0) {
        s = sqrt(-2.0 * log(u));
        sv = sin(ang);
        cv = cos(ang);
    } else {
        s = sqrt(-2.0 * orc_log(u));
        orc_sincos(ang, &sv, &cv);
    }
    (void)cv;
    return s * sv;
}

/* Backend.h:209-236, float branch (two words). */
float orc_normal_f32(uint64_t seed, uint32_t global, uint64_t counter, int mathmode)
{
    uint32_t b[4];
    orc_philox_block(seed, global, counter, b);
    const float k2pi = 2.3283064e-10f * 6.2831855f; /* folded in float, as the reference's constant */
    const float u = u01_float_from_word(b[0]);
    const float v = (float)b[1] * k2pi + (k2pi / 2.0f);
    float s, sv;
    if (mathmode == 0) {
        s = sqrtf(-2.0f * logf(u));
        sv = sinf(v);
    } else {
        s = sqrtf(-2.0f * orc_logf(u));
        sv = orc_sinf(v);
    }
    return s * sv;
}

/* Backend.h:75-81 */
uint64_t orc_splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* Generator.h:64-67 */
uint64_t orc_generator_key(uint64_t seed, uint64_t stream_id)
{
    return orc_splitmix64(seed ^ orc_splitmix64(stream_id));
}

/* Generator.h:77-81, salt at :193 */
uint64_t orc_substream_key(uint64_t key, uint64_t k)
{
    return orc_splitmix64(key ^ orc_splitmix64(k ^ 0x2545F4914F6CDD1DULL));
}

/* ---- bulk streams ---- */

void orc_fill_bits32(uint64_t seed, uint64_t first, uint64_t n, uint32_t* out)
{
    /* word w of the stream is word (w & 1) of bits64(seed, w >> 1, 0) */
#pragma omp parallel for schedule(static)
    for (uint64_t k = 0; k < n; ++k) {
        const uint64_t w = first + k;
        uint32_t b[4];
        orc_philox_block(seed, (uint32_t)(w >> 1), 0, b);
        out[k] = b[w & 1];
    }
}

void orc_fill_uniform_f64(uint64_t seed, uint64_t first, uint64_t n, double* out)
{
#pragma omp parallel for schedule(static)
    for (uint64_t k = 0; k < n; ++k) out[k] = orc_uniform01_f64(seed, (uint32_t)(first + k), 0);
}

void orc_fill_normal_f64(uint64_t seed, uint64_t first, uint64_t n, int mathmode, double* out)
{
#pragma omp parallel for schedule(static)
    for (uint64_t k = 0; k < n; ++k) out[k] = orc_normal_f64(seed, (uint32_t)(first + k), 0, mathmode);
}

void orc_fill_uniform_f32(uint64_t seed, uint64_t first, uint64_t n, float* out)
{
#pragma omp parallel for schedule(static)
    for (uint64_t k = 0; k < n; ++k) out[k] = orc_uniform01_f32(seed, (uint32_t)(first + k), 0);
}

void orc_fill_normal_f32(uint64_t seed, uint64_t first, uint64_t n, int mathmode, float* out)
{
#pragma omp parallel for schedule(static)
    for (uint64_t k = 0; k < n; ++k) out[k] = orc_normal_f32(seed, (uint32_t)(first + k), 0, mathmode);
}

void orc_fill_array_uniform_f64(uint64_t seed, double lo, double hi, int ncomp, uint64_t n, uint64_t pitch,
                                double* out)
{
    for (int c = 0; c < ncomp; ++c) {
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < n; ++i)
            out[(uint64_t)c * pitch + i] = lo + (hi - lo) * orc_uniform01_f64(seed, (uint32_t)i, (uint64_t)c);
    }
}

void orc_fill_array_uniform_f32(uint64_t seed, float lo, float hi, int ncomp, uint64_t n, uint64_t pitch,
                                float* out)
{
    for (int c = 0; c < ncomp; ++c) {
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < n; ++i)
            out[(uint64_t)c * pitch + i] = lo + (hi - lo) * orc_uniform01_f32(seed, (uint32_t)i, (uint64_t)c);
    }
}

void orc_fill_array_normal_f64(uint64_t seed, double mean, double sd, int ncomp, uint64_t n, uint64_t pitch,
                               int mathmode, double* out)
{
    for (int c = 0; c < ncomp; ++c) {
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < n; ++i)
            out[(uint64_t)c * pitch + i] = mean + sd * orc_normal_f64(seed, (uint32_t)i, (uint64_t)c, mathmode);
    }
}

/* float planes: mean + sd * the float normal (random/detail/Backend.h:209-236, two words), as the float
 * array fill computes it */
void orc_fill_array_normal_f32(uint64_t seed, float mean, float sd, int ncomp, uint64_t n, uint64_t pitch,
                               int mathmode, float* out)
{
    for (int c = 0; c < ncomp; ++c) {
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < n; ++i)
            out[(uint64_t)c * pitch + i] = mean + sd * orc_normal_f32(seed, (uint32_t)i, (uint64_t)c, mathmode);
    }
}
