/*
 * oracle.h -- CPU restatement of aether's data-parallel hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (aether_amd/, include/)
 * links, includes or calls this code; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it, and only as the checker or as the
 * timed CPU baseline. Built with -ffp-contract=off so every floating-point
 * operation is rounded exactly where the source says.
 *
 * Each function cites the reference file:line whose behaviour it restates
 * (paths relative to the aether source tree).
 *
 * Pinning (what makes each function trustworthy; see tests/test_oracle_*.py):
 *   philox / bits / uniforms / keys : 315 core rows + 45 key rows of
 *        tests/random/golden_random.h (exact), plus the Random123 KAT.
 *   normals (libm route)             : golden_random.h normals within 1e-12 / 1e-5.
 *   own math (sin/cos/exp/log)       : MPFR golden rows (tests/bandmath/golden_*.h),
 *        ULP bound stated in the test; brute-force vs mpmath on random inputs.
 *   normals (own-math route)         : libm-route normals within 1e-12 / 1e-5.
 *   fused expression, Vec3           : reference-built fixtures (oracle/_ref) bit-exact
 *        for the libm route; mpmath closed form per element.
 *   compensated sums                 : exact-by-construction fixtures (<= 1 ulp),
 *        math.fsum on random data, Inf/NaN rules, pairCombine symmetry.
 *   DFT                              : closed forms (delta, constant, single tone),
 *        numpy.fft, Parseval.
 */
#ifndef AETHER_ORACLE_H
#define AETHER_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Philox4x32-10 and the draw primitives ---------------- */
/* random/detail/Philox.h:124-133, random/detail/Backend.h:144-150 */
void orc_philox4x32_10(const uint32_t ctr[4], uint32_t key0, uint32_t key1, uint32_t out[4]);
void orc_philox_block(uint64_t seed, uint32_t global, uint64_t counter, uint32_t out[4]);
/* random/detail/Backend.h:170-176 */
uint64_t orc_random_bits64(uint64_t seed, uint32_t global, uint64_t counter);
/* random/detail/Backend.h:184-195 (+ :123-134 conversions) */
double orc_uniform01_f64(uint64_t seed, uint32_t global, uint64_t counter);
double orc_u01_f64_from_word(uint32_t w);
float orc_u01_f32_from_word(uint32_t w);
float orc_uniform01_f32(uint64_t seed, uint32_t global, uint64_t counter);
/* random/detail/Backend.h:209-236; mathmode 0 = libm (the reference host route),
 * 1 = own math (the algorithm the device kernels use). */
double orc_normal_f64(uint64_t seed, uint32_t global, uint64_t counter, int mathmode);
float orc_normal_f32(uint64_t seed, uint32_t global, uint64_t counter, int mathmode);
/* random/detail/Backend.h:75-81, random/Generator.h:64-81 */
uint64_t orc_splitmix64(uint64_t x);
uint64_t orc_generator_key(uint64_t seed, uint64_t stream_id);
uint64_t orc_substream_key(uint64_t key, uint64_t k);

/* Bulk streams (bench inputs and parity): out[i] for i in [0, n). */
/* 32-bit words: word 2*i is lo(bits64(seed, i, 0)) ... see orc_fill_bits32. */
void orc_fill_bits32(uint64_t seed, uint64_t first, uint64_t n, uint32_t* out);
void orc_fill_uniform_f64(uint64_t seed, uint64_t first, uint64_t n, double* out);
void orc_fill_normal_f64(uint64_t seed, uint64_t first, uint64_t n, int mathmode, double* out);
void orc_fill_uniform_f32(uint64_t seed, uint64_t first, uint64_t n, float* out);
void orc_fill_normal_f32(uint64_t seed, uint64_t first, uint64_t n, int mathmode, float* out);
/* Array fill "key=seed, counter=component": plane c of an (ncomp, pitch) SoA
 * array gets lo + (hi-lo)*u01(seed, i, c) -- UniformLeaf::eval,
 * random/detail/DistributionLeaf.h:127-134 (seed already salted by caller). */
void orc_fill_array_uniform_f64(uint64_t seed, double lo, double hi, int ncomp, uint64_t n, uint64_t pitch,
                                double* out);
void orc_fill_array_uniform_f32(uint64_t seed, float lo, float hi, int ncomp, uint64_t n, uint64_t pitch,
                                float* out);
void orc_fill_array_normal_f64(uint64_t seed, double mean, double sd, int ncomp, uint64_t n, uint64_t pitch,
                               int mathmode, double* out);
void orc_fill_array_normal_f32(uint64_t seed, float mean, float sd, int ncomp, uint64_t n, uint64_t pitch,
                               int mathmode, float* out);

/* ---------------- own math (shared algorithm with the device) ---------- */
double orc_sin(double x);
double orc_cos(double x);
void orc_sincos(double x, double* s, double* c);
double orc_exp(double x);
double orc_log(double x);
float orc_sinf(float x);
float orc_cosf(float x);
float orc_expf(float x);
float orc_logf(float x);
float orc_tanf(float x);
float orc_atanf(float x);
float orc_asinf(float x);
float orc_acosf(float x);
float orc_tanhf(float x);
float orc_powf(float x, float y);
/* vector forms for the tests: fn 0=sin 1=cos 2=exp 3=log; mode 0 = libm, 1 = own */
double orc_tan(double x);
double orc_atan(double x);
double orc_asin(double x);
double orc_acos(double x);
double orc_tanh(double x);
double orc_pow(double x, double y);
double orc_rsqrt(double x);
double orc_sign(double x);
double orc_fmin(double a, double b);
double orc_fmax(double a, double b);
double orc_selgt(double a, double b);
/* fn: 0 sin 1 cos 2 exp 3 log 4 tan 5 atan 6 asin 7 acos 8 tanh 9 rsqrt 10 sign (mode 0: libm for 0-3) */
void orc_math_f64(int fn, int mode, uint64_t n, const double* x, double* y);
/* fn: 0 pow 1 fmin 2 fmax 3 ordered max (a > b ? a : b) */
void orc_math2_f64(int fn, uint64_t n, const double* x, const double* y, double* z);
void orc_math2_f32(int fn, uint64_t n, const float* x, const float* y, float* z);
void orc_math_f32(int fn, int mode, uint64_t n, const float* x, float* y);

/* ---------------- fused SoA expressions ---------------- */
/* d = a + b*c + sin(a)*exp(-b*b) per component (expr/Operations.h:25-53,
 * expr/nodes/Elementwise.h:291-357,513-533, expr/Assign.h:124-160).
 * Spelled exactly as aether evaluates it: Sum(Sum(a, Product(b,c)),
 * Product(CWiseSin(a), CWiseExp(CWiseScale(-1, Product(b,b))))). */
void orc_eval_cfg0_f64(const double* a, const double* b, const double* c, double* d, int ncomp, uint64_t n,
                       uint64_t pitch, int mathmode, int nthreads);
void orc_eval_cfg0_f32(const float* a, const float* b, const float* c, float* d, int ncomp, uint64_t n,
                       uint64_t pitch, int mathmode, int nthreads);
/* e = cross(a,b).dot(c) + sqrt(a.dot(a) + 1) (expr/nodes/Geometric.h:32-73,
 * expr/Reduce.h:150-190 right-nested folds, math/math.h:132). */
void orc_eval_vec3geo_f64(const double* a, const double* b, const double* c, double* e, uint64_t n, uint64_t pitch,
                          int nthreads);

/* ---------------- compensated accumulation ---------------- */
/* accum/atomic.h:142-149, :157-164, :171-188; accum/TwoSum.h:87-154 */
void orc_neumaier_add_f64(double* hi, double* comp, double term);
void orc_pair_combine_f64(double ah, double ac, double bh, double bc, double* oh, double* oc);
void orc_neumaier_add_f32(float* hi, float* comp, float term);
void orc_pair_combine_f32(float ah, float ac, float bh, float bc, float* oh, float* oc);
/* 32 interleaved Neumaier walks, then the 32-slot xor tree
 * (tests/test_CompensatedSum_common.h:71-80 -- the reference's canonical form). */
double orc_tree_sum32_f64(const double* x, uint64_t n);
float orc_tree_sum32_f32(const float* x, uint64_t n);
void orc_block_sum_f64(const double* x, uint64_t n, int nthreads, double* oh, double* oc);
/* The fixed, size-only reduction shape the device kernels use (restated
 * here so the device result can be checked bit for bit):
 *   tile = 2048 samples; lane l in 0..255 walks its 16-byte vectors in
 *   rounds: double: 2l, 2l+1, 512+2l, 512+2l+1, ... (4 rounds of 512);
 *   float: 4l..4l+3, 1024+4l..1024+4l+3 (2 rounds of 1024); from (+0,+0),
 *   skipping samples >= n; the 256 lane pairs are combined by pairCombine_
 *   in a butterfly over lanes (xor 1,2,4,...,32) inside each 64-lane wave,
 *   then ((w0,w1),(w2,w3)); tile partials are then cut into groups of 1024
 *   consecutive partials, each reduced by the halving tree (t, t+stride),
 *   stride = 512, 256, ..., 1, absent partner = nothing; the group results
 *   are reduced the same way until one pair is left. Result: hi and comp. */
void orc_device_sum_f64(const double* x, uint64_t n, double* hi, double* comp);
void orc_device_sum_f32(const float* x, uint64_t n, float* hi, float* comp);
/* Same shape, summing x*x (squared L2 norm before sqrt). */
void orc_device_sumsq_f64(const double* x, uint64_t n, double* hi, double* comp);
void orc_device_sumsq_f32(const float* x, uint64_t n, float* hi, float* comp);
/* float-float (Ff2-style, banded/Ff2.h:71-380) accumulation in the same shape. */
void orc_device_sum_ff_f32(const float* x, uint64_t n, float* hi, float* lo);
void orc_device_sumsq_ff_f32(const float* x, uint64_t n, float* hi, float* lo);
/* plain (no compensation) in the same shape. */
void orc_device_sum_plain_f64(const double* x, uint64_t n, double* hi);
void orc_device_sumsq_plain_f64(const double* x, uint64_t n, double* hi);
void orc_device_sum_plain_f32(const float* x, uint64_t n, float* hi);
void orc_device_sumsq_plain_f32(const float* x, uint64_t n, float* hi);
/* Exact-enough reference: long double Neumaier walk (tests compare with math.fsum). */
long double orc_sum_ld(const double* x, uint64_t n);

/* ---------------- DFT ---------------- */
/* Naive DFT in long double (tests/test_Fft_common.h:52-80 oracle form):
 * out[k] = scale * sum_j in[j] * exp(sign*2*pi*i*j*k/n), sign = -1 forward. */
void orc_dft_c128(const double* in, double* out, uint64_t n, int forward, double scale);

#ifdef __cplusplus
}
#endif
#endif
