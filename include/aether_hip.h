/*
 * aether_hip.h -- the C ABI of libaether_hip.so, the MI355X (gfx950) engine
 * behind aether's data-parallel hot path.
 *
 * Conventions (the reference's C seam, dsc/src/seam.h:1-57 and
 * dsc/src/plugin.cpp:22-39): every entry returns 0 on success, 1 on failure
 * with a per-thread message in aeth_last_error(), 2 when the request is valid
 * but this build does not support it. `stream` is a raw hipStream_t (NULL =
 * the null stream); every launch is stream-ordered and returns without
 * synchronising. Device pointers are plain pointers; SoA arrays are `ncomp`
 * planes of `pitch` elements, element (c, i) at c*pitch + i
 * (array/Array.h:80,100-110,279-295). Sample counts are limited by the
 * reference's 32-bit offset_t (index/Offset.h:73-91): ncomp*pitch <= 2^32-1.
 * No torch types; callable from C, C++, ctypes.
 */
#ifndef AETHER_HIP_H
#define AETHER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AETH_ABI_VERSION 1

enum { AETH_F64 = 0, AETH_F32 = 1 };

/* ------------------------------------------------------------------ the memory seam */
/* `aether-memory/1`: the reference's C seam between aether._core and its
 * device plugin, same table layout, field order, kinds and return codes
 * (dsc/src/seam.h:16-51), so the reference's loader (dsc/src/aether_core.cpp:
 * 95-113: dlopen, abi check, probe) binds this library unchanged. Kinds:
 * CPU, DEVICE (HIP device memory), PINNED (hipHostMalloc). FFT plans cover
 * C2C, R2C and C2R over axes of power-of-two length <= 2^34 (Stockham up to
 * 8192, four-step above, two levels past 2^26) or any other length < 2^32
 * (Bluestein); longer axes return 2. For R2C/C2R `shape` is
 * the real side's and the last axis is n/2 + 1 long on the complex side
 * (dsc/src/cufft_plan.h:9-12). fft_plan is refused while its stream is
 * captured; fft_exec allocates nothing and may be captured; fft_destroy is
 * safe during a capture (dsc/src/seam.h:48-50): the plan's device buffers are
 * freed at once when no capture stands in the way, otherwise by a later
 * destroy or drain(). chunk_release defers the same way. */
#define AETHER_MEMORY_ABI 2
enum { AETHER_MEM_CPU = 0, AETHER_MEM_DEVICE = 1, AETHER_MEM_PINNED = 2 };
enum { AETHER_FFT_C2C = 0, AETHER_FFT_R2C = 1, AETHER_FFT_C2R = 2 };
enum { AETHER_FFT_BACKWARD = 0, AETHER_FFT_FORWARD = 1, AETHER_FFT_ORTHO = 2 };

typedef struct AetherMemoryV1 {
    uint32_t abi;
    int (*probe)(int* device_count);
    int (*runtime_version)(int* version);
    int (*current_device)(int* dev);
    const char* (*last_error)(void);
    int (*chunk_alloc)(int kind, int dev, size_t bytes, void* stream, void** chunk, void** ptr);
    void (*chunk_release)(void* chunk);
    void (*drain)(void);
    int (*fill)(int kind, int dev, void* ptr, size_t bytes, const void* value, size_t width, void* stream);
    int (*copy)(int dst_kind, int dst_dev, void* dst, int src_kind, int src_dev, const void* src, size_t bytes,
                void* stream);
    int (*stream_wait)(int dev, void* producer, void* consumer);
    int (*stream_sync)(int dev, void* stream);
    int (*pointer_device)(const void* ptr, int* dev);
    int (*fft_plan)(int dev, int kind, int is_double, int norm, int forward, int ndim, const int64_t* shape, int naxes,
                    const int32_t* axes, void* stream, void** plan);
    int (*fft_exec)(void* plan, const void* in, void* out, void* stream);
    void (*fft_destroy)(void* plan);
} AetherMemoryV1;

/* dsc/src/plugin.cpp:156 (the plugin's only export) */
const AetherMemoryV1* aether_memory_v1(void);

/* ------------------------------------------------------------------ errors */
/* dsc/src/plugin.cpp:22-39 (thread-local last_error) */
const char* aeth_last_error(void);
int aeth_abi_version(void);
/* Number of HIP devices; 0 only if a driver answers and >= 1 device exists
 * (dsc/src/seam.h:24 probe). */
int aeth_probe(int* device_count);

/* ------------------------------------------------------------------ RNG */
/* Philox4x32-10 streams (random/detail/Backend.h:144-236). Element k of the
 * fill is element first+k of the stream:
 *   BITS32      : word (j & 1) of block(seed, j >> 1, 0)  -- randomBits64's lo then hi
 *   UNIFORM_F64 : uniform01<double>(seed, j, 0)            (Backend.h:184-195)
 *   UNIFORM_F32 : uniform01<float>(seed, j, 0)
 *   NORMAL_F64  : standardNormal<double>(seed, j, 0)       (Backend.h:209-236)
 *   NORMAL_F32  : standardNormal<float>(seed, j, 0)
 * j must stay below 2^32 (BITS32: 2^33). `out` must be 16-byte aligned for
 * the vector path (any alignment works, slower). */
enum { AETH_RNG_BITS32 = 0, AETH_RNG_UNIFORM_F64 = 1, AETH_RNG_UNIFORM_F32 = 2, AETH_RNG_NORMAL_F64 = 3,
       AETH_RNG_NORMAL_F32 = 4 };
int aeth_rng_fill(int kind, uint64_t seed, uint64_t first, uint64_t n, void* out, void* stream);

/* Array fill, the lowering of `view[i] = leaf` for aether's distribution
 * leaves (random/detail/DistributionLeaf.h:111-236): plane c, sample i gets
 *   UNIFORM: p0 + (p1 - p0) * uniform01(seed, i, c)
 *   NORMAL : p0 + p1 * standardNormal(seed, i, c)
 * `seed` is the leaf's seed after salting: Generator(s).uniform(...) draws
 * with seed = key ^ 0x243F6A8885A308D3, normal with key ^ 0x13198A2E03707344
 * (random/Generator.h:64-67, DistributionLeaf.h:117,184); the raw backend
 * stream ("key = seed") passes the seed itself. */
enum { AETH_DIST_UNIFORM = 0, AETH_DIST_NORMAL = 1 };
int aeth_array_fill(int dist, int dtype, uint64_t seed, double p0, double p1, int ncomp, uint64_t n,
                    uint64_t pitch, void* out, void* stream);

/* ------------------------------------------------------------------ fused evaluation */
/* Compensated accumulation modes (accum/atomic.h:142-188, banded/Ff2.h:132-145). */
enum { AETH_SUM_NONE = -1, AETH_SUM_NEUMAIER = 0, AETH_SUM_PLAIN = 1, AETH_SUM_FLOATFLOAT = 2 };

/* d = a + b*c + sin(a)*exp(-b*b) per component, one pass, fused with an
 * optional per-component sum of d in `sum_mode`. The lowering of
 *   dv[i] = av + bv*cv + cwiseSin(av)*cwiseExp(-1.0*(bv*bv))
 * (expr/Assign.h:124-160,264-271; expr/Operations.h:25-71;
 * expr/nodes/Elementwise.h:291-357,513-533) over whole arrays.
 * sums (device memory, may be NULL when sum_mode is NONE): 2*ncomp values of
 * dtype, (hi, comp) per component; the sum is hi + comp. */
int aeth_eval_cfg0(int dtype, const void* a, const void* b, const void* c, void* d, int ncomp, uint64_t n,
                   uint64_t pitch, int sum_mode, void* sums, void* stream);

/* aeth_eval_cfg0 for callers that cannot allocate: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. No
 * counterpart in the reference (which has no C entry for evaluation). Same
 * kernels and bits as aeth_eval_cfg0; the sums' tile partials live in a fixed
 * 4 MiB device array per GPU (held until the process ends) instead of a
 * stream-ordered allocation, so the call makes no hipMalloc/hipFree. Sizes
 * whose partials need more (about 2^28 samples of FP64 over 3 components)
 * return 1. Every call on one GPU shares that array: calls that overlap on one
 * GPU (different streams, threads) race. */
int aeth_eval_cfg0_noalloc(int dtype, const void* a, const void* b, const void* c, void* d, int ncomp, uint64_t n,
                           uint64_t pitch, int sum_mode, void* sums, void* stream);

/* Generic lowering of `dest = expr` (expr/Assign.h:124-160,264-271): each
 * output component is a PROGRAM -- a postfix tape over the planes it reads.
 * The tape is evaluated in registers for every sample; all planes a program
 * reads are loaded first (16-byte lanes), so one assignment is one kernel
 * whose HBM traffic is exactly the planes read plus the planes written.
 * Node order follows aether's evaluation order; nothing is contracted, so a
 * tape gives the same bits as the specialised kernels above for the same
 * tree (e.g. the cfg0 expression). Operations (stack effect in brackets):
 *   PUSH   arg = index into the program's `loads` [-> x]
 *   CONST  arg = index into `consts`                 [-> c]
 *   RNG    arg = index into `rngs`: a distribution leaf drawing at (seed,
 *          sample, comp) as aeth_array_fill does     [-> x]
 *   ADD SUB MUL DIV                                  [a b -> a op b]
 *   NEG SIN COS EXP LOG SQRT ABS                     [a -> f(a)]
 *   SIGN TAN ATAN ASIN ACOS TANH                     [a -> f(a)]
 *   FMIN FMAX                                        [a b -> fmin/fmax(a, b)]
 *   SELGT  (a > b ? a : b, the reductions' max)      [a b -> .]
 *   POW                                              [a b -> a^b]
 *   (expr/nodes/Elementwise.h:200-851, expr/Reduce.h:74-130, math/math.h:161-171)
 * Limits: 16 loads, 64 instructions, stack depth 8, 8 programs, 16 leaves,
 * 16 constants, 8 rng leaves per call. */
enum { AETH_OP_PUSH = 0, AETH_OP_CONST = 1, AETH_OP_RNG = 2, AETH_OP_ADD = 3, AETH_OP_SUB = 4, AETH_OP_MUL = 5,
       AETH_OP_DIV = 6, AETH_OP_NEG = 7, AETH_OP_SIN = 8, AETH_OP_COS = 9, AETH_OP_EXP = 10, AETH_OP_LOG = 11,
       AETH_OP_SQRT = 12, AETH_OP_ABS = 13, AETH_OP_SIGN = 14, AETH_OP_TAN = 15, AETH_OP_ATAN = 16,
       AETH_OP_ASIN = 17, AETH_OP_ACOS = 18, AETH_OP_TANH = 19, AETH_OP_FMIN = 20, AETH_OP_FMAX = 21,
       AETH_OP_SELGT = 22, AETH_OP_POW = 23 };
#define AETH_TAPE_MAX_LOADS 16
#define AETH_TAPE_MAX_INSNS 64
#define AETH_TAPE_MAX_DEPTH 8
#define AETH_TAPE_MAX_PROGRAMS 8
#define AETH_TAPE_MAX_LEAVES 16
#define AETH_TAPE_MAX_CONSTS 16
#define AETH_TAPE_MAX_RNGS 8
typedef struct {
    uint8_t op;
    uint8_t arg;
} aeth_insn;
typedef struct {
    uint8_t n_loads;
    uint8_t n_insns;
    uint8_t out_leaf;   /* destination leaf slot */
    uint8_t out_comp;   /* destination component (plane) */
    uint8_t load_leaf[AETH_TAPE_MAX_LOADS];
    uint8_t load_comp[AETH_TAPE_MAX_LOADS];
    aeth_insn insns[AETH_TAPE_MAX_INSNS];
} aeth_program;
typedef struct {
    uint64_t seed;  /* salted leaf seed, as aeth_array_fill */
    double p0, p1;  /* (lo, hi) or (mean, sd) */
    uint32_t comp;  /* Philox counter = component */
    uint32_t dist;  /* AETH_DIST_UNIFORM / AETH_DIST_NORMAL */
} aeth_rng_leaf;
/* leaves: device base pointers (planes at leaf_pitch elements apart); all
 * leaves share dtype and the sample count n. sums as in aeth_eval_cfg0, one
 * (hi, comp) pair per program in program order. */
int aeth_eval_tape(int dtype, int n_programs, const aeth_program* programs, int n_leaves, void* const* leaves,
                   const uint64_t* leaf_pitch, int n_consts, const double* consts, int n_rngs,
                   const aeth_rng_leaf* rngs, uint64_t n, int sum_mode, void* sums, void* stream);

/* Trees of any size: programs whose loads, instructions and arguments are
 * unbounded (32-bit indices), any number of leaves, constants and random
 * leaves, up to 64 programs. What fits the limits above runs as aeth_eval_tape
 * (one kernel). A larger program is cut into pieces that fit: its largest
 * subtrees are evaluated first into temporary planes of dtype (stream-ordered
 * allocations), the pieces above them load those planes. Every node still
 * rounds once to dtype, so the values are the same as a single evaluation;
 * sums apply to the outputs as above. Programs of an assignment that is cut
 * run one after another in component order. No counterpart beyond the
 * reference's templates, which lower any tree (expr/Assign.h:124-160). */
typedef struct {
    uint8_t op;
    uint8_t pad[3];
    uint32_t arg;
} aeth_insn_ex;
typedef struct {
    uint32_t n_loads;
    uint32_t n_insns;
    uint32_t out_leaf;
    uint32_t out_comp;
    const uint32_t* load_leaf;  /* n_loads */
    const uint32_t* load_comp;  /* n_loads */
    const aeth_insn_ex* insns;  /* n_insns */
} aeth_program_ex;
int aeth_eval_tape_ex(int dtype, int n_programs, const aeth_program_ex* programs, int n_leaves, void* const* leaves,
                      const uint64_t* leaf_pitch, int n_consts, const double* consts, int n_rngs,
                      const aeth_rng_leaf* rngs, uint64_t n, int sum_mode, void* sums, void* stream);

/* Each distinct tape structure is lowered at run time (hipRTC) to its own
 * straight-line kernel, compiled once per process and cached; the resident
 * interpreter kernel runs a tape when the runtime compiler is unavailable or
 * AETHER_TAPE_INTERPRET is set. Both give the same bits. With the environment
 * variable AETHER_KERNEL_CACHE naming a writable directory, compiled kernels
 * are also kept there (one code object per architecture and generated source)
 * and later processes load them instead of compiling. compiled: hipRTC
 * compiles; cache_hits: kernels found in this process's cache or on disk. */
int aeth_tape_jit_stats(uint64_t* compiled, uint64_t* cache_hits, uint64_t* fallbacks);
/* Per-launch timing of the tape kernels (measurement; no counterpart in the
 * reference). While enabled, each tape kernel launched outside a stream
 * capture is bracketed by two HIP events on its stream (the sums' group tree
 * that follows it is not). aeth_tape_timing_read sets *count to the number of
 * launches recorded and not yet read, then waits for the oldest `capacity` of
 * them and hands back their elapsed milliseconds in launch order (capacity 0
 * only counts); the rest stay pending for the next read. */
int aeth_tape_timing(int enable);
int aeth_tape_timing_read(float* ms, int capacity, int* count);

/* Compile and load, without running, every kernel that aeth_eval_tape_ex
 * with these arguments would launch (the same structure, launch shape and
 * cache policy), so that the first real call pays no runtime compile (0.1-0.5 s
 * per tape structure and process). Nothing is allocated, read or written; the
 * pointers only decide which planes alias. No counterpart in the reference,
 * whose kernels are compiled ahead of time: it is for callers whose first
 * call must not stall, such as servers warming up and timed loops. */
int aeth_tape_prepare(int dtype, int n_programs, const aeth_program_ex* programs, int n_leaves,
                      void* const* leaves, const uint64_t* leaf_pitch, int n_consts, const double* consts,
                      int n_rngs, const aeth_rng_leaf* rngs, uint64_t n, int sum_mode, void* stream);

/* Compile (without loading or running) the kernel of a tape for `arch`
 * ("gfx950"); works without a GPU. code_bytes: size of the code object. */
int aeth_tape_jit_build(int dtype, int sum_mode, int n_programs, const aeth_program* programs, const char* arch,
                        uint64_t* code_bytes);

/* Launch-shape autotuner (no counterpart in the reference, whose kernels take a caller-chosen grid:
 * expr/Assign.h:124-160). Times the runtime-compiled kernel of this assignment (the arguments of
 * aeth_eval_tape, without sums) at each candidate shape -- samples per workgroup tile 1024, 2048 or 4096,
 * 16- or 8-byte vector accesses, and a minimum of waves per SIMD asked of the register allocator (none, or
 * 8) -- on `stream`, on the caller's arrays, and keeps the fastest for this
 * tape structure on this device for the rest of the process: later aeth_eval_tape calls of the same
 * structure without sums launch with it (calls with sums keep the 2048-sample, 16-byte shape, whose
 * reduction order the sums are defined by; aeth_tape_tune_sums below tunes their occupancy target). The destination holds the assignment's result afterwards.
 * Refused inside a stream capture and for an assignment whose destination is also one of its inputs.
 * Synchronises `stream`. chosen (may be NULL): the shape picked and its average time per launch. */
typedef struct {
    uint32_t tile;
    uint32_t vector_bytes;
    uint32_t waves_per_simd; /* 0: the compiler's choice */
    float ms;
} aeth_tape_shape;
int aeth_tape_tune(int dtype, int n_programs, const aeth_program* programs, int n_leaves, void* const* leaves,
                   const uint64_t* leaf_pitch, int n_consts, const double* consts, int n_rngs,
                   const aeth_rng_leaf* rngs, uint64_t n, void* stream, aeth_tape_shape* chosen);
/* The same for an assignment with sums (sum_mode != AETH_SUM_NONE; the arguments of aeth_eval_tape):
 * the 2048-sample, 16-byte shape stays (it defines the sums' bits) and only the register allocator's
 * occupancy target varies (none, 4 or 8 waves per SIMD), which changes no operation. Later calls of the
 * structure with that sum mode launch with the pick. `sums` receives the assignment's sums. No
 * counterpart in the reference. */
int aeth_tape_tune_sums(int dtype, int n_programs, const aeth_program* programs, int n_leaves,
                        void* const* leaves, const uint64_t* leaf_pitch, int n_consts, const double* consts,
                        int n_rngs, const aeth_rng_leaf* rngs, uint64_t n, int sum_mode, void* sums, void* stream,
                        aeth_tape_shape* chosen);

/* e = cross(a, b).dot(c) + sqrt(a.dot(a) + 1) over Array<double,3> a, b, c
 * (pitch `pitch`) into a scalar plane e (expr/nodes/Geometric.h:32-73,
 * expr/Expression.h:162-173, expr/Reduce.h:150-190, math/math.h:132). */
int aeth_eval_vec3geo(const void* a, const void* b, const void* c, void* e, uint64_t n, uint64_t pitch,
                      void* stream);

/* Per-component sum (square = 0) or sum of squares (square = 1) of an SoA
 * array in the fixed reduction shape; out: 2*ncomp values (hi, comp|lo). */
int aeth_reduce(int dtype, int mode, int square, const void* x, int ncomp, uint64_t n, uint64_t pitch, void* out,
                void* stream);

/* Elementwise math (math/math.h:228-287, math/explog.h:39,61).
 * fn: 0 sin, 1 cos, 2 exp, 3 log: y[i] = fn(x[i]).
 * fn 4 (fused stress form): y holds 4 planes of n: sin, cos, exp, log of x. */
enum { AETH_FN_SIN = 0, AETH_FN_COS = 1, AETH_FN_EXP = 2, AETH_FN_LOG = 3, AETH_FN_ALL4 = 4 };
int aeth_math(int fn, int dtype, const void* x, void* y, uint64_t n, void* stream);

/* ------------------------------------------------------------------ batched FFT */
/* Many complex transforms of uneven power-of-two lengths (1..8192) in ONE
 * launch driven by a device-resident work-descriptor table. Replaces, for
 * power-of-two lines, fft::c2c (fft/Fft.h:186-197) on the host and the
 * seam's fft_plan/fft_exec/fft_destroy (dsc/src/seam.h:41-50,
 * dsc/src/cufft_plan.h:122-320) on the device. Data are interleaved
 * (re, im) pairs; signal i occupies `lengths[i]` consecutive complex elements
 * starting at complex offset in_offsets[i] (input) / out_offsets[i] (output).
 * prec: C128 (double pairs), C64 (float pairs), C64_FF (float pairs in
 * memory, float-float arithmetic inside: the compensated FP32 mode),
 * C64_WIDE (float pairs in memory, FP64 arithmetic inside). norm follows
 * numpy / fft::Norm
 * (fft/Fft.h:58,73-76): BACKWARD scales the inverse by 1/N, FORWARD the
 * forward by 1/N, ORTHO both by 1/sqrt(N). Plan creation allocates and
 * copies (like the seam's); exec allocates nothing and may be captured.
 * In-place (in == out with equal offsets) is allowed. Lengths that are not
 * powers of two, or above 8192, return 2. */
enum { AETH_FFT_C128 = 0, AETH_FFT_C64 = 1, AETH_FFT_C64_FF = 2, AETH_FFT_C64_WIDE = 3 };
enum { AETH_FFT_BACKWARD = 0, AETH_FFT_FORWARD = 1, AETH_FFT_ORTHO = 2 };
int aeth_fft_batch_plan(int prec, uint64_t count, const uint64_t* lengths, const uint64_t* in_offsets,
                        const uint64_t* out_offsets, void** plan);
int aeth_fft_batch_exec(void* plan, const void* in, void* out, int forward, int norm, void* stream);
/* Safe during a stream capture, like the seam's fft_destroy. */
int aeth_fft_batch_destroy(void* plan);

/* ------------------------------------------------------------------ serving */
/* A continuous-batching job server on one GPU (no counterpart in the
 * reference, which has no serving loop). Jobs:
 *   EVAL  : a, b, c = Array<T,3>(n) of uniforms on (-1,1) drawn from seeds
 *           seed, seed+1, seed+2 (counter = component); d = a + b*c +
 *           sin(a)*exp(-b*b); result: per-component compensated sums of d
 *           (Neumaier for F64, float-float for F32) as 3 (hi, comp) pairs.
 *   SAMPLE: n standard normals standardNormal<T>(seed, i, 0); result: (hi,
 *           comp) of their sum and of their sum of squares.
 *   FFT   : n (1..64) signals of lengths 2^fft_log2n[k], interleaved entries
 *           standardNormal<T>(seed, j, 0); forward transform (complex128 for
 *           F64, complex64 with float-float arithmetic for F32); result: (hi,
 *           comp) of the sum of every output real and imaginary part.
 * Each job's result is bit for bit the result of aeth_run_job_plain on the
 * same job: batching never changes a value. Arrays live in a device pool of
 * hbm_budget_bytes; a job is admitted (in arrival order) when all its buffers
 * fit and released when its batch completes. fp32_compensated runs every job
 * in FP32 storage with compensated accumulation. */
enum { AETH_JOB_EVAL = 0, AETH_JOB_SAMPLE = 1, AETH_JOB_FFT = 2 };
typedef struct {
    uint32_t kind;
    uint32_t dtype;          /* AETH_F64 / AETH_F32 */
    uint64_t n;              /* EVAL / SAMPLE: samples; FFT: number of signals */
    uint64_t seed;
    double arrival;          /* seconds after the start of aeth_server_replay */
    uint8_t fft_log2n[64];
} aeth_job;
typedef struct {
    uint64_t id;
    uint32_t kind;
    int32_t status;          /* 0 ok; 1 failed; 2 cannot fit the budget */
    double value[8];         /* (hi, comp) pairs as described above, widened to double */
    double submit_time;      /* seconds (server clock; replay: since its start) */
    double done_time;
} aeth_job_result;
typedef struct {
    uint64_t hbm_budget_bytes;
    uint32_t streams;        /* batches in flight, 1..4 */
    uint32_t max_batch_jobs; /* coalescing cap per batch */
    uint32_t fp32_compensated;
    uint32_t flags;          /* AETH_SERVER_GRAPH_BATCHES, AETH_SERVER_TUNE: see below */
} aeth_server_config;
/* flags: AETH_SERVER_GRAPH_BATCHES -- every batch of a stream slot is ONE
 * replay of a graph captured when the server is created: a staging kernel
 * that pulls the batch's descriptor tables from pinned host memory, then
 * persistent kernels (fixed grids, descriptor counts read on the device) for
 * the fills, fused evaluations, FFTs, reductions and trees, and the copy of
 * the results. The same kernels' per-unit bodies run either way, so results
 * are bit for bit those of direct launches.
 * AETH_SERVER_TUNE (with AETH_SERVER_GRAPH_BATCHES) -- the autotuner of the
 * batched launch: each slot's graph is captured for persistent grids of 2,
 * 4, 8 and 16 blocks per CU; the first batches rotate over them, timed with
 * events, and once each has 8 timed batches the lowest median time per
 * algorithmic byte is kept for the rest of the server's life. Every grid
 * gives the same bits. Without the flag, 8 blocks per CU. */
enum { AETH_SERVER_GRAPH_BATCHES = 1, AETH_SERVER_TUNE = 2 };
typedef struct {
    uint64_t submitted, completed, failed, batches;
    uint64_t pool_bytes_in_use, pool_bytes_budget;
} aeth_server_stats;
int aeth_server_create(const aeth_server_config* cfg, void** server);
int aeth_server_submit(void* server, const aeth_job* job, uint64_t* id);
int aeth_server_poll(void* server, aeth_job_result* out, uint32_t max, uint32_t* count);
int aeth_server_wait_idle(void* server);
int aeth_server_get_stats(void* server, aeth_server_stats* stats);
/* The graph form's persistent grid in use (blocks per CU) and whether the
 * tuner has settled (AETH_SERVER_TUNE; without it: 8 and 0). No counterpart
 * in the reference. */
int aeth_server_tuning(void* server, uint32_t* blocks_per_cu, uint32_t* settled);
int aeth_server_destroy(void* server);
/* Open-loop replay: submits jobs[i] at jobs[i].arrival seconds after the call
 * starts, waits for all, results[i] is job i's result. */
int aeth_server_replay(void* server, const aeth_job* jobs, uint64_t count, aeth_job_result* results);
/* The plain per-call form of one job: its own hipMalloc/hipFree, the single-
 * call entries above, synchronous (how a caller of the reference serves a
 * request). */
int aeth_run_job_plain(const aeth_job* job, aeth_job_result* out, void* stream);
/* The same, with the job's arrays carved from a caller-owned device workspace
 * (256-byte aligned pieces; an EVAL job of n samples needs about 12*n*esize +
 * 1 MiB) instead of hipMalloc/hipFree per call: the plain form without its
 * allocation cost (an FFT job still builds its plan per call). */
int aeth_run_job_plain_ws(const aeth_job* job, aeth_job_result* out, void* workspace, uint64_t workspace_bytes,
                          void* stream);

/* ------------------------------------------------------------------ accumulation */
/* The device primitives are in include/aether/device/accum.hpp for HIP kernels of a caller's own;
 * these entries run them over whole arrays, one lane per element. */

/* Keyed sum: target[keys[l]] += terms[l] for l < n, with ONE atomic per run of lanes of a wavefront
 * that share a slot (accum/run.h:1-461; wave64 forms):
 *   AETH_RUN_ALIGNED : the slot is equal over every aligned group of `width` lanes (2..64), given
 *                      that lane l's own sample is l + own; the vote checks 64 consecutive own samples
 *                      starting on a multiple of 64;
 *   AETH_RUN_TESTED  : one vote checks that the full wave shares one slot;
 *   AETH_RUN_SORTED  : keys are non-decreasing in lane order; each maximal run commits once.
 * A wave that fails its vote, or is not full, commits lane by lane: correct for any keys. With
 * `compensated`, each commit adds its two-sum residual into comp[key] (delivered: target + comp).
 * Keys outside [0, slots) take no part. commits (device, may be NULL): (run, lane) commit counters,
 * incremented atomically (diagnostics). dtype: AETH_F64 / AETH_F32 (terms, target, comp). */
enum { AETH_RUN_ALIGNED = 0, AETH_RUN_TESTED = 1, AETH_RUN_SORTED = 2 };
int aeth_keyed_sum(int dtype, int form, int width, int compensated, const int64_t* keys, const void* terms,
                   uint64_t n, int64_t own, uint64_t slots, void* target, void* comp, unsigned long long* commits,
                   void* stream);

/* Accumulation plane (accum/AccumPlane.h:36-470): value[d*pitch + i] (+ comp) += terms[d*term_pitch + i]
 * for i < n and d < vd (1, 2, 3, 4, 6 or 8 lanes), components in ascending order. COMPENSATED is one
 * hardware atomic add plus one of its two-sum residual into comp (accum/atomic.h:120-137): safe for
 * concurrent producers on any streams, the delivered value + comp independent of their order to within
 * the residuals' own rounding. SERIALIZED is a plain read-add-write (comp unused): the caller orders
 * the producers of a sample. */
enum { AETH_PLANE_COMPENSATED = 0, AETH_PLANE_SERIALIZED = 1 };
int aeth_plane_add(int policy, int vd, const double* terms, uint64_t term_pitch, uint64_t n, double* value,
                   double* comp, uint64_t pitch, void* stream);

/* Append counter (accum/AppendCounter.h:30-192): every i < n with x[i] > threshold claims a slot of
 * `out` (capacity places) from the device counter and writes i there; a claim past capacity sets
 * *overflow = 1 and writes nothing (the counter itself keeps the raw count). claim: the granularity
 * of the atomics on the counter -- LANE (one per claim, the reference's tryAppend), WAVE (one per
 * wavefront; its claimants take consecutive slots in lane order) or TILE (one per block of 16384
 * elements, which keep their order). counter and overflow are device memory the caller zeroes
 * between steps; which slots a wave or tile gets depends on the order the atomics land. */
enum { AETH_CLAIM_LANE = 0, AETH_CLAIM_WAVE = 1, AETH_CLAIM_TILE = 2 };
int aeth_append_where(int dtype, const void* x, uint64_t n, double threshold, uint32_t* out, uint64_t capacity,
                      uint32_t* counter, uint32_t* overflow, int claim, void* stream);

/* ------------------------------------------------------------------ residency */
/* How a copy between two memory places travels (residency/PeerAccess.h:
 * 57-129, peerRoute): HOST unless both places are DEVICE memory; STAGED when
 * force_staging is set (through a pinned bounce buffer, the path a GPU pair
 * without peer access takes); SAME for one device; P2P when each device can
 * reach the other (peer access is then enabled both ways, once per process);
 * STAGED otherwise. kinds are the seam's AETHER_MEM_* kinds. */
enum { AETH_ROUTE_SAME = 0, AETH_ROUTE_P2P = 1, AETH_ROUTE_STAGED = 2, AETH_ROUTE_HOST = 3 };
int aeth_peer_route(int kind_a, int dev_a, int kind_b, int dev_b, int force_staging, int* route);
/* Completion markers for asynchronous transfers (residency/Transport.h:
 * 175-205, the transport's lastEvent/wait/fence): record creates an event
 * and records it on `stream`; synchronize waits for it; destroy frees it. */
int aeth_event_record(void* stream, void** event);
int aeth_event_synchronize(void* event);
int aeth_event_destroy(void* event);

/* ------------------------------------------------------------------ runtime evaluator */
/* eval::runtimeEval over RuntimeViews (eval/Runtime.h:1-342, view/RuntimeView.h:71-120): elementwise
 * ops over strided N-d views of device memory, each operand addressed through its own element
 * strides over the output's logical row-major shape (rank 0..6; a rank-0 view is one element):
 *   ASSIGN     out = a          (the one op that may cast double <-> float)
 *   ADD / SUB  out = a +- b
 *   SCALE      out = s * a
 *   ADD_SCALED / SUB_SCALED   out = a +- s * b   (s rounded to the operands' type; not contracted)
 * Shapes must agree; dtypes (AETH_F64 / AETH_F32) must agree except for ASSIGN. Dense row-major
 * operands run 16-byte lanes; other layouts the general strided kernel. Strides are in elements,
 * non-negative. */
#define AETH_RUNTIME_RANK_CAP 6
typedef struct {
    void* data;
    int32_t dtype;
    int32_t rank;
    uint32_t extents[AETH_RUNTIME_RANK_CAP];
    int64_t strides[AETH_RUNTIME_RANK_CAP];
} aeth_runtime_view;
enum { AETH_RT_ASSIGN = 0, AETH_RT_ADD = 1, AETH_RT_SUB = 2, AETH_RT_SCALE = 3, AETH_RT_ADD_SCALED = 4,
       AETH_RT_SUB_SCALED = 5 };
int aeth_runtime_eval(int op, const aeth_runtime_view* a, const aeth_runtime_view* b, const aeth_runtime_view* out,
                      double scalar, void* stream);

#ifdef __cplusplus
}
#endif
#endif
