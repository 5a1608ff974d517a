// serve.hip -- continuous-batching job server over the hot path (one GPU).
//
// The reference has no serving loop; this is the serving extension. Jobs
// (evaluate / sample / FFT) arrive from any thread. One scheduler thread
// keeps up to `streams` batches in flight: whenever a stream
// slot is free it takes the pending jobs in arrival order, admits each one
// whose buffers fit in the device pool (an HBM arena carved by a first-fit
// allocator: the budget), and launches the whole batch as a handful of
// descriptor-driven kernels -- fill inputs, fused eval + sums, sample sums,
// FFTs, checksums, per-job trees -- then one small copy of the results to
// pinned memory and an event. Completed slots free their jobs' memory, which
// admits the next jobs (residency: chunk/Chunk.h:111-160, residency/*).
//
// Every kernel runs the per-tile bodies of the single-call path (units.hpp,
// fft.hip) with each job's own sample count, so a job's result is bit for
// bit the same alone and inside any batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/aether_hip.h"
#include "accum.hpp"
#include "common.hpp"
#include "fft.hpp"
#include "rng.hpp"
#include "units.hpp"

namespace aeth {
namespace serve {

using units::kGroup;
using units::kThreads;
using units::kTile;

constexpr int kFillPerBlock = 4096;

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------- descriptors
struct FillDesc {
    void* out;
    uint64_t n;      // samples per component
    uint64_t pitch;  // plane pitch (elements)
    uint64_t seed;
    double p0, p1;
    uint64_t unit_start;
    uint32_t ncomp;
    uint32_t dtype;  // AETH_F64 / AETH_F32
    uint32_t kind;   // 0: uniform array leaf (counter = component); 1: normal stream (counter 0)
    uint32_t pad_;
};

struct EvalDesc {
    const void* a;
    const void* b;
    const void* c;
    void* d;
    void* part;      // 3 x tiles pairs
    uint64_t pitch;
    uint64_t unit_start;
    uint32_t n, tiles, dtype, pad_;
};

struct RedDesc {
    const void* x;
    void* part;      // tiles pairs
    uint64_t unit_start;
    uint32_t n, tiles, dtype, square;
};

struct TreeDesc {
    const void* part;
    void* scratch;
    double* out;     // (hi, comp) widened to double
    uint32_t tiles, dtype;
};

// The batch's counts, at the start of the slot's staging area on both sides. Kernels read them from device
// memory, so a launch with a fixed grid (one node of a graph captured once per slot) serves every batch.
struct BatchHeader {
    uint32_t fills, evals, reds, trees;
    uint32_t fft_ndesc[2]; // complex128, complex64 float-float
    uint32_t jobs, dtypes; // dtypes: bit AETH_F64 / AETH_F32 set when some job of the batch has that dtype
    uint64_t fill_units, eval_units, red_units;
    uint64_t region_bytes[9]; // bytes used in each region of the staging layout (header first)
};

// Which dtypes a batch kernel is compiled for: a batch whose jobs all have one dtype runs the variant of that
// dtype (fewer registers: the other dtype's arithmetic is not compiled in), a mixed batch the variant for
// both. The per-unit bodies are the same code either way, so the bits do not depend on the variant.
enum DtVariant : int { kDtF64 = 0, kDtF32 = 1, kDtBoth = 2 };

template<int DT>
__device__ __forceinline__ bool is_f64(uint32_t dtype)
{
    if constexpr (DT == kDtBoth) return dtype == AETH_F64;
    else return DT == kDtF64;
}

template<class D>
__device__ __forceinline__ uint32_t find_desc(const D* d, uint32_t count, uint64_t unit)
{
    uint32_t lo = 0, hi = count - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (d[mid].unit_start <= unit) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One unit: kFillPerBlock consecutive samples of one component (units run component-major), so the
// component and the sample range come from one 32-bit division per block, and a wave's stores are
// consecutive.
template<int DT>
__device__ __forceinline__ void fill_unit(const FillDesc& D, uint64_t u)
{
    const uint32_t chunks = (uint32_t)((D.n + kFillPerBlock - 1) / kFillPerBlock);
    const uint32_t local = (uint32_t)(u - D.unit_start);
    const uint32_t c = local / chunks;
    const uint32_t i0 = (local - c * chunks) * (uint32_t)kFillPerBlock;
    const uint64_t stop = (uint64_t)i0 + kFillPerBlock;
    const uint32_t end = (uint32_t)(stop < D.n ? stop : D.n);
    if (D.kind == 0) {
        if (is_f64<DT>(D.dtype)) {
            double* out = static_cast<double*>(D.out) + (uint64_t)c * D.pitch;
            for (uint64_t i = (uint64_t)i0 + threadIdx.x; i < end; i += kThreads)
                out[i] = D.p0 + (D.p1 - D.p0) * rng::u01_from_word_d(rng::block(D.seed, (uint32_t)i, c).x);
        } else {
            float* out = static_cast<float*>(D.out) + (uint64_t)c * D.pitch;
            const float p0 = (float)D.p0, p1 = (float)D.p1;
            for (uint64_t i = (uint64_t)i0 + threadIdx.x; i < end; i += kThreads)
                out[i] = p0 + (p1 - p0) * rng::u01_from_word_f(rng::block(D.seed, (uint32_t)i, c).x);
        }
    } else {
        for (uint64_t i = (uint64_t)i0 + threadIdx.x; i < end; i += kThreads) {
            const rng::U4 b = rng::block(D.seed, (uint32_t)i, 0);
            if (is_f64<DT>(D.dtype)) static_cast<double*>(D.out)[i] = rng::normal_from_block_d(b);
            else static_cast<float*>(D.out)[i] = rng::normal_from_block_f(b);
        }
    }
}

template<int DT>
__global__ __launch_bounds__(kThreads) void fill_batch_kernel(const BatchHeader* __restrict__ hdr,
                                                             const FillDesc* __restrict__ descs)
{
    const uint32_t count = hdr->fills;
    const uint64_t units = hdr->fill_units;
    for (uint64_t u = blockIdx.x; u < units; u += gridDim.x)
        fill_unit<DT>(descs[find_desc(descs, count, u)], u);
}

template<int DT>
__device__ __forceinline__ void eval_unit(const EvalDesc& D, uint64_t u, acc::Pair<double>* ldsd,
                                          acc::Pair<float>* ldsf)
{
    const uint32_t local = (uint32_t)(u - D.unit_start);
    const uint32_t comp = local / D.tiles, tile = local - comp * D.tiles;
    if (is_f64<DT>(D.dtype)) {
        const auto p = units::cfg0_tile<double, acc::kNeumaier, true, true, false>(
            static_cast<const double*>(D.a), static_cast<const double*>(D.b), static_cast<const double*>(D.c),
            static_cast<double*>(D.d), D.n, D.pitch, comp, tile);
        const auto r = acc::block_combine<acc::kNeumaier>(p, ldsd);
        if (threadIdx.x == 0) static_cast<acc::Pair<double>*>(D.part)[local] = r;
    } else {
        const auto p = units::cfg0_tile<float, acc::kFloatFloat, true, true, false>(
            static_cast<const float*>(D.a), static_cast<const float*>(D.b), static_cast<const float*>(D.c),
            static_cast<float*>(D.d), D.n, D.pitch, comp, tile);
        const auto r = acc::block_combine<acc::kFloatFloat>(p, ldsf);
        if (threadIdx.x == 0) static_cast<acc::Pair<float>*>(D.part)[local] = r;
    }
}

template<int DT>
__global__ __launch_bounds__(kThreads) void eval_batch_kernel(const BatchHeader* __restrict__ hdr,
                                                             const EvalDesc* __restrict__ descs)
{
    __shared__ acc::Pair<double> ldsd[4];
    __shared__ acc::Pair<float> ldsf[4];
    const uint32_t count = hdr->evals;
    const uint64_t units = hdr->eval_units;
    for (uint64_t u = blockIdx.x; u < units; u += gridDim.x)
        eval_unit<DT>(descs[find_desc(descs, count, u)], u, ldsd, ldsf);
}

template<int DT>
__device__ __forceinline__ void reduce_unit(const RedDesc& D, uint64_t u, acc::Pair<double>* ldsd,
                                            acc::Pair<float>* ldsf)
{
    const uint32_t tile = (uint32_t)(u - D.unit_start);
    if (is_f64<DT>(D.dtype)) {
        const double* x = static_cast<const double*>(D.x);
        const auto p = D.square ? units::reduce_tile<double, acc::kNeumaier, true, true, false>(x, D.n, 0, 0, tile)
                                : units::reduce_tile<double, acc::kNeumaier, false, true, false>(x, D.n, 0, 0, tile);
        const auto r = acc::block_combine<acc::kNeumaier>(p, ldsd);
        if (threadIdx.x == 0) static_cast<acc::Pair<double>*>(D.part)[tile] = r;
    } else {
        const float* x = static_cast<const float*>(D.x);
        const auto p = D.square ? units::reduce_tile<float, acc::kFloatFloat, true, true, false>(x, D.n, 0, 0, tile)
                                : units::reduce_tile<float, acc::kFloatFloat, false, true, false>(x, D.n, 0, 0, tile);
        const auto r = acc::block_combine<acc::kFloatFloat>(p, ldsf);
        if (threadIdx.x == 0) static_cast<acc::Pair<float>*>(D.part)[tile] = r;
    }
}

template<int DT>
__global__ __launch_bounds__(kThreads) void reduce_batch_kernel(const BatchHeader* __restrict__ hdr,
                                                               const RedDesc* __restrict__ descs)
{
    __shared__ acc::Pair<double> ldsd[4];
    __shared__ acc::Pair<float> ldsf[4];
    const uint32_t count = hdr->reds;
    const uint64_t units = hdr->red_units;
    for (uint64_t u = blockIdx.x; u < units; u += gridDim.x)
        reduce_unit<DT>(descs[find_desc(descs, count, u)], u, ldsd, ldsf);
}

// The group tree of the single-call path, all levels inside one block.
template<class T, int M>
__device__ __forceinline__ void tree_one(const TreeDesc& D, acc::Pair<T>* s)
{
    const uint32_t t = threadIdx.x;
    const acc::Pair<T>* src = static_cast<const acc::Pair<T>*>(D.part);
    acc::Pair<T>* out = static_cast<acc::Pair<T>*>(D.scratch);
    uint32_t count = D.tiles;
    while (count > 1) {
        const uint32_t groups = (count + kGroup - 1) / kGroup;
        for (uint32_t g = 0; g < groups; ++g) {
            const uint32_t m = min((uint32_t)kGroup, count - g * kGroup);
            if (t < m) s[t] = src[g * kGroup + t];
            if (t + 512 < m) s[t + 512] = src[g * kGroup + t + 512];
            units::group_halving<T, M>(s, m);
            if (t == 0) out[g] = s[0];
            __threadfence_block();
            __syncthreads();
        }
        src = out;
        out += groups;
        count = groups;
    }
    if (t == 0) {
        const acc::Pair<T> r = src[0];
        D.out[0] = (double)r.h;
        D.out[1] = (double)r.c;
    }
}

__global__ __launch_bounds__(512) void tree_batch_kernel(const BatchHeader* __restrict__ hdr,
                                                         const TreeDesc* __restrict__ descs)
{
    __shared__ acc::Pair<double> s[kGroup];
    const uint32_t count = hdr->trees;
    for (uint32_t k = blockIdx.x; k < count; k += gridDim.x) {
        __syncthreads(); // the previous tree is done with the LDS
        const TreeDesc& D = descs[k];
        if (D.dtype == AETH_F64) tree_one<double, acc::kNeumaier>(D, s);
        else tree_one<float, acc::kFloatFloat>(D, reinterpret_cast<acc::Pair<float>*>(s));
    }
}

// The staging area's used bytes, region by region, from the pinned host copy (read through its device
// mapping) into device memory: one kernel in place of a copy whose size the batch decides.
constexpr int kRegions = 9;
__global__ __launch_bounds__(kThreads) void stage_in_kernel(const unsigned char* __restrict__ host,
                                                           unsigned char* __restrict__ dev,
                                                           const uint64_t* __restrict__ region_offset)
{
    __shared__ uint64_t used[kRegions];
    if (threadIdx.x < kRegions) used[threadIdx.x] = reinterpret_cast<const BatchHeader*>(host)->region_bytes[threadIdx.x];
    __syncthreads();
    for (int r = 0; r < kRegions; ++r) {
        const uint64_t at = region_offset[r], vecs = (used[r] + 15) / 16;
        const uint4* src = reinterpret_cast<const uint4*>(host + at);
        uint4* dst = reinterpret_cast<uint4*>(dev + at);
        for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < vecs; i += (uint64_t)gridDim.x * kThreads)
            dst[i] = src[i];
    }
}

// ---------------------------------------------------------------- device pool
class Pool {
public:
    explicit Pool(uint64_t bytes) : size_(bytes)
    {
        check(hipMalloc(&base_, bytes), "hipMalloc(pool)", bytes);
        free_[0] = bytes;
    }
    ~Pool() { (void)hipFree(base_); }
    // first fit, 256-byte granules; returns nullptr when nothing fits
    void* base() const { return base_; }
    void* alloc(uint64_t bytes)
    {
        bytes = (bytes + 255) & ~255ull;
        if (bytes == 0) bytes = 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const uint64_t off = it->first, left = it->second - bytes;
                free_.erase(it);
                if (left) free_[off + bytes] = left;
                used_ += bytes;
                return static_cast<char*>(base_) + off;
            }
        }
        return nullptr;
    }
    void release(void* p, uint64_t bytes)
    {
        bytes = (bytes + 255) & ~255ull;
        if (bytes == 0) bytes = 256;
        uint64_t off = static_cast<char*>(p) - static_cast<char*>(base_);
        used_ -= bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
    uint64_t used() const { return used_; }
    uint64_t size() const { return size_; }

private:
    void* base_ = nullptr;
    uint64_t size_;
    uint64_t used_ = 0;
    std::map<uint64_t, uint64_t> free_;
};

// ---------------------------------------------------------------- jobs
struct Job {
    aeth_job spec;
    uint64_t id;
    double submit_t = 0;
    std::vector<std::pair<void*, uint64_t>> mem;
    // resolved pointers
    void* buf[4] = {nullptr, nullptr, nullptr, nullptr};
    void* part = nullptr;
    void* scratch = nullptr;
    uint32_t tiles = 0;
    uint64_t total_points = 0; // FFT
};

size_t esize(uint32_t dtype) { return dtype == AETH_F64 ? 8 : 4; }
uint32_t tiles_of(uint64_t n) { return n ? (uint32_t)((n + kTile - 1) / kTile) : 1u; }
uint64_t cap32(uint64_t n) { return (n + 31) / 32 * 32; }
uint64_t tree_pairs(uint32_t tiles)
{
    uint64_t t = tiles, s = 0;
    while (t > 1) {
        t = (t + kGroup - 1) / kGroup;
        s += t;
    }
    return s + 1;
}

// The results a job reports: number of (hi, comp) pairs.
int result_pairs(uint32_t kind) { return kind == AETH_JOB_EVAL ? 3 : kind == AETH_JOB_SAMPLE ? 2 : 1; }

void validate(const aeth_job& j)
{
    if (j.dtype != AETH_F64 && j.dtype != AETH_F32) fail("aeth_job", "dtype must be F64 or F32");
    if (j.kind == AETH_JOB_EVAL || j.kind == AETH_JOB_SAMPLE) {
        const uint64_t span = (j.kind == AETH_JOB_EVAL ? 3 : 1) * cap32(j.n);
        if (j.n == 0 || span > 0xFFFFFFFFull) fail("aeth_job", "sample count out of range");
    } else if (j.kind == AETH_JOB_FFT) {
        if (j.n < 1 || j.n > 64) fail("aeth_job", "FFT jobs carry 1..64 signals");
        for (uint64_t k = 0; k < j.n; ++k)
            if (j.fft_log2n[k] > 13) fail("aeth_job", "FFT lengths are 2^0..2^13");
    } else {
        fail("aeth_job", "unknown job kind");
    }
}

uint64_t fft_points(const aeth_job& j)
{
    uint64_t t = 0;
    for (uint64_t k = 0; k < j.n; ++k) t += 1ull << j.fft_log2n[k];
    return t;
}

// Allocate every buffer of a job from the pool; false (and nothing held) if it does not fit.
bool admit(Pool& pool, Job& J)
{
    const aeth_job& s = J.spec;
    const size_t es = esize(s.dtype);
    std::vector<uint64_t> sizes;
    if (s.kind == AETH_JOB_EVAL) {
        J.tiles = tiles_of(s.n);
        const uint64_t arr = 3 * cap32(s.n) * es;
        sizes = {arr, arr, arr, arr, 3ull * J.tiles * 2 * es, 3ull * tree_pairs(J.tiles) * 2 * es};
    } else if (s.kind == AETH_JOB_SAMPLE) {
        J.tiles = tiles_of(s.n);
        sizes = {cap32(s.n) * es, 2ull * J.tiles * 2 * es, 2ull * tree_pairs(J.tiles) * 2 * es};
    } else {
        J.total_points = fft_points(s);
        J.tiles = tiles_of(2 * J.total_points);
        const uint64_t buf = 2 * J.total_points * es;
        sizes = {buf, buf, (uint64_t)J.tiles * 2 * es, tree_pairs(J.tiles) * 2 * es};
    }
    J.mem.clear();
    for (uint64_t b : sizes) {
        void* p = pool.alloc(b);
        if (!p) {
            for (auto& m : J.mem) pool.release(m.first, m.second);
            J.mem.clear();
            return false;
        }
        J.mem.emplace_back(p, b);
    }
    if (s.kind == AETH_JOB_EVAL) {
        for (int k = 0; k < 4; ++k) J.buf[k] = J.mem[k].first;
        J.part = J.mem[4].first;
        J.scratch = J.mem[5].first;
    } else if (s.kind == AETH_JOB_SAMPLE) {
        J.buf[0] = J.mem[0].first;
        J.part = J.mem[1].first;
        J.scratch = J.mem[2].first;
    } else {
        J.buf[0] = J.mem[0].first;
        J.buf[1] = J.mem[1].first;
        J.part = J.mem[2].first;
        J.scratch = J.mem[3].first;
    }
    return true;
}

uint64_t job_need(const aeth_job& s)
{
    const size_t es = esize(s.dtype);
    if (s.kind == AETH_JOB_EVAL) return 12 * cap32(s.n) * es + 8 * tiles_of(s.n) * es + (1 << 16);
    if (s.kind == AETH_JOB_SAMPLE) return cap32(s.n) * es + 4 * tiles_of(s.n) * es + (1 << 16);
    return 4 * fft_points(s) * es + (1 << 16);
}

// ---------------------------------------------------------------- batches
// The staging area of a slot: a header and eight tables at fixed offsets, sized for cfg.max_batch_jobs
// (the same layout in pinned host memory and in device memory).
enum { kRegHeader = 0, kRegFill, kRegEval, kRegRed, kRegTree, kRegFd0, kRegFs0, kRegFd2, kRegFs2 };
struct Layout {
    uint64_t at[kRegions];
    uint64_t cap[kRegions]; // bytes
    uint64_t total;
};

Layout make_layout(uint32_t max_jobs)
{
    const uint64_t sigs = 64ull * max_jobs; // at most 64 signals a job, each at most one descriptor
    const uint64_t caps[kRegions] = {sizeof(BatchHeader),           3ull * max_jobs * sizeof(FillDesc),
                                     max_jobs * sizeof(EvalDesc),   2ull * max_jobs * sizeof(RedDesc),
                                     3ull * max_jobs * sizeof(TreeDesc), sigs * sizeof(fft::Desc),
                                     sigs * sizeof(fft::Sig),       sigs * sizeof(fft::Desc),
                                     sigs * sizeof(fft::Sig)};
    Layout L{};
    uint64_t o = 0;
    for (int r = 0; r < kRegions; ++r) {
        o = (o + 255) & ~uint64_t(255);
        L.at[r] = o;
        L.cap[r] = caps[r];
        o += caps[r];
    }
    L.total = (o + 255) & ~uint64_t(255);
    return L;
}

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    unsigned char* host_stage = nullptr;     // pinned, mapped
    unsigned char* host_stage_dev = nullptr; // its device address
    unsigned char* dev_stage = nullptr;
    uint64_t* region_offset = nullptr;       // device copy of the layout's offsets
    double* host_res = nullptr;
    double* dev_res = nullptr;
    hipGraphExec_t graph[4] = {};            // the batch as one captured graph per persistent-grid candidate
    hipEvent_t t0 = nullptr, t1 = nullptr;   // timing of a batch the tuner measures
    int cand = -1;                           // the candidate the in-flight batch was timed with (-1: untimed)
    double work = 0;                         // its algorithmic bytes
    std::vector<Job> jobs;
    bool busy = false;
    double launch_t = 0;
};

// The persistent grids' candidates of the graph form: blocks of the batch kernels per CU. Every candidate
// gives the same bits (grid-stride kernels over position-independent units); the tuner keeps the fastest.
constexpr uint32_t kUnitsPerCu[4] = {2, 4, 8, 16};
constexpr int kDefaultCand = 2; // 8 blocks a CU
constexpr int kTuneSamples = 8; // timed batches per candidate before the pick

// Algorithmic bytes of a job in the precision it executes in (eval: 3 loads + 1 store per (component,
// sample); sampling: values written; FFT: each point read and written once).
double job_work(const aeth_job& j)
{
    const double es = (double)esize(j.dtype);
    if (j.kind == AETH_JOB_EVAL) return 12.0 * (double)j.n * es;
    if (j.kind == AETH_JOB_SAMPLE) return (double)j.n * es;
    double pts = 0;
    for (uint64_t k = 0; k < j.n; ++k) pts += (double)(1ull << j.fft_log2n[k]);
    return 4.0 * pts * es;
}

template<class T>
void put(const Layout& L, unsigned char* base, int region, const std::vector<T>& v, BatchHeader& h)
{
    const uint64_t bytes = v.size() * sizeof(T);
    if (bytes > L.cap[region]) fail("server", "batch descriptors exceed the staging area");
    if (bytes) std::memcpy(base + L.at[region], v.data(), bytes);
    h.region_bytes[region] = bytes;
}

// Write the slot's batch into its staging area (descriptor tables and header); returns the header.
BatchHeader build_batch(Slot& S, const Layout& L, void* base)
{
    std::vector<FillDesc> fills;
    std::vector<EvalDesc> evals;
    std::vector<RedDesc> reds;
    std::vector<TreeDesc> trees;
    std::vector<uint32_t> l2[3];
    std::vector<uint64_t> off[3];
    uint64_t fill_units = 0, eval_units = 0, red_units = 0;
    // FFT buffers are addressed by complex-element offsets from the pool base (pool blocks are 256-byte aligned).
    const uintptr_t b0 = reinterpret_cast<uintptr_t>(base);
    for (size_t j = 0; j < S.jobs.size(); ++j) {
        Job& J = S.jobs[j];
        const aeth_job& s = J.spec;
        const size_t es = esize(s.dtype);
        double* res = S.dev_res + 8 * j;
        if (s.kind == AETH_JOB_EVAL) {
            const uint64_t cap = cap32(s.n);
            for (int k = 0; k < 3; ++k) {
                FillDesc f{};
                f.out = J.buf[k];
                f.n = s.n;
                f.pitch = cap;
                f.seed = s.seed + (uint64_t)k;
                f.p0 = -1.0;
                f.p1 = 1.0;
                f.ncomp = 3;
                f.dtype = s.dtype;
                f.kind = 0;
                f.unit_start = fill_units;
                fill_units += 3 * ((s.n + kFillPerBlock - 1) / kFillPerBlock);
                fills.push_back(f);
            }
            EvalDesc e{};
            e.a = J.buf[0];
            e.b = J.buf[1];
            e.c = J.buf[2];
            e.d = J.buf[3];
            e.part = J.part;
            e.pitch = cap;
            e.n = (uint32_t)s.n;
            e.tiles = J.tiles;
            e.dtype = s.dtype;
            e.unit_start = eval_units;
            eval_units += 3ull * J.tiles;
            evals.push_back(e);
            for (int k = 0; k < 3; ++k)
                trees.push_back(TreeDesc{static_cast<char*>(J.part) + (size_t)k * J.tiles * 2 * es,
                                         static_cast<char*>(J.scratch) + (size_t)k * tree_pairs(J.tiles) * 2 * es,
                                         res + 2 * k, J.tiles, s.dtype});
        } else if (s.kind == AETH_JOB_SAMPLE) {
            FillDesc f{};
            f.out = J.buf[0];
            f.n = s.n;
            f.pitch = s.n;
            f.seed = s.seed;
            f.ncomp = 1;
            f.dtype = s.dtype;
            f.kind = 1;
            f.unit_start = fill_units;
            fill_units += (s.n + kFillPerBlock - 1) / kFillPerBlock;
            fills.push_back(f);
            for (uint32_t q = 0; q < 2; ++q) {
                reds.push_back(RedDesc{J.buf[0], static_cast<char*>(J.part) + (size_t)q * J.tiles * 2 * es, red_units,
                                       (uint32_t)s.n, J.tiles, s.dtype, q});
                red_units += J.tiles;
                trees.push_back(TreeDesc{static_cast<char*>(J.part) + (size_t)q * J.tiles * 2 * es,
                                         static_cast<char*>(J.scratch) + (size_t)q * tree_pairs(J.tiles) * 2 * es,
                                         res + 2 * q, J.tiles, s.dtype});
            }
        } else {
            FillDesc f{};
            f.out = J.buf[0];
            f.n = 2 * J.total_points;
            f.pitch = f.n;
            f.seed = s.seed;
            f.ncomp = 1;
            f.dtype = s.dtype;
            f.kind = 1;
            f.unit_start = fill_units;
            fill_units += (f.n + kFillPerBlock - 1) / kFillPerBlock;
            fills.push_back(f);
            const int prec = s.dtype == AETH_F64 ? 0 : 2;
            const size_t ce = 2 * es; // bytes per complex element
            uint64_t pos = 0;
            for (uint64_t k = 0; k < s.n; ++k) {
                l2[prec].push_back(s.fft_log2n[k]);
                off[prec].push_back((reinterpret_cast<uintptr_t>(J.buf[0]) - b0) / ce + pos);
                pos += 1ull << s.fft_log2n[k];
            }
            reds.push_back(RedDesc{J.buf[1], J.part, red_units, (uint32_t)(2 * J.total_points), J.tiles, s.dtype, 0});
            red_units += J.tiles;
            trees.push_back(TreeDesc{J.part, J.scratch, res, J.tiles, s.dtype});
        }
    }
    // FFT tables: input offsets are relative to the pool base; outputs sit at the same offset in buf[1]
    std::vector<fft::Desc> fdesc[3];
    std::vector<fft::Sig> fsig[3];
    for (int prec : {0, 2}) {
        if (l2[prec].empty()) continue;
        std::vector<uint64_t> out_off(off[prec].size());
        size_t q = 0;
        for (size_t j = 0; j < S.jobs.size(); ++j) {
            const Job& J = S.jobs[j];
            if (J.spec.kind != AETH_JOB_FFT || (J.spec.dtype == AETH_F64 ? 0 : 2) != prec) continue;
            const size_t ce = 2 * esize(J.spec.dtype);
            const uint64_t delta =
                (reinterpret_cast<uintptr_t>(J.buf[1]) - reinterpret_cast<uintptr_t>(J.buf[0])) / ce;
            for (uint64_t k = 0; k < J.spec.n; ++k, ++q) out_off[q] = off[prec][q] + delta;
        }
        fft::build_tables(l2[prec].data(), off[prec].data(), out_off.data(), l2[prec].size(), fdesc[prec], fsig[prec]);
    }
    BatchHeader h{};
    h.fills = (uint32_t)fills.size();
    h.evals = (uint32_t)evals.size();
    h.reds = (uint32_t)reds.size();
    h.trees = (uint32_t)trees.size();
    h.fft_ndesc[0] = (uint32_t)fdesc[0].size();
    h.fft_ndesc[1] = (uint32_t)fdesc[2].size();
    h.jobs = (uint32_t)S.jobs.size();
    for (const Job& J : S.jobs) h.dtypes |= 1u << J.spec.dtype;
    h.fill_units = fill_units;
    h.eval_units = eval_units;
    h.red_units = red_units;
    put(L, S.host_stage, kRegFill, fills, h);
    put(L, S.host_stage, kRegEval, evals, h);
    put(L, S.host_stage, kRegRed, reds, h);
    put(L, S.host_stage, kRegTree, trees, h);
    put(L, S.host_stage, kRegFd0, fdesc[0], h);
    put(L, S.host_stage, kRegFs0, fsig[0], h);
    put(L, S.host_stage, kRegFd2, fdesc[2], h);
    put(L, S.host_stage, kRegFs2, fsig[2], h);
    h.region_bytes[kRegHeader] = sizeof(BatchHeader);
    std::memcpy(S.host_stage + L.at[kRegHeader], &h, sizeof h);
    return h;
}

template<class T>
const T* at(const Slot& S, const Layout& L, int region)
{
    return reinterpret_cast<const T*>(S.dev_stage + L.at[region]);
}

// The batch's kernels on the slot's stream. grid_exact: size each grid to the batch (direct launches);
// otherwise fixed persistent grids that read the counts on the device (the form captured into a graph).
struct PersistentGrids {
    int units = 0;    // fill / eval / reduce / tree / staging kernels
    int fft[2] = {0, 0};
};

template<int DT>
void enqueue_kernels(Slot& S, const Layout& L, void* base, const BatchHeader* h, const PersistentGrids& pg);

// `dtypes`: the dtype mask the batch kernels must cover (a direct launch takes its batch's; a captured
// graph the mask of every batch it will replay).
void enqueue_batch(Slot& S, const Layout& L, void* base, const BatchHeader* h, uint32_t max_jobs,
                   const PersistentGrids& pg, uint32_t dtypes)
{
    if (h) dtypes = h->dtypes;
    if (dtypes == (1u << AETH_F64)) enqueue_kernels<kDtF64>(S, L, base, h, pg);
    else if (dtypes == (1u << AETH_F32)) enqueue_kernels<kDtF32>(S, L, base, h, pg);
    else enqueue_kernels<kDtBoth>(S, L, base, h, pg);
    const size_t rb = 8 * sizeof(double) * (h ? std::max<size_t>(1, h->jobs) : max_jobs);
    check(hipMemcpyAsync(S.host_res, S.dev_res, rb, hipMemcpyDeviceToHost, S.stream), "hipMemcpyAsync(results)");
}

template<int DT>
void enqueue_kernels(Slot& S, const Layout& L, void* base, const BatchHeader* h, const PersistentGrids& pg)
{
    hipStream_t st = S.stream;
    const BatchHeader* hdr = at<BatchHeader>(S, L, kRegHeader);
    const int persist = pg.units;
    auto grid = [&](uint64_t units) { return h ? (uint32_t)std::max<uint64_t>(1, units) : (uint32_t)persist; };
    constexpr int kStageBlocks = 64; // a few MB at most cross PCIe; 64 blocks keep the reads in flight
    hipLaunchKernelGGL(stage_in_kernel, dim3(kStageBlocks), dim3(kThreads), 0, st, S.host_stage_dev, S.dev_stage,
                       S.region_offset);
    if (!h || h->fills)
        hipLaunchKernelGGL(fill_batch_kernel<DT>, dim3(grid(h ? h->fill_units : 0)), dim3(kThreads), 0, st, hdr,
                           at<FillDesc>(S, L, kRegFill));
    if (!h || h->evals)
        hipLaunchKernelGGL(eval_batch_kernel<DT>, dim3(grid(h ? h->eval_units : 0)), dim3(kThreads), 0, st, hdr,
                           at<EvalDesc>(S, L, kRegEval));
    for (int q = 0; q < 2; ++q) {
        const int prec = q == 0 ? 0 : 2;
        const auto* fd = at<fft::Desc>(S, L, q == 0 ? kRegFd0 : kRegFd2);
        const auto* fs = at<fft::Sig>(S, L, q == 0 ? kRegFs0 : kRegFs2);
        if (h) {
            if (!h->fft_ndesc[q]) continue;
            fft::BatchPlan p;
            p.prec = prec;
            p.descs = const_cast<fft::Desc*>(fd);
            p.sigs = const_cast<fft::Sig*>(fs);
            p.ndesc = h->fft_ndesc[q];
            fft::launch(p, base, base, 1, 0, st);
        } else {
            fft::launch_persistent(prec, pg.fft[q], fd, fs, &hdr->fft_ndesc[q], base, base, 1, 0, st);
        }
    }
    if (!h || h->reds)
        hipLaunchKernelGGL(reduce_batch_kernel<DT>, dim3(grid(h ? h->red_units : 0)), dim3(kThreads), 0, st, hdr,
                           at<RedDesc>(S, L, kRegRed));
    if (!h || h->trees)
        hipLaunchKernelGGL(tree_batch_kernel, dim3(h ? std::max<uint32_t>(1, h->trees) : (uint32_t)persist),
                           dim3(512), 0, st, hdr, at<TreeDesc>(S, L, kRegTree));
    check(hipGetLastError(), "serve batch launch");
}

// Capture the persistent form of a batch on the slot's stream into a graph (once per grid candidate).
hipGraphExec_t capture_batch_graph(Slot& S, const Layout& L, void* base, uint32_t max_jobs, uint32_t units_per_cu,
                                   uint32_t dtypes)
{
    // everything the launches look up is ready before the capture: twiddle tables, grid sizes
    fft::prepare(S.stream);
    PersistentGrids pg;
    pg.units = device_cus() * (int)units_per_cu;
    pg.fft[0] = fft::persistent_grid(0);
    pg.fft[1] = fft::persistent_grid(2);
    check(hipStreamSynchronize(S.stream), "hipStreamSynchronize");
    check(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipGraph_t g = nullptr;
    try {
        enqueue_batch(S, L, base, nullptr, max_jobs, pg, dtypes);
    } catch (...) {
        (void)hipStreamEndCapture(S.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    check(hipStreamEndCapture(S.stream, &g), "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    check(e, "hipGraphInstantiate");
    return exec;
}

// Launch every job of the slot as one batch on the slot's stream; in graph form, with the graph of grid
// candidate `cand`, timed between S.t0 and S.t1 when `timed`.
void launch_batch(Slot& S, const Layout& L, void* base, bool graph, int cand, bool timed)
{
    const BatchHeader h = build_batch(S, L, base);
    if (graph) {
        if (timed) check(hipEventRecord(S.t0, S.stream), "hipEventRecord");
        check(hipGraphLaunch(S.graph[cand], S.stream), "hipGraphLaunch");
        if (timed) check(hipEventRecord(S.t1, S.stream), "hipEventRecord");
    } else {
        enqueue_batch(S, L, base, &h, 0, PersistentGrids{}, 0);
    }
    check(hipEventRecord(S.done, S.stream), "hipEventRecord");
}

struct Server {
    aeth_server_config cfg;
    Layout layout{};
    bool graph = false;
    // online tuning of the graph form's persistent grids (AETH_SERVER_TUNE): timed batches rotate over the
    // candidates until each has kTuneSamples, then the lowest median time per byte is kept (under mu)
    bool tune = false;
    bool settled = false;
    int chosen = kDefaultCand;
    std::vector<float> samples[4];
    int dev = 0;
    std::unique_ptr<Pool> pool;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_in, cv_out;
    std::deque<Job> incoming;
    std::deque<Job> pending;  // scheduler-owned
    std::deque<aeth_job_result> done;
    std::atomic<uint64_t> next_id{1};
    std::atomic<uint64_t> outstanding{0};
    std::atomic<bool> stop{false};
    std::thread worker;
    aeth_server_stats stats{};
    std::string fatal;

    void record_timing(Slot& S)
    {
        if (S.cand < 0) return;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, S.t0, S.t1) == hipSuccess && S.work > 0) samples[S.cand].push_back(ms / (float)(S.work * 1e-9));
        S.cand = -1;
        for (const auto& v : samples)
            if ((int)v.size() < kTuneSamples) return;
        int best = kDefaultCand;
        float best_med = -1.0f;
        for (int c = 0; c < 4; ++c) {
            std::vector<float> v = samples[c];
            std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
            const float med = v[v.size() / 2];
            if (best_med < 0.0f || med < best_med) {
                best_med = med;
                best = c;
            }
        }
        std::lock_guard<std::mutex> lock(mu);
        chosen = best;
        settled = true;
    }

    // the grid candidate of the next batch: the chosen one, or while tuning the one with fewest samples
    int next_candidate(bool& timed) const
    {
        timed = false;
        if (!graph || !tune || settled) return chosen;
        int c = 0;
        for (int k = 1; k < 4; ++k)
            if (samples[k].size() < samples[c].size()) c = k;
        timed = true;
        return c;
    }

    void complete(Slot& S)
    {
        record_timing(S);
        const double t = now_s();
        std::vector<aeth_job_result> out;
        for (size_t j = 0; j < S.jobs.size(); ++j) {
            Job& J = S.jobs[j];
            aeth_job_result r{};
            r.id = J.id;
            r.kind = J.spec.kind;
            r.status = 0;
            const int used = 2 * result_pairs(J.spec.kind);
            for (int k = 0; k < 8; ++k) r.value[k] = k < used ? S.host_res[8 * j + k] : 0.0;
            r.submit_time = J.submit_t;
            r.done_time = t;
            out.push_back(r);
            for (auto& m : J.mem) pool->release(m.first, m.second);
        }
        {
            std::lock_guard<std::mutex> lock(mu);
            for (auto& r : out) done.push_back(r);
            stats.completed += out.size();
            stats.batches += 1;
            stats.pool_bytes_in_use = pool->used();
        }
        outstanding -= out.size();
        cv_out.notify_all();
        S.jobs.clear();
        S.busy = false;
    }

    void fail_job(const Job& J, int status)
    {
        aeth_job_result r{};
        r.id = J.id;
        r.kind = J.spec.kind;
        r.status = status;
        r.submit_time = J.submit_t;
        r.done_time = now_s();
        {
            std::lock_guard<std::mutex> lock(mu);
            done.push_back(r);
            stats.failed += 1;
        }
        outstanding -= 1;
        cv_out.notify_all();
    }

    void loop()
    {
        (void)hipSetDevice(dev);
        while (true) {
            {
                std::unique_lock<std::mutex> lock(mu);
                bool any_busy = false;
                for (auto& s : slots) any_busy = any_busy || s.busy;
                if (incoming.empty() && pending.empty() && !any_busy) {
                    if (stop) break;
                    cv_in.wait_for(lock, std::chrono::milliseconds(2));
                }
                while (!incoming.empty()) {
                    pending.push_back(std::move(incoming.front()));
                    incoming.pop_front();
                }
            }
            bool progressed = false;
            for (auto& S : slots) {
                if (S.busy && hipEventQuery(S.done) == hipSuccess) {
                    complete(S);
                    progressed = true;
                }
            }
            for (auto& S : slots) {
                if (S.busy || pending.empty()) continue;
                // admission in arrival order against the pool (the HBM budget)
                while (!pending.empty() && S.jobs.size() < cfg.max_batch_jobs) {
                    Job& J = pending.front();
                    if (job_need(J.spec) > pool->size()) {
                        fail_job(J, 2);
                        pending.pop_front();
                        continue;
                    }
                    if (!admit(*pool, J)) break;
                    S.jobs.push_back(std::move(J));
                    pending.pop_front();
                }
                if (S.jobs.empty()) break; // pool full: wait for completions
                try {
                    bool timed = false;
                    const int cand = next_candidate(timed);
                    S.work = 0;
                    for (const auto& J : S.jobs) S.work += job_work(J.spec);
                    launch_batch(S, layout, pool->base(), graph, cand, timed);
                    S.cand = timed ? cand : -1;
                } catch (const Failure&) {
                    fatal = last_error();
                    for (auto& J : S.jobs) {
                        for (auto& m : J.mem) pool->release(m.first, m.second);
                        fail_job(J, 1);
                    }
                    S.jobs.clear();
                    continue;
                }
                S.busy = true;
                progressed = true;
            }
            if (!progressed) std::this_thread::yield();
        }
    }
};

} // namespace serve
} // namespace aeth

using namespace aeth;
using namespace aeth::serve;

extern "C" int aeth_server_create(const aeth_server_config* cfg, void** server)
{
    return guarded([&] {
        if (!cfg || !server) fail("aeth_server_create", "NULL argument");
        *server = nullptr;
        auto S = std::make_unique<Server>();
        S->cfg = *cfg;
        if (S->cfg.streams < 1) S->cfg.streams = 1;
        if (S->cfg.streams > 4) S->cfg.streams = 4;
        if (S->cfg.max_batch_jobs < 1) S->cfg.max_batch_jobs = 256;
        if (S->cfg.max_batch_jobs > 4096) S->cfg.max_batch_jobs = 4096;
        check(hipGetDevice(&S->dev), "hipGetDevice");
        S->pool = std::make_unique<Pool>(cfg->hbm_budget_bytes);
        S->graph = (S->cfg.flags & AETH_SERVER_GRAPH_BATCHES) != 0;
        S->tune = S->graph && (S->cfg.flags & AETH_SERVER_TUNE) != 0;
        S->layout = make_layout(S->cfg.max_batch_jobs);
        const uint64_t stage = S->layout.total;
        S->slots.resize(S->cfg.streams);
        for (auto& sl : S->slots) {
            check(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
            check(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "hipEventCreateWithFlags");
            check(hipHostMalloc(reinterpret_cast<void**>(&sl.host_stage), stage, hipHostMallocMapped),
                  "hipHostMalloc(stage)", stage);
            check(hipHostGetDevicePointer(reinterpret_cast<void**>(&sl.host_stage_dev), sl.host_stage, 0),
                  "hipHostGetDevicePointer(stage)");
            check(hipMalloc(reinterpret_cast<void**>(&sl.dev_stage), stage), "hipMalloc(stage)", stage);
            check(hipMalloc(reinterpret_cast<void**>(&sl.region_offset), sizeof S->layout.at), "hipMalloc(layout)");
            check(hipMemcpy(sl.region_offset, S->layout.at, sizeof S->layout.at, hipMemcpyHostToDevice),
                  "hipMemcpy(layout)");
            const size_t rb = 8 * sizeof(double) * S->cfg.max_batch_jobs;
            check(hipHostMalloc(reinterpret_cast<void**>(&sl.host_res), rb, hipHostMallocDefault), "hipHostMalloc(res)");
            check(hipMalloc(reinterpret_cast<void**>(&sl.dev_res), rb), "hipMalloc(res)");
            // an FP32-compensated server runs every job in FP32; otherwise batches may mix both dtypes
            const uint32_t dtypes = S->cfg.fp32_compensated ? 1u << AETH_F32 : (1u << AETH_F64) | (1u << AETH_F32);
            if (S->tune) {
                check(hipEventCreate(&sl.t0), "hipEventCreate");
                check(hipEventCreate(&sl.t1), "hipEventCreate");
                for (int c = 0; c < 4; ++c)
                    sl.graph[c] = capture_batch_graph(sl, S->layout, S->pool->base(), S->cfg.max_batch_jobs,
                                                      kUnitsPerCu[c], dtypes);
            } else if (S->graph) {
                sl.graph[kDefaultCand] = capture_batch_graph(sl, S->layout, S->pool->base(), S->cfg.max_batch_jobs,
                                                             kUnitsPerCu[kDefaultCand], dtypes);
            }
        }
        Server* raw = S.get();
        S->worker = std::thread([raw] { raw->loop(); });
        *server = S.release();
    });
}

extern "C" int aeth_server_submit(void* server, const aeth_job* job, uint64_t* id)
{
    return guarded([&] {
        auto* S = static_cast<Server*>(server);
        if (!S || !job) fail("aeth_server_submit", "NULL argument");
        validate(*job);
        Job J;
        J.spec = *job;
        if (S->cfg.fp32_compensated) J.spec.dtype = AETH_F32;
        J.id = S->next_id++;
        J.submit_t = now_s();
        if (id) *id = J.id;
        S->outstanding += 1;
        {
            std::lock_guard<std::mutex> lock(S->mu);
            S->incoming.push_back(std::move(J));
            S->stats.submitted += 1;
        }
        S->cv_in.notify_one();
    });
}

extern "C" int aeth_server_poll(void* server, aeth_job_result* out, uint32_t max, uint32_t* count)
{
    return guarded([&] {
        auto* S = static_cast<Server*>(server);
        if (!S || !count) fail("aeth_server_poll", "NULL argument");
        std::lock_guard<std::mutex> lock(S->mu);
        uint32_t k = 0;
        while (k < max && !S->done.empty()) {
            out[k++] = S->done.front();
            S->done.pop_front();
        }
        *count = k;
    });
}

extern "C" int aeth_server_wait_idle(void* server)
{
    return guarded([&] {
        auto* S = static_cast<Server*>(server);
        std::unique_lock<std::mutex> lock(S->mu);
        S->cv_out.wait(lock, [&] { return S->outstanding.load() == 0; });
    });
}

extern "C" int aeth_server_get_stats(void* server, aeth_server_stats* st)
{
    return guarded([&] {
        auto* S = static_cast<Server*>(server);
        std::lock_guard<std::mutex> lock(S->mu);
        *st = S->stats;
        st->pool_bytes_budget = S->pool->size();
    });
}

extern "C" int aeth_server_tuning(void* server, uint32_t* blocks_per_cu, uint32_t* settled)
{
    return guarded([&] {
        auto* S = static_cast<Server*>(server);
        if (!S || !blocks_per_cu || !settled) fail("aeth_server_tuning", "NULL argument");
        std::lock_guard<std::mutex> lock(S->mu);
        *blocks_per_cu = kUnitsPerCu[S->chosen];
        *settled = S->settled ? 1u : 0u;
    });
}

extern "C" int aeth_server_destroy(void* server)
{
    return guarded([&] {
        auto* S = static_cast<Server*>(server);
        if (!S) return;
        S->stop = true;
        S->cv_in.notify_all();
        if (S->worker.joinable()) S->worker.join();
        for (auto& sl : S->slots) {
            (void)hipStreamSynchronize(sl.stream);
            (void)hipStreamDestroy(sl.stream);
            (void)hipEventDestroy(sl.done);
            (void)hipHostFree(sl.host_stage);
            (void)hipFree(sl.dev_stage);
            (void)hipHostFree(sl.host_res);
            (void)hipFree(sl.dev_res);
            (void)hipFree(sl.region_offset);
            for (auto& g : sl.graph)
                if (g) (void)hipGraphExecDestroy(g);
            if (sl.t0) (void)hipEventDestroy(sl.t0);
            if (sl.t1) (void)hipEventDestroy(sl.t1);
        }
        delete S;
    });
}

extern "C" int aeth_server_replay(void* server, const aeth_job* jobs, uint64_t count, aeth_job_result* results)
{
    return guarded([&] {
        auto* S = static_cast<Server*>(server);
        if (!S) fail("aeth_server_replay", "NULL server");
        std::vector<uint64_t> ids(count);
        const double t0 = now_s();
        for (uint64_t i = 0; i < count; ++i) {
            const double due = t0 + jobs[i].arrival;
            while (now_s() < due) {
                const double wait = due - now_s();
                if (wait > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(wait * 0.5e6)));
            }
            const int rc = aeth_server_submit(server, &jobs[i], &ids[i]);
            if (rc) throw Failure{rc};
        }
        if (aeth_server_wait_idle(server)) throw Failure{1};
        std::map<uint64_t, uint64_t> pos;
        for (uint64_t i = 0; i < count; ++i) pos[ids[i]] = i;
        std::vector<aeth_job_result> buf(4096);
        uint64_t got = 0;
        while (got < count) {
            uint32_t k = 0;
            if (aeth_server_poll(server, buf.data(), (uint32_t)buf.size(), &k)) throw Failure{1};
            if (k == 0) break;
            for (uint32_t q = 0; q < k; ++q) {
                auto it = pos.find(buf[q].id);
                if (it != pos.end()) {
                    results[it->second] = buf[q];
                    results[it->second].submit_time -= t0;
                    results[it->second].done_time -= t0;
                    ++got;
                }
            }
        }
        if (got != count) fail("aeth_server_replay", "missing results");
    });
}

// The plain per-call form: one job, its own allocations, synchronous -- how a
// caller of the reference runs a request (Array ctor allocates, fill, launch,
// download, free).
namespace {
int run_plain(const aeth_job* job, aeth_job_result* out, void* ws, uint64_t ws_bytes, void* stream)
{
    return guarded([&] {
        validate(*job);
        const aeth_job& s = *job;
        hipStream_t st = as_stream(stream);
        const size_t es = esize(s.dtype);
        std::vector<void*> allocs;
        uint64_t ws_used = 0;
        auto galloc = [&](size_t b) {
            void* p = nullptr;
            if (ws) { // carve from the caller's workspace (256-byte aligned pieces)
                const uint64_t need = (b + 255) / 256 * 256;
                if (ws_used + need > ws_bytes) fail("aeth_run_job_plain_ws", "workspace too small", ws_used + need);
                p = static_cast<char*>(ws) + ws_used;
                ws_used += need;
                return p;
            }
            check(hipMalloc(&p, b ? b : 16), "hipMalloc", b);
            allocs.push_back(p);
            return p;
        };
        const double t_submit = now_s();
        double res[8] = {0};
        try {
            void* sums = galloc(8 * es);
            auto widen = [&](int pairs) {
                std::vector<unsigned char> h(2 * pairs * es);
                check(hipMemcpyAsync(h.data(), sums, h.size(), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
                check(hipStreamSynchronize(st), "hipStreamSynchronize");
                for (int k = 0; k < 2 * pairs; ++k)
                    res[k] = es == 8 ? reinterpret_cast<double*>(h.data())[k] : (double)reinterpret_cast<float*>(h.data())[k];
            };
            const int sum_mode = s.dtype == AETH_F64 ? AETH_SUM_NEUMAIER : AETH_SUM_FLOATFLOAT;
            if (s.kind == AETH_JOB_EVAL) {
                const uint64_t cap = cap32(s.n);
                void* a = galloc(3 * cap * es);
                void* b = galloc(3 * cap * es);
                void* c = galloc(3 * cap * es);
                void* d = galloc(3 * cap * es);
                void* in[3] = {a, b, c};
                for (int k = 0; k < 3; ++k)
                    if (aeth_array_fill(AETH_DIST_UNIFORM, (int)s.dtype, s.seed + k, -1.0, 1.0, 3, s.n, cap, in[k], stream))
                        throw Failure{1};
                if (aeth_eval_cfg0((int)s.dtype, a, b, c, d, 3, s.n, cap, sum_mode, sums, stream)) throw Failure{1};
                widen(3);
            } else if (s.kind == AETH_JOB_SAMPLE) {
                void* x = galloc(cap32(s.n) * es);
                if (aeth_rng_fill(s.dtype == AETH_F64 ? AETH_RNG_NORMAL_F64 : AETH_RNG_NORMAL_F32, s.seed, 0, s.n, x,
                                  stream))
                    throw Failure{1};
                if (aeth_reduce((int)s.dtype, sum_mode, 0, x, 1, s.n, s.n, sums, stream)) throw Failure{1};
                if (aeth_reduce((int)s.dtype, sum_mode, 1, x, 1, s.n, s.n, static_cast<char*>(sums) + 2 * es, stream))
                    throw Failure{1};
                widen(2);
            } else {
                const uint64_t total = fft_points(s);
                void* x = galloc(2 * total * es);
                void* y = galloc(2 * total * es);
                if (aeth_rng_fill(s.dtype == AETH_F64 ? AETH_RNG_NORMAL_F64 : AETH_RNG_NORMAL_F32, s.seed, 0,
                                  2 * total, x, stream))
                    throw Failure{1};
                std::vector<uint64_t> len(s.n), offs(s.n);
                uint64_t pos = 0;
                for (uint64_t k = 0; k < s.n; ++k) {
                    len[k] = 1ull << s.fft_log2n[k];
                    offs[k] = pos;
                    pos += len[k];
                }
                void* plan = nullptr;
                if (aeth_fft_batch_plan(s.dtype == AETH_F64 ? AETH_FFT_C128 : AETH_FFT_C64_FF, s.n, len.data(),
                                        offs.data(), offs.data(), &plan))
                    throw Failure{1};
                const int rc = aeth_fft_batch_exec(plan, x, y, 1, AETH_FFT_BACKWARD, stream);
                check(hipStreamSynchronize(st), "hipStreamSynchronize");
                aeth_fft_batch_destroy(plan);
                if (rc) throw Failure{rc};
                if (aeth_reduce((int)s.dtype, sum_mode, 0, y, 1, 2 * total, 2 * total, sums, stream)) throw Failure{1};
                widen(1);
            }
        } catch (...) {
            for (void* p : allocs) (void)hipFree(p);
            throw;
        }
        for (void* p : allocs) (void)hipFree(p);
        if (out) {
            std::memset(out, 0, sizeof *out);
            out->kind = s.kind;
            for (int k = 0; k < 8; ++k) out->value[k] = res[k];
            out->submit_time = t_submit;
            out->done_time = now_s();
        }
    });
}
} // namespace

extern "C" int aeth_run_job_plain(const aeth_job* job, aeth_job_result* out, void* stream)
{
    return run_plain(job, out, nullptr, 0, stream);
}

extern "C" int aeth_run_job_plain_ws(const aeth_job* job, aeth_job_result* out, void* workspace,
                                     uint64_t workspace_bytes, void* stream)
{
    if (!workspace) return guarded([] { fail("aeth_run_job_plain_ws", "workspace is NULL"); });
    return run_plain(job, out, workspace, workspace_bytes, stream);
}
