// seam.hip -- the `aether-memory/1` C seam (dsc/src/seam.h:22-51) over HIP.
//
// Same table shape, field order and return conventions as the reference's
// CUDA plugin (dsc/src/plugin.cpp:20-156), exported as aether_memory_v1().
// Memory kinds CPU / DEVICE / PINNED; device chunks are stream-ordered
// (hipMallocAsync from the device's default pool); a chunk released while its
// stream is being captured is queued and freed by drain() after the capture
// ends (chunk/StreamOrdered.h:160-286). fill and device-to-device copy are
// hand-written kernels (16 KiB units of 16-byte lanes), legal inside a graph capture
// and free of memset nodes; other copies go through hipMemcpyAsync.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/aether_hip.h"
#include "common.hpp"
#include "units.hpp"

namespace aeth {
namespace fft {
int seam_plan(int dev, int kind, int is_double, int norm, int forward, int ndim, const int64_t* shape, int naxes,
              const int32_t* axes, void* stream, void** plan);
int seam_exec(void* plan, const void* in, void* out, void* stream);
void seam_destroy(void* plan);
} // namespace fft

namespace {

struct Chunk {
    int kind;
    int dev;
    void* ptr;
    size_t bytes;
    hipStream_t stream;
};

std::mutex g_deferred_mu;
std::vector<Chunk*> g_deferred;

struct OnDevice {
    int prev = 0;
    bool switched = false;
    explicit OnDevice(int dev)
    {
        check(hipGetDevice(&prev), "hipGetDevice");
        switched = prev != dev;
        if (switched) check(hipSetDevice(dev), "hipSetDevice");
    }
    ~OnDevice()
    {
        if (switched) (void)hipSetDevice(prev);
    }
};

// Never throws: it runs from chunk_release, which returns nothing.
void free_now(Chunk* c)
{
    RelaxedCapture relaxed;
    try {
        if (c->kind == AETHER_MEM_DEVICE) {
            OnDevice d(c->dev);
            (void)hipFreeAsync(c->ptr, c->stream);
        } else if (c->kind == AETHER_MEM_PINNED) {
            (void)hipHostFree(c->ptr);
        } else {
            std::free(c->ptr);
        }
    } catch (...) {
    }
    (void)hipGetLastError();
    delete c;
}

// A release a capture makes illegal now: a device free is ordered on its own stream, a pinned free
// synchronises the device (the legacy stream).
bool blocked(const Chunk* c)
{
    if (c->kind == AETHER_MEM_CPU) return false;
    return capturing(nullptr) || (c->kind == AETHER_MEM_DEVICE && capturing(c->stream));
}

constexpr int kThreads = 256;

// Fill `bytes` with a repeating `width`-byte pattern (width divides 16 or is
// 1, 2, 4, 8, 16): the 16-byte lane pattern is pre-expanded on the host.
// Byte k of the region takes pat16[k % 16]. The `head` bytes before the first 16-byte boundary and the
// tail after the last one are stored bytewise; the aligned body takes 16-byte stores of body16, which is
// pat16 rotated by `head` (so body byte j is pat16[(head + j) % 16]).
//
// Bulk fills and device-to-device copies move 16 KiB units (four 16-byte vectors a lane, a copy's loads
// all issued before its stores), one workgroup a unit, in the XCD-aware unit order of the evaluation
// kernels; the vectors past the last whole unit go grid-stride. NT selects the streaming cache policy, taken
// once the bytes moved are far beyond the 256 MB last-level cache. Measured on MI355X over 3 GiB: a copy at
// one vector a lane grid-stride moved 4.77 TB/s.
constexpr uint32_t kUnitVecs = kThreads * 4;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template<bool NT>
__device__ __forceinline__ u32x4 ld16(const u32x4* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template<bool NT>
__device__ __forceinline__ void st16(u32x4* p, u32x4 x)
{
    if constexpr (NT) __builtin_nontemporal_store(x, p);
    else *p = x;
}

template<bool NT>
__global__ __launch_bounds__(kThreads) void fill_kernel(unsigned char* __restrict__ p, size_t bytes, uint4 pat16,
                                                        uint4 body16, uint32_t head, uint32_t units)
{
    constexpr uint32_t K = units::xcd_chunk(kUnitVecs * 16);
    u32x4* body = reinterpret_cast<u32x4*>(p + head);
    const size_t vecs = (bytes - head) / 16;
    const u32x4 b = {body16.x, body16.y, body16.z, body16.w};
    for (uint32_t v = blockIdx.x; v < units; v += gridDim.x) {
        const size_t base = (size_t)units::xcd_unit(v, units, K) * kUnitVecs + threadIdx.x;
#pragma unroll
        for (int r = 0; r < 4; ++r) st16<NT>(body + base + r * kThreads, b);
    }
    for (size_t i = (size_t)units * kUnitVecs + (size_t)blockIdx.x * kThreads + threadIdx.x; i < vecs;
         i += (size_t)gridDim.x * kThreads)
        body[i] = b;
    const size_t tail = bytes - head - vecs * 16;
    if (blockIdx.x == 0) {
        if (threadIdx.x < head) p[threadIdx.x] = reinterpret_cast<const unsigned char*>(&pat16)[threadIdx.x];
        if (threadIdx.x < tail)
            reinterpret_cast<unsigned char*>(body + vecs)[threadIdx.x] =
                reinterpret_cast<const unsigned char*>(&body16)[threadIdx.x];
    }
}

template<bool NT>
__global__ __launch_bounds__(kThreads) void copy16_kernel(u32x4* __restrict__ dst, const u32x4* __restrict__ src,
                                                          size_t vecs, uint32_t units, uint32_t tail)
{
    constexpr uint32_t K = units::xcd_chunk(kUnitVecs * 16);
    for (uint32_t v = blockIdx.x; v < units; v += gridDim.x) {
        const size_t base = (size_t)units::xcd_unit(v, units, K) * kUnitVecs + threadIdx.x;
        u32x4 x[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) x[r] = ld16<NT>(src + base + r * kThreads);
#pragma unroll
        for (int r = 0; r < 4; ++r) st16<NT>(dst + base + r * kThreads, x[r]);
    }
    for (size_t i = (size_t)units * kUnitVecs + (size_t)blockIdx.x * kThreads + threadIdx.x; i < vecs;
         i += (size_t)gridDim.x * kThreads)
        dst[i] = src[i];
    if (blockIdx.x == 0 && threadIdx.x < tail)
        reinterpret_cast<unsigned char*>(dst + vecs)[threadIdx.x] =
            reinterpret_cast<const unsigned char*>(src + vecs)[threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void copy_bytes_kernel(unsigned char* __restrict__ dst,
                                                              const unsigned char* __restrict__ src, size_t bytes)
{
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < bytes; i += (size_t)gridDim.x * kThreads)
        dst[i] = src[i];
}

int grid_for(size_t work)
{
    const size_t blocks = (work + kThreads - 1) / kThreads;
    return (int)std::max<size_t>(1, std::min<size_t>(blocks, (size_t)device_cus() * 32));
}

// Whole 16 KiB units of `vecs` 16-byte vectors (at most 2^31 - 1 of them, 32 TiB, so a unit index plus the
// grid never wraps; vectors beyond go grid-stride) and the grid that runs them one workgroup a unit, or
// grid-stride over the vectors when there is no whole unit.
uint32_t units_of(size_t vecs) { return (uint32_t)std::min<size_t>(vecs / kUnitVecs, 0x7FFFFFFFull); }

int unit_grid(size_t vecs)
{
    const uint32_t u = units_of(vecs);
    return u ? (int)u : grid_for(vecs + 1);
}

// Streaming policy once the bytes moved dwarf the last-level cache.
bool streaming_bytes(size_t moved) { return moved >= (512ull << 20); }

int s_probe(int* count) { return aeth_probe(count); }

int s_runtime_version(int* v)
{
    return guarded([&] { check(hipRuntimeGetVersion(v), "hipRuntimeGetVersion"); });
}

int s_current_device(int* d)
{
    return guarded([&] { check(hipGetDevice(d), "hipGetDevice"); });
}

const char* s_last_error() { return last_error(); }

int s_chunk_alloc(int kind, int dev, size_t bytes, void* stream, void** chunk, void** ptr)
{
    return guarded([&] {
        if (!chunk || !ptr) fail("chunk_alloc", "NULL out-parameter");
        auto* c = new Chunk{kind, dev, nullptr, bytes, as_stream(stream)};
        try {
            if (kind == AETHER_MEM_DEVICE) {
                OnDevice d(dev);
                if (bytes) check(hipMallocAsync(&c->ptr, bytes, c->stream), "hipMallocAsync", bytes);
            } else if (kind == AETHER_MEM_PINNED) {
                if (bytes) check(hipHostMalloc(&c->ptr, bytes, hipHostMallocDefault), "hipHostMalloc", bytes);
            } else if (kind == AETHER_MEM_CPU) {
                c->ptr = bytes ? std::aligned_alloc(256, (bytes + 255) / 256 * 256) : nullptr;
                if (bytes && !c->ptr) fail("chunk_alloc", "host allocation failed", bytes);
            } else {
                fail("chunk_alloc", "unknown memory kind");
            }
        } catch (...) {
            delete c;
            throw;
        }
        *chunk = c;
        *ptr = c->ptr;
    });
}

// Release every deferred chunk whose capture has ended; those still blocked stay queued.
void s_drain()
{
    drain_device_frees(); // buffers of FFT plans destroyed during a capture
    std::vector<Chunk*> keep, now;
    {
        std::lock_guard<std::mutex> lock(g_deferred_mu);
        if (g_deferred.empty()) return;
        for (Chunk* c : g_deferred) (blocked(c) ? keep : now).push_back(c);
        g_deferred.swap(keep);
    }
    for (Chunk* c : now) free_now(c);
}

void s_chunk_release(void* chunk)
{
    auto* c = static_cast<Chunk*>(chunk);
    if (!c) return;
    s_drain();
    if (blocked(c)) {
        std::lock_guard<std::mutex> lock(g_deferred_mu);
        g_deferred.push_back(c);
        return;
    }
    free_now(c);
}

int s_fill(int kind, int dev, void* ptr, size_t bytes, const void* value, size_t width, void* stream)
{
    return guarded([&] {
        if (bytes == 0) return;
        if (!ptr || !value) fail("fill", "NULL pointer");
        if (width == 0 || width > 16 || (16 % width) != 0 || bytes % width != 0)
            fail("fill", "width must divide 16 and the byte count", bytes);
        unsigned char pat[16];
        for (size_t i = 0; i < 16; ++i) pat[i] = static_cast<const unsigned char*>(value)[i % width];
        if (kind == AETHER_MEM_DEVICE) {
            OnDevice d(dev);
            if (reinterpret_cast<uintptr_t>(ptr) % width != 0) fail("fill", "device pointer is not aligned to the width", bytes);
            const uint32_t head =
                (uint32_t)std::min<size_t>(bytes, (16 - reinterpret_cast<uintptr_t>(ptr) % 16) % 16);
            unsigned char rot[16];
            for (size_t i = 0; i < 16; ++i) rot[i] = pat[(head + i) % 16];
            uint4 p16, b16;
            std::memcpy(&p16, pat, 16);
            std::memcpy(&b16, rot, 16);
            const size_t vecs = (bytes - head) / 16;
            if (streaming_bytes(bytes))
                hipLaunchKernelGGL(fill_kernel<true>, dim3(unit_grid(vecs)), dim3(kThreads), 0, as_stream(stream),
                                   static_cast<unsigned char*>(ptr), bytes, p16, b16, head, units_of(vecs));
            else
                hipLaunchKernelGGL(fill_kernel<false>, dim3(unit_grid(vecs)), dim3(kThreads), 0, as_stream(stream),
                                   static_cast<unsigned char*>(ptr), bytes, p16, b16, head, units_of(vecs));
            check(hipGetLastError(), "fill_kernel launch", bytes);
        } else {
            // host memory: ordered after the stream's queued work
            check(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize", bytes);
            unsigned char* p = static_cast<unsigned char*>(ptr);
            for (size_t i = 0; i < bytes; ++i) p[i] = pat[i % 16];
        }
    });
}

int s_copy(int dst_kind, int dst_dev, void* dst, int src_kind, int src_dev, const void* src, size_t bytes,
           void* stream)
{
    return guarded([&] {
        if (bytes == 0) return;
        hipStream_t st = as_stream(stream);
        if (dst_kind == AETHER_MEM_DEVICE && src_kind == AETHER_MEM_DEVICE && dst_dev == src_dev) {
            OnDevice d(dst_dev);
            if (aligned16(dst) && aligned16(src)) {
                const size_t vecs = bytes / 16;
                u32x4* d16 = static_cast<u32x4*>(dst);
                const u32x4* s16 = static_cast<const u32x4*>(src);
                const uint32_t tail = (uint32_t)(bytes - vecs * 16);
                if (streaming_bytes(2 * bytes))
                    hipLaunchKernelGGL(copy16_kernel<true>, dim3(unit_grid(vecs)), dim3(kThreads), 0, st, d16, s16,
                                       vecs, units_of(vecs), tail);
                else
                    hipLaunchKernelGGL(copy16_kernel<false>, dim3(unit_grid(vecs)), dim3(kThreads), 0, st, d16, s16,
                                       vecs, units_of(vecs), tail);
            } else {
                hipLaunchKernelGGL(copy_bytes_kernel, dim3(grid_for(bytes)), dim3(kThreads), 0, st,
                                   static_cast<unsigned char*>(dst), static_cast<const unsigned char*>(src), bytes);
            }
            check(hipGetLastError(), "copy_kernel launch", bytes);
            return;
        }
        if (dst_kind == AETHER_MEM_DEVICE && src_kind == AETHER_MEM_DEVICE) {
            check(hipMemcpyPeerAsync(dst, dst_dev, src, src_dev, bytes, st), "hipMemcpyPeerAsync", bytes);
            return;
        }
        const int dev = dst_kind == AETHER_MEM_DEVICE ? dst_dev : src_dev;
        OnDevice d(dev);
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, st), "hipMemcpyAsync", bytes);
        // pageable host memory: the copy must be complete before the caller may touch it
        if (dst_kind == AETHER_MEM_CPU || src_kind == AETHER_MEM_CPU)
            check(hipStreamSynchronize(st), "hipStreamSynchronize", bytes);
    });
}

int s_stream_wait(int dev, void* producer, void* consumer)
{
    return guarded([&] {
        OnDevice d(dev);
        hipEvent_t ev;
        check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreateWithFlags");
        hipError_t e = hipEventRecord(ev, as_stream(producer));
        if (e == hipSuccess) e = hipStreamWaitEvent(as_stream(consumer), ev, 0);
        (void)hipEventDestroy(ev);
        check(e, "stream_wait");
    });
}

int s_stream_sync(int dev, void* stream)
{
    return guarded([&] {
        OnDevice d(dev);
        check(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
    });
}

int s_pointer_device(const void* p, int* dev)
{
    return guarded([&] {
        hipPointerAttribute_t a;
        check(hipPointerGetAttributes(&a, p), "hipPointerGetAttributes");
        if (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged) fail("pointer_device", "not a device pointer");
        *dev = a.device;
    });
}

int s_fft_plan(int dev, int kind, int is_double, int norm, int forward, int ndim, const int64_t* shape, int naxes,
               const int32_t* axes, void* stream, void** plan)
{
    return fft::seam_plan(dev, kind, is_double, norm, forward, ndim, shape, naxes, axes, stream, plan);
}

int s_fft_exec(void* plan, const void* in, void* out, void* stream) { return fft::seam_exec(plan, in, out, stream); }

void s_fft_destroy(void* plan) { fft::seam_destroy(plan); }

const AetherMemoryV1 kTable = {AETHER_MEMORY_ABI, s_probe,        s_runtime_version, s_current_device, s_last_error,
                               s_chunk_alloc,     s_chunk_release, s_drain,           s_fill,           s_copy,
                               s_stream_wait,     s_stream_sync,  s_pointer_device,  s_fft_plan,       s_fft_exec,
                               s_fft_destroy};

} // namespace

namespace {
std::mutex g_later_mu;
std::vector<void*> g_later; // device allocations whose free a capture refused or would have broken

bool capture_error(hipError_t e)
{
    return e == hipErrorStreamCaptureUnsupported || e == hipErrorStreamCaptureImplicit ||
           e == hipErrorStreamCaptureInvalidated || e == hipErrorStreamCaptureWrongThread ||
           e == hipErrorCapturedEvent;
}
} // namespace

void free_device_later(std::vector<void*> ptrs)
{
    std::lock_guard<std::mutex> lock(g_later_mu);
    for (void* p : g_later) ptrs.push_back(p);
    g_later.clear();
    RelaxedCapture relaxed;
    // the legacy stream captured (or its query refused): a free, which synchronises the device, must wait
    if (capturing(nullptr)) {
        for (void* p : ptrs)
            if (p) g_later.push_back(p);
        return;
    }
    for (void* p : ptrs) {
        if (!p) continue;
        const hipError_t e = hipFree(p);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            if (capture_error(e)) g_later.push_back(p);
        }
    }
}

void drain_device_frees()
{
    bool any = false;
    {
        std::lock_guard<std::mutex> lock(g_later_mu);
        any = !g_later.empty();
    }
    if (any) free_device_later({});
}

} // namespace aeth

extern "C" __attribute__((visibility("default"))) const AetherMemoryV1* aether_memory_v1(void) { return &aeth::kTable; }

// ---------------------------------------------------------------- residency support
namespace {
std::mutex g_peer_mu;
std::vector<std::pair<int, int>> g_peer_enabled; // (from, to) pairs with peer access turned on

void enable_peer(int from, int to)
{
    for (const auto& e : g_peer_enabled)
        if (e.first == from && e.second == to) return;
    aeth::OnDevice on(from);
    const hipError_t e = hipDeviceEnablePeerAccess(to, 0);
    if (e != hipSuccess && e != hipErrorPeerAccessAlreadyEnabled) aeth::check(e, "hipDeviceEnablePeerAccess");
    (void)hipGetLastError();
    g_peer_enabled.emplace_back(from, to);
}
} // namespace

extern "C" int aeth_peer_route(int kind_a, int dev_a, int kind_b, int dev_b, int force_staging, int* route)
{
    return aeth::guarded([&] {
        if (!route) aeth::fail("aeth_peer_route", "route is NULL");
        if (kind_a != AETHER_MEM_DEVICE || kind_b != AETHER_MEM_DEVICE) {
            *route = AETH_ROUTE_HOST;
            return;
        }
        if (force_staging) {
            *route = AETH_ROUTE_STAGED;
            return;
        }
        if (dev_a == dev_b) {
            *route = AETH_ROUTE_SAME;
            return;
        }
        int ab = 0, ba = 0;
        aeth::check(hipDeviceCanAccessPeer(&ab, dev_a, dev_b), "hipDeviceCanAccessPeer");
        aeth::check(hipDeviceCanAccessPeer(&ba, dev_b, dev_a), "hipDeviceCanAccessPeer");
        if (!(ab && ba)) {
            *route = AETH_ROUTE_STAGED;
            return;
        }
        std::lock_guard<std::mutex> lock(g_peer_mu);
        enable_peer(dev_a, dev_b);
        enable_peer(dev_b, dev_a);
        *route = AETH_ROUTE_P2P;
    });
}

extern "C" int aeth_event_record(void* stream, void** event)
{
    return aeth::guarded([&] {
        if (!event) aeth::fail("aeth_event_record", "event is NULL");
        hipEvent_t ev = nullptr;
        aeth::check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreateWithFlags");
        const hipError_t e = hipEventRecord(ev, aeth::as_stream(stream));
        if (e != hipSuccess) {
            (void)hipEventDestroy(ev);
            aeth::check(e, "hipEventRecord");
        }
        *event = ev;
    });
}

extern "C" int aeth_event_synchronize(void* event)
{
    return aeth::guarded([&] {
        if (event) aeth::check(hipEventSynchronize(static_cast<hipEvent_t>(event)), "hipEventSynchronize");
    });
}

extern "C" int aeth_event_destroy(void* event)
{
    return aeth::guarded([&] {
        if (event) aeth::check(hipEventDestroy(static_cast<hipEvent_t>(event)), "hipEventDestroy");
    });
}
