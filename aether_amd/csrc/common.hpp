// common.hpp -- error convention and launch helpers shared by the C-ABI entry points.
//
// Every entry point returns 0 on success, 1 on failure (message in the
// per-thread last error), 2 when the request is valid but not supported by
// this build -- the convention of the reference's C seam (dsc/src/seam.h:6-7,
// dsc/src/plugin.cpp:22-39). Messages follow aether::err::fail's shape
// "<op> failed on <where>, size=<bytes> bytes (<detail>)" (err/Error.h:63).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>
#include <vector>

namespace aeth {

void set_error(const std::string& msg);
const char* last_error();

struct Failure {
    int code;
};

inline void fail(const char* op, const std::string& detail, size_t bytes = 0, int code = 1)
{
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s failed on device, size=%zu bytes (%s)", op, bytes, detail.c_str());
    set_error(buf);
    throw Failure{code};
}

inline void check(hipError_t e, const char* op, size_t bytes = 0)
{
    if (e != hipSuccess) fail(op, hipGetErrorString(e), bytes);
}

// Run `body`; translate failures into the 0/1/2 return convention.
template<class F>
int guarded(F&& body)
{
    try {
        body();
        return 0;
    } catch (const Failure& f) {
        return f.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("unknown failure");
        return 1;
    }
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Number of compute units of the current device (cached per device).
int device_cus();

// Stream-ordered scratch: allocated from the device's default memory pool on
// `stream` and freed on it when the guard dies. Legal inside a stream capture
// (the graph owns the allocation). On this ROCm the pair holds the host until
// the stream has drained, so calls that run outside a capture take a
// StreamLease instead.
class StreamScratch {
public:
    StreamScratch(size_t bytes, hipStream_t stream) : stream_(stream)
    {
        if (bytes) check(hipMallocAsync(&ptr_, bytes, stream), "hipMallocAsync(scratch)", bytes);
    }
    ~StreamScratch()
    {
        if (ptr_) (void)hipFreeAsync(ptr_, stream_);
    }
    StreamScratch(const StreamScratch&) = delete;
    StreamScratch& operator=(const StreamScratch&) = delete;
    void* get() const { return ptr_; }

private:
    void* ptr_ = nullptr;
    hipStream_t stream_;
};

// Scratch kept per (device, stream): one growable device buffer per stream handle, reused by every call
// on that stream. Calls on one stream run in stream order, so one call's kernels never meet another's in
// the buffer; the lease holds the slot's lock while the call enqueues its kernels, so two host threads
// issuing to one stream take turns. A buffer that must grow is released and replaced in stream order.
// Measured on MI355X: a stream-ordered allocation and free per call held the host for ~1.9 ms of a 2.1 ms
// fused evaluation and left the GPU idle 7-16 us before the next call's kernel; a lease costs neither.
// Inside a stream capture, or when this thread already holds the stream's slot, the lease falls back to a
// StreamScratch. A stream destroyed with work still queued must not have its handle reused while that work
// runs (the slot is keyed by the handle's value). Each lease records an event on its stream as it ends (the
// buffer's last use); once more than a few slots hold buffers, a lease on a stream new to the library
// releases the buffers of idle slots (unlocked, their last use complete), so a program that creates and
// destroys streams keeps a bounded amount of scratch.
class StreamLease {
public:
    StreamLease(size_t bytes, hipStream_t stream);
    ~StreamLease();
    StreamLease(const StreamLease&) = delete;
    StreamLease& operator=(const StreamLease&) = delete;
    void* get() const { return ptr_; }

private:
    void* ptr_ = nullptr;
    void* slot_ = nullptr;
    hipStream_t stream_ = nullptr;
    StreamScratch* fallback_ = nullptr;
};

// Queries and releases made while another thread or stream captures in global mode: the thread's
// capture mode is relaxed for the call's duration, so they are legal and do not invalidate the capture.
struct RelaxedCapture {
    hipStreamCaptureMode prev = hipStreamCaptureModeRelaxed;
    RelaxedCapture() { (void)hipThreadExchangeStreamCaptureMode(&prev); }
    ~RelaxedCapture() { (void)hipThreadExchangeStreamCaptureMode(&prev); }
};

// Whether `s` is being captured, or a capture elsewhere makes work on it illegal (a failed query, such
// as the implicit capture of the legacy stream under a blocking capture, answers yes).
inline bool capturing(hipStream_t s)
{
    RelaxedCapture relaxed;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    return st != hipStreamCaptureStatusNone;
}

// Frees device allocations (hipMalloc) in a way that is safe while a stream capture is in progress:
// with the thread's capture mode relaxed, and deferred -- kept on a process-wide list -- while the legacy
// stream is captured or a free is refused for a capture. Every call (and drain_device_frees, which the
// seam's drain() calls) first retries what earlier calls deferred. Never throws.
void free_device_later(std::vector<void*> ptrs);
void drain_device_frees();

} // namespace aeth
