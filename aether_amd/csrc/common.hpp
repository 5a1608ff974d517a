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
// `stream` and freed on it when the guard dies. No host synchronisation.
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
