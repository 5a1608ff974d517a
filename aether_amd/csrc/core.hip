// core.hip -- error state, device queries and library identity.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/aether_hip.h"
#include "common.hpp"

namespace aeth {

namespace {
thread_local std::string g_last_error;
std::mutex g_cu_mutex;
std::vector<int> g_cus;
} // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

const char* last_error() { return g_last_error.c_str(); }

int device_cus()
{
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_cu_mutex);
    if ((int)g_cus.size() <= dev) g_cus.resize(dev + 1, 0);
    if (g_cus[dev] == 0) {
        int cus = 0;
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        g_cus[dev] = cus > 0 ? cus : 1;
    }
    return g_cus[dev];
}

namespace {

struct LeaseSlot {
    std::mutex mu;
    void* ptr = nullptr;
    size_t bytes = 0;
    hipEvent_t last = nullptr; // recorded on the stream as each lease ends: the buffer's last use
    bool kept = false;         // a record of `last` failed: it no longer marks the last use, never release
};

std::mutex g_slots_mu;
// never erased (leases point in); only the buffers of idle slots are released
std::map<std::pair<int, hipStream_t>, std::unique_ptr<LeaseSlot>> g_slots;
size_t g_slot_buffers = 0; // slots holding a buffer (guarded by g_slots_mu)
constexpr size_t kKeptBuffers = 16;
thread_local std::vector<LeaseSlot*> t_held;

// Release the buffers of idle slots of device `dev` in the order of `st` (a stream of that device, not
// capturing): a slot no lease holds (its lock is free) whose last use has completed. Called with
// g_slots_mu held; try_lock never waits, so the lock order slot -> g_slots_mu elsewhere cannot deadlock.
void release_idle(int dev, hipStream_t st, const LeaseSlot* keep)
{
    for (auto& kv : g_slots) {
        LeaseSlot* s = kv.second.get();
        if (kv.first.first != dev || s == keep || !s->ptr || !s->mu.try_lock()) continue;
        if (s->ptr && !s->kept && s->last && hipEventQuery(s->last) == hipSuccess &&
            hipFreeAsync(s->ptr, st) == hipSuccess) {
            s->ptr = nullptr;
            s->bytes = 0;
            --g_slot_buffers;
        }
        (void)hipGetLastError();
        s->mu.unlock();
    }
}

LeaseSlot* slot_for(int dev, hipStream_t st)
{
    std::lock_guard<std::mutex> lock(g_slots_mu);
    auto& p = g_slots[{dev, st}];
    if (!p) {
        p.reset(new LeaseSlot);
        if (g_slot_buffers >= kKeptBuffers) release_idle(dev, st, p.get());
    }
    return p.get();
}

} // namespace

StreamLease::StreamLease(size_t bytes, hipStream_t stream)
{
    if (bytes == 0) return;
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    LeaseSlot* slot = capturing(stream) ? nullptr : slot_for(dev, stream);
    if (!slot || std::find(t_held.begin(), t_held.end(), slot) != t_held.end()) {
        fallback_ = new StreamScratch(bytes, stream);
        ptr_ = fallback_->get();
        return;
    }
    std::unique_lock<std::mutex> lock(slot->mu);
    if (!slot->last) check(hipEventCreateWithFlags(&slot->last, hipEventDisableTiming), "hipEventCreateWithFlags");
    if (slot->bytes < bytes) {
        // earlier calls' kernels on this stream may still read the old buffer: both steps are stream-ordered
        const size_t grown = (std::max(bytes, slot->bytes + slot->bytes / 2) + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
        if (slot->ptr) {
            check(hipFreeAsync(slot->ptr, stream), "hipFreeAsync(lease)", slot->bytes);
            slot->ptr = nullptr;
            slot->bytes = 0;
            std::lock_guard<std::mutex> g(g_slots_mu);
            --g_slot_buffers;
        }
        check(hipMallocAsync(&slot->ptr, grown, stream), "hipMallocAsync(lease)", grown);
        slot->bytes = grown;
        std::lock_guard<std::mutex> g(g_slots_mu);
        ++g_slot_buffers;
    }
    ptr_ = slot->ptr;
    slot_ = slot;
    stream_ = stream;
    t_held.push_back(slot);
    lock.release(); // held until the lease ends
}

StreamLease::~StreamLease()
{
    delete fallback_;
    if (slot_) {
        auto* slot = static_cast<LeaseSlot*>(slot_);
        // the call's kernels are queued: this event completes after its last use of the buffer
        if (hipEventRecord(slot->last, stream_) != hipSuccess) {
            slot->kept = true;
            (void)hipGetLastError();
        }
        t_held.erase(std::find(t_held.begin(), t_held.end(), slot));
        slot->mu.unlock();
    }
}

} // namespace aeth

extern "C" const char* aeth_last_error(void) { return aeth::last_error(); }

extern "C" int aeth_abi_version(void) { return AETH_ABI_VERSION; }

extern "C" int aeth_probe(int* device_count)
{
    return aeth::guarded([&] {
        int n = 0;
        aeth::check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device_count) *device_count = n;
        if (n < 1) aeth::fail("aeth_probe", "no HIP device");
    });
}
