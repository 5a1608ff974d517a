// core.hip -- error state, device queries and library identity.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
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
