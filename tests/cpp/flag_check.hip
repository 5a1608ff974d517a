// flag_check.hip -- drives aether::DeviceFlag / DeviceFlagView the way a user's own HIP kernels would, for
// tests/test_device_flag.py (mode "host", no GPU) and tests/test_gpu_device_flag.py (mode "device"). The
// cases are the reference's own (tests/test_DeviceFlag.cpp, tests/test_DeviceFlag.cu): initially clear;
// set from a kernel; checkAndClear throws once, then clears; clear without throwing; many lanes racing to
// set it; plus the consumer (AppendCounter's overflow) and a move. Built by the package build with hipcc
// against libaether_hip.so. Prints "ok" or the first failed check and exits 1.
#include <aether/aether.hpp>
#include <aether/device/flag.hpp>

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

int failures = 0;

#define CHECK(cond)                                                                                                   \
    do {                                                                                                              \
        if (!(cond)) {                                                                                                \
            std::printf("check failed at line %d: %s\n", __LINE__, #cond);                                           \
            ++failures;                                                                                               \
            return;                                                                                                   \
        }                                                                                                             \
    } while (0)

template<class F>
bool throws_error(F&& f, const char* needle)
{
    try {
        f();
    } catch (const aether::Error& e) {
        return std::strstr(e.what(), needle) != nullptr;
    }
    return false;
}

__global__ void set_flag_kernel(aether::DeviceFlagView flag) { flag.set(); }

// every lane of the grid races to set the same word
__global__ void race_set_flag_kernel(aether::DeviceFlagView flag) { flag.set(); }

// sets the flag only where the input says so (a device error path)
__global__ void set_where_kernel(aether::DeviceFlagView flag, const int* bad, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && bad[i]) flag.set();
}

void launch_check(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw aether::Error(std::string(what) + ": " + hipGetErrorString(e));
}

void host_mode()
{
    // the host arm over a word of host memory: a relaxed atomic store of 1
    uint32_t word = 0;
    aether::DeviceFlagView v(&word);
    CHECK(v.data() == &word);
    v.set();
    CHECK(word == 1u);
    word = 0;
    std::vector<std::thread> ts;
    for (int t = 0; t < 8; ++t)
        ts.emplace_back([v] {
            for (int k = 0; k < 512; ++k) v.set();
        });
    for (auto& t : ts) t.join();
    CHECK(word == 1u);
}

void device_mode()
{
    {   // InitiallyClear
        aether::DeviceFlag flag;
        CHECK(!flag.isSet());
    }
    {   // SetMakesIsSetTrue; a read does not clear it
        aether::DeviceFlag flag;
        hipLaunchKernelGGL(set_flag_kernel, dim3(1), dim3(1), 0, nullptr, flag.deviceView());
        launch_check("set_flag_kernel");
        CHECK(flag.isSet());
        CHECK(flag.isSet());
    }
    {   // CheckAndClearThrowsWhenSetThenClears
        aether::DeviceFlag flag;
        hipLaunchKernelGGL(set_flag_kernel, dim3(1), dim3(1), 0, nullptr, flag.deviceView());
        launch_check("set_flag_kernel");
        CHECK(flag.isSet());
        CHECK(throws_error([&] { flag.checkAndClear("DeviceFlagTest"); }, "DeviceFlagTest failed on DeviceFlag"));
        CHECK(!flag.isSet());
        CHECK(!throws_error([&] { flag.checkAndClear("DeviceFlagTest"); }, ""));
    }
    {   // CheckAndClearDoesNotThrowWhenClear
        aether::DeviceFlag flag;
        CHECK(!throws_error([&] { flag.checkAndClear("DeviceFlagTest"); }, ""));
    }
    {   // ClearResetsWithoutThrowing
        aether::DeviceFlag flag;
        hipLaunchKernelGGL(set_flag_kernel, dim3(1), dim3(1), 0, nullptr, flag.deviceView());
        launch_check("set_flag_kernel");
        CHECK(flag.isSet());
        flag.clear();
        CHECK(!flag.isSet());
    }
    {   // ConcurrentSetFromManyThreadsIsIdempotent
        aether::DeviceFlag flag;
        hipLaunchKernelGGL(race_set_flag_kernel, dim3(64), dim3(256), 0, nullptr, flag.deviceView());
        launch_check("race_set_flag_kernel");
        CHECK(flag.isSet());
        uint32_t raw = 0;
        CHECK(hipMemcpy(&raw, flag.deviceView().data(), 4, hipMemcpyDeviceToHost) == hipSuccess);
        CHECK(raw == 1u);
    }
    {   // a device error path: set only by the lane that meets a bad input; sticky across launches
        aether::DeviceFlag flag;
        const int n = 4096;
        int* bad = nullptr;
        CHECK(hipMalloc(&bad, n * sizeof(int)) == hipSuccess);
        CHECK(hipMemset(bad, 0, n * sizeof(int)) == hipSuccess);
        hipLaunchKernelGGL(set_where_kernel, dim3(n / 256), dim3(256), 0, nullptr, flag.deviceView(), bad, n);
        launch_check("set_where_kernel");
        CHECK(!flag.isSet());
        const int one = 1;
        CHECK(hipMemcpy(bad + 3001, &one, sizeof one, hipMemcpyHostToDevice) == hipSuccess);
        hipLaunchKernelGGL(set_where_kernel, dim3(n / 256), dim3(256), 0, nullptr, flag.deviceView(), bad, n);
        launch_check("set_where_kernel");
        CHECK(hipMemset(bad, 0, n * sizeof(int)) == hipSuccess);
        hipLaunchKernelGGL(set_where_kernel, dim3(n / 256), dim3(256), 0, nullptr, flag.deviceView(), bad, n);
        launch_check("set_where_kernel");
        CHECK(flag.isSet()); // a clean launch after the bad one leaves it set
        CHECK(hipFree(bad) == hipSuccess);
    }
    {   // a moved flag keeps its word
        aether::DeviceFlag a;
        hipLaunchKernelGGL(set_flag_kernel, dim3(1), dim3(1), 0, nullptr, a.deviceView());
        launch_check("set_flag_kernel");
        aether::DeviceFlag b(std::move(a));
        CHECK(b.isSet());
        aether::DeviceFlag c;
        c = std::move(b);
        CHECK(c.isSet());
    }
    {   // the consumer: an AppendCounter's overflow is a DeviceFlag
        aether::Array<double> x(8192);
        x.deviceView() = aether::random::Generator(5).uniform<double, 1>(0.0, 1.0);
        aether::accum::AppendCounter ctr(100);
        aether::Array<std::uint32_t> out(0);
        out.reserve(100);
        aether::accum::appendWhere(x.deviceView(), 0.5, ctr, out);
        CHECK(ctr.overflowed());
        CHECK(ctr.overflow().isSet());
        CHECK(throws_error([&] { ctr.checkOverflow("appendWhere"); }, "appendWhere failed on DeviceFlag"));
        CHECK(!ctr.overflowed());
        CHECK(ctr.commit(out) == 100);
    }
}

} // namespace

int main(int argc, char** argv)
{
    const std::string mode = argc > 1 ? argv[1] : "";
    try {
        if (mode == "host") host_mode();
        else if (mode == "device") device_mode();
        else {
            std::printf("usage: flag_check host|device\n");
            return 2;
        }
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 1;
    }
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
