// math.hip -- elementwise transcendental maps (the trig stress workload).
// fn 0..3 apply one function; fn 4 reads x once and writes sin, cos, exp and
// log of it to four planes (4 bytes in, 16 bytes out per float input).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/aether_hip.h"
#include "common.hpp"
#include "dmath.hpp"

namespace aeth {
namespace {

constexpr int kThreads = 256;

template<int FN>
__device__ __forceinline__ double apply(double x)
{
    if constexpr (FN == AETH_FN_SIN) return dm::sin(x);
    else if constexpr (FN == AETH_FN_COS) return dm::cos(x);
    else if constexpr (FN == AETH_FN_EXP) return dm::exp(x);
    else return dm::log(x);
}

template<int FN>
__device__ __forceinline__ float apply(float x)
{
    if constexpr (FN == AETH_FN_SIN) return dm::sinf(x);
    else if constexpr (FN == AETH_FN_COS) return dm::cosf(x);
    else if constexpr (FN == AETH_FN_EXP) return dm::expf(x);
    else return dm::logf(x);
}

template<class T, int FN>
__global__ __launch_bounds__(kThreads) void map_kernel(const T* __restrict__ x, T* __restrict__ y, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        const T v = x[i];
        if constexpr (FN == AETH_FN_ALL4) {
            if constexpr (sizeof(T) == 8) {
                double s, c;
                dm::sincos(v, &s, &c);
                y[i] = s;
                y[n + i] = c;
            } else {
                float s, c;
                dm::sincosf(v, &s, &c);
                y[i] = s;
                y[n + i] = c;
            }
            y[2 * n + i] = apply<AETH_FN_EXP>(v);
            y[3 * n + i] = apply<AETH_FN_LOG>(v);
        } else {
            y[i] = apply<FN>(v);
        }
    }
}

template<class T>
void launch(int fn, const T* x, T* y, uint64_t n, hipStream_t st)
{
    const uint64_t blocks = (n + kThreads - 1) / kThreads;
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(blocks, (uint64_t)device_cus() * 32u));
    switch (fn) {
    case AETH_FN_SIN: hipLaunchKernelGGL((map_kernel<T, AETH_FN_SIN>), dim3(grid), dim3(kThreads), 0, st, x, y, n); break;
    case AETH_FN_COS: hipLaunchKernelGGL((map_kernel<T, AETH_FN_COS>), dim3(grid), dim3(kThreads), 0, st, x, y, n); break;
    case AETH_FN_EXP: hipLaunchKernelGGL((map_kernel<T, AETH_FN_EXP>), dim3(grid), dim3(kThreads), 0, st, x, y, n); break;
    case AETH_FN_LOG: hipLaunchKernelGGL((map_kernel<T, AETH_FN_LOG>), dim3(grid), dim3(kThreads), 0, st, x, y, n); break;
    case AETH_FN_ALL4:
        hipLaunchKernelGGL((map_kernel<T, AETH_FN_ALL4>), dim3(grid), dim3(kThreads), 0, st, x, y, n);
        break;
    default: fail("aeth_math", "unknown function");
    }
    check(hipGetLastError(), "aeth_math launch");
}

} // namespace
} // namespace aeth

using namespace aeth;

extern "C" int aeth_math(int fn, int dtype, const void* x, void* y, uint64_t n, void* stream)
{
    return guarded([&] {
        if (n == 0) return;
        if (!x || !y) fail("aeth_math", "NULL pointer");
        if (dtype == AETH_F64)
            launch<double>(fn, static_cast<const double*>(x), static_cast<double*>(y), n, as_stream(stream));
        else if (dtype == AETH_F32)
            launch<float>(fn, static_cast<const float*>(x), static_cast<float*>(y), n, as_stream(stream));
        else
            fail("aeth_math", "unknown dtype");
    });
}
