// flag.hpp -- DeviceFlagView: the device-side half of the error contract (err/DeviceFlag.h:49-84).
//
// Device code cannot throw; it sets a sticky 32-bit word instead, and the host reads it back, clears it and
// turns it into an aether::Error (aether::DeviceFlag in aether/aether.hpp owns the word). The view is what a
// kernel takes by value. Included by HIP sources (device arm: an atomic OR of 1 into the word, so any number
// of lanes may set it at once) and by host translation units (host arm: a relaxed atomic store of 1 into a
// word of host memory -- OR with 1 into a word that only ever holds 0 or 1 is a store).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AETH_FLAG_HD __host__ __device__
#else
#define AETH_FLAG_HD
#endif

#if !defined(__HIP_DEVICE_COMPILE__)
#include <atomic>
#endif

namespace aether {

class DeviceFlagView {
public:
    DeviceFlagView() = default;
    AETH_FLAG_HD constexpr explicit DeviceFlagView(uint32_t* word) : word_(word) {}

    AETH_FLAG_HD uint32_t* data() const { return word_; }

    /** Set the flag (sticky: only the host clears it). */
    AETH_FLAG_HD void set() const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        atomicOr(word_, 1u);
#else
        std::atomic_ref<uint32_t>(*word_).store(1u, std::memory_order_relaxed);
#endif
    }

private:
    uint32_t* word_ = nullptr;
};

} // namespace aether
