// rng.hpp -- Philox4x32-10 and aether's draw conversions, device side.
//
// Counter layout, key and conversions are aether's (random/detail/Philox.h:
// 76-133, random/detail/Backend.h:75-236): block(seed, global, counter) =
// philox4x32_10({lo(counter), hi(counter), lo(global), hi(global)},
// key = (lo(seed), hi(seed))). The bijection is Salmon et al., SC'11.
// Pure 32-bit integer VALU: each round is two 32x32->64 products
// (v_mad_u64_u32) plus two three-input xors (gfx950's v_bitop3_b32 with the
// XOR truth table, one instruction where two-input xors take two); the key
// bump is two scalar adds when the key is uniform.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmath.hpp"

namespace aeth {
namespace rng {

struct U4 {
    uint32_t x, y, z, w;
};

// x ^ y ^ z in one instruction: truth table 0x96 is the parity of the three input bits.
__device__ __forceinline__ uint32_t xor3(uint32_t x, uint32_t y, uint32_t z)
{
    return __builtin_amdgcn_bitop3_b32(x, y, z, 0x96);
}

__device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        // the first two rounds keep plain xors: callers' counters are often zero or uniform there, which the
        // compiler folds through a xor but not through the three-input op
        if (round < 2) c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        else c = U4{xor3(hi1, c.y, k0), lo1, xor3(hi0, c.w, k1), lo0};
    }
    return c;
}

// Backend.h:144-150 (global is the 32-bit offset_t of the reference, so its
// high counter word is zero).
__device__ __forceinline__ U4 block(uint64_t seed, uint32_t global, uint64_t counter)
{
    return philox(U4{(uint32_t)counter, (uint32_t)(counter >> 32), global, 0u}, (uint32_t)seed,
                  (uint32_t)(seed >> 32));
}

// Backend.h:123-126: w 2^-32 + 2^-33 (exact)
__device__ __forceinline__ double u01_from_word_d(uint32_t w) { return (double)w * 0x1p-32 + 0x1p-33; }
// Backend.h:131-134: (2(w>>9)+1) 2^-24 (exact)
__device__ __forceinline__ float u01_from_word_f(uint32_t w) { return (float)(((w >> 9) << 1) | 1u) * 0x1p-24f; }

// Backend.h:209-236 (double): four words, Box-Muller, first normal only.
__device__ __forceinline__ double normal_from_block_d(U4 b)
{
    const uint64_t zx = (uint64_t)b.x ^ ((uint64_t)b.y << 21);
    const double u = (double)zx * 0x1p-53 + 0x1p-54;
    const uint64_t zy = (uint64_t)b.z ^ ((uint64_t)b.w << 21);
    const double v = (double)zy * (0x1p-53 * 2.0) + 0x1p-53;
    const double s = __builtin_sqrt(-2.0 * dm::log(u));
    return s * dm::sin(v * 3.1415926535897932);
}

// Backend.h:209-236 (float): two words.
__device__ __forceinline__ float normal_from_block_f(U4 b)
{
    const float k2pi = 2.3283064e-10f * 6.2831855f;
    const float u = u01_from_word_f(b.x);
    const float v = (float)b.y * k2pi + (k2pi / 2.0f);
    const float s = __builtin_sqrtf(-2.0f * dm::logf(u));
    return s * dm::sinf(v);
}

} // namespace rng
} // namespace aeth
