// Compile-time checks of the full-tile predicate that every tiled kernel (the fused evaluation, the
// reductions, the resident tape interpreter and the runtime-compiled tape kernels) uses to choose its
// unpredicated 16-byte path. Sample counts reach 2^32 - 1 (the 32-bit offset_t span), where the last
// tile's end s0 + tile passes 2^32: a 32-bit comparison wraps to 0 and wrongly takes the tile as full.
#include "../../aether_amd/csrc/units.hpp"

using aeth::units::full_tile;

// ordinary sizes
static_assert(full_tile(0, 2048, 2048));
static_assert(!full_tile(0, 2048, 2047));
static_assert(full_tile(2048, 2048, 4096));
static_assert(!full_tile(2048, 2048, 4095));
static_assert(!full_tile(0, 1024, 0));

// n = 2^32 - 1 and n = 2^32 - 4: the last 2048-sample tile starts at 2^32 - 2048 and is partial
static_assert(!full_tile(0xFFFFF800u, 2048, 0xFFFFFFFFull));
static_assert(!full_tile(0xFFFFF800u, 2048, 0xFFFFFFFCull));
// the tile before it is whole
static_assert(full_tile(0xFFFFF000u, 2048, 0xFFFFFFFFull));
static_assert(full_tile(0xFFFFF000u, 2048, 0xFFFFFFFCull));
// the tuned 4096-sample tiles of the runtime-compiled kernels
static_assert(!full_tile(0xFFFFF000u, 4096, 0xFFFFFFFFull));
static_assert(full_tile(0xFFFFE000u, 4096, 0xFFFFFFFCull));
// an exact multiple just below 2^32: n = 2^32 - 2048 ends the tile at n exactly
static_assert(full_tile(0xFFFFF000u, 2048, 0xFFFFF800ull));
static_assert(!full_tile(0xFFFFF800u, 2048, 0xFFFFF800ull));

// what the 32-bit form would have computed for the last tile: a wrap to 0, which is <= any n
static_assert((uint32_t)(0xFFFFF800u + 2048u) == 0u);

// The XCD-aware unit order is a permutation of [0, units): every unit is run exactly once, whatever
// the span and whether or not the last span is whole.
using aeth::units::xcd_chunk;
using aeth::units::xcd_unit;

constexpr bool is_permutation(uint32_t units, uint32_t K)
{
    bool seen[5000] = {};
    for (uint32_t v = 0; v < units; ++v) {
        const uint32_t u = xcd_unit(v, units, K);
        if (u >= units || seen[u]) return false;
        seen[u] = true;
    }
    return true;
}

static_assert(is_permutation(0, 32) && is_permutation(1, 32) && is_permutation(7, 1));
static_assert(is_permutation(255, 32) && is_permutation(256, 32) && is_permutation(257, 32));
static_assert(is_permutation(4097, 16) && is_permutation(4999, 64) && is_permutation(4096, 64));
static_assert(is_permutation(3 * 1024 + 5, 32) && is_permutation(1000, 2));
// inside a whole span, XCD x (workgroups b with b mod 8 = x) takes K consecutive units
static_assert(xcd_unit(0, 512, 32) == 0 && xcd_unit(8, 512, 32) == 1 && xcd_unit(1, 512, 32) == 32);
static_assert(xcd_unit(255, 512, 32) == 255 && xcd_unit(256, 512, 32) == 256 && xcd_unit(257, 512, 32) == 288);
// the partial last span keeps launch order
static_assert(xcd_unit(300, 300, 32) == 300 && xcd_unit(299, 300, 32) == 299 && xcd_unit(256, 300, 32) == 256);
// chunks of about 512 KiB of one plane
static_assert(xcd_chunk(2048 * 8) == 32 && xcd_chunk(2048 * 4) == 64 && xcd_chunk(1024 * 8) == 64);
static_assert(xcd_chunk(1u << 20) == 1);

int main() { return 0; }
