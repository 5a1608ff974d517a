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

int main() { return 0; }
