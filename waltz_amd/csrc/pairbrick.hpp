// Chunk <-> brick map of the paired-brick path. With full 16x16x16 bricks
// (nx % 32 == 0, ny % 16 == 0, nz % 16 == 0, z band 16) WZP chunk c, i.e.
// magnitudes [8192c, 8192c + 8192), is exactly bricks 2c and 2c+1 of the
// row-major brick grid (brick base = brick index * 4096), and the 32 ranks
// [32t, 32t + 32) of that chunk form one 4x4x2 sub-block: in a 16^3 brick the
// recursive octant rank (cpu/src/codec.cpp:96-107) has bits
// z3 y3 x3 z2 y2 x2 z1 y1 x1 z0 y0 x0 from most to least significant, so the
// low five bits are x0 y0 z0 x1 y1. Host+device so a CPU test checks the map
// against the recursive rank.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define WZP_HD __host__ __device__ __forceinline__
#else
#define WZP_HD inline
#endif

namespace wz {

constexpr bool pair_bricks_fit(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t bz) {
    return bz == 16 && nx % 32 == 0 && ny % 16 == 0 && nz % 16 == 0 && nx && ny && nz;
}

// Origin of the sub-block holding ranks [32t, 32t+32) of chunk c (t < 256).
WZP_HD void pair_subblock(uint32_t nx, uint32_t ny, uint32_t c, uint32_t t, uint32_t& x, uint32_t& y,
                          uint32_t& z) {
    const uint32_t nbx = nx / 16, nby = ny / 16;
    const uint32_t b = 2 * c + (t >> 7);
    const uint32_t xi = b % nbx, r = b / nbx, yi = r % nby, zi = r / nby;
    const uint32_t u = t & 127u;  // rank bits 5..11: z1 x2 y2 z2 x3 y3 z3
    x = 16 * xi + 4 * ((u >> 1) & 1u) + 8 * ((u >> 4) & 1u);
    y = 16 * yi + 4 * ((u >> 2) & 1u) + 8 * ((u >> 5) & 1u);
    z = 16 * zi + 2 * (u & 1u) + 4 * ((u >> 3) & 1u) + 8 * ((u >> 6) & 1u);
}

// Rank (0..31) inside the sub-block of offset (dx, dy, dz), dx, dy < 4, dz < 2.
WZP_HD constexpr uint32_t subblock_rank(uint32_t dx, uint32_t dy, uint32_t dz) {
    return (dx & 1u) | ((dy & 1u) << 1) | (dz << 2) | ((dx >> 1) << 3) | ((dy >> 1) << 4);
}

// Chunk <-> brick map of the other fields of full bricks (nx % 16 == 0,
// ny % 16 == 0, nz % bz == 0, z band bz <= 16): brick b of the row-major
// brick grid holds packed positions [bV, bV + V), V = 256 bz, in recursive
// octant rank order (one table of V ranks, the full brick's), and chunk c
// holds packed positions [8192c, 8192c + E), so the bricks of a chunk are
// consecutive (a chunk starts or ends inside a brick when V does not divide
// 8192). The (de)quantizing coders visit a chunk's bricks in row-major
// order, four x-consecutive samples at a time, and place each sample at its
// rank in the chunk.
constexpr bool rank_bricks_fit(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t bz) {
    return bz >= 1 && bz <= 16 && nx % 16 == 0 && ny % 16 == 0 && nz % bz == 0 && nx && ny && nz &&
           !pair_bricks_fit(nx, ny, nz, bz);
}
// Bricks that can overlap one chunk (V = 256 at least, starting anywhere).
constexpr uint32_t kRankBricksMax = 8192 / 256 + 2;

}  // namespace wz
