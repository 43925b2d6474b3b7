// XCD-aware tile order for 1-D grids (device code only).
//
// An MI355X launch deals its workgroups to the 8 XCDs round-robin (block b
// -> XCD b % 8), and each XCD has its own L2.  Block b takes tile
// (b % 8) * (nblocks / 8) + b / 8 in row-major tile order, so each XCD
// sweeps one contiguous band of the image and neighbouring tiles, which
// gather from the same rows of the other views, share that XCD's L2.
#pragma once

#include <hip/hip_runtime.h>

namespace hp3d {

// The grid must be 1-D with a multiple of 8 blocks.  Returns false for the
// padding blocks past the last tile.
__device__ __forceinline__ bool xcd_tile(int tiles_x, int tiles, int& bx, int& by)
{
    const int nb = gridDim.x, b = blockIdx.x;
    const int t = (b & 7) * (nb >> 3) + (b >> 3);
    if (t >= tiles) return false;
    by = t / tiles_x;
    bx = t - by * tiles_x;
    return true;
}

// Blocks for a 1-D XCD-ordered grid over `tiles` tiles: a multiple of 8.
inline unsigned xcd_grid(int tiles) { return (unsigned)((tiles + 7) / 8 * 8); }

}  // namespace hp3d
