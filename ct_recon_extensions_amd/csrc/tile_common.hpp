// Pixel tiling shared by the pixel-driven BP kernels and their transposes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace ctr {

constexpr int kAnchorChunk = 64;  // anchors computed per batch (one per lane of wave 0)

template <int TW, int TH>
struct TileShape {
    static constexpr int kPix = TW * TH;
    static constexpr int kBlock = kPix >= 256 ? 256 : kPix;
    static constexpr int kPPT = kPix / kBlock;
    static_assert(kPix % kBlock == 0, "tile must be a multiple of the block");
};

// Workgroups are dispatched to the 8 XCDs round-robin by linear id, and
// each XCD has its own L2.  Mapping the id so that every XCD receives one
// contiguous band of tiles keeps the detector spans a band needs within one
// L2 (instead of every XCD reading every sinogram row).  Bijective for any
// tile count; ids >= ntiles (none in practice) map to themselves.
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_tile(int id, int ntiles)
{
    if (id >= ntiles) return id;
    const int q = ntiles / kXcds, r = ntiles % kXcds;
    const int x = id % kXcds, k = id / kXcds;
    return x * q + min(x, r) + k;
}

// Slices per workgroup for the BP engine: slice-interleaved staging
// (a multiple of 4 slices) addresses four slices of na x ns floats through
// one 32-bit buffer range, so at 512 MiB per slice it gives way to 2.
inline int engine_nb(int nb, long long na, long long ns)
{
    return (nb > 2 && na * ns * 4 >= (1LL << 29)) ? 2 : nb;
}

// Tile order of the BP engine's forward kernel: supertiles of kSuper x
// kSuper tiles, row-major, tiles row-major inside each, and every XCD given a
// contiguous run of that order (xcd_tile over all slice groups).  An XCD then
// holds a compact block of the image, whose detector spans overlap far more
// than those of a band of whole tile rows, so more of its staging hits its
// own L2.  The grid is padded to whole supertiles; padding workgroups get -1.
constexpr int kSuper = 8;

__host__ __device__ inline int super_grid(int tiles_x, int tiles_y)
{
    return ((tiles_x + kSuper - 1) / kSuper) * ((tiles_y + kSuper - 1) / kSuper) * kSuper * kSuper;
}

// Workgroup (bx, by) of a grid of super_grid(...) x gridDim.y -> tile index
// (or -1) and slice group.
__device__ __forceinline__ int super_tile(int tiles_x, int tiles_y, int& group)
{
    const int g = gridDim.x;
    const int l = xcd_tile(blockIdx.x + blockIdx.y * g, g * gridDim.y);
    group = l / g;
    const int w = l - group * g;
    const int sx = (tiles_x + kSuper - 1) / kSuper;
    const int s = w / (kSuper * kSuper), i = w - s * (kSuper * kSuper);
    const int tx = (s % sx) * kSuper + i % kSuper, ty = (s / sx) * kSuper + i / kSuper;
    return (tx < tiles_x && ty < tiles_y) ? ty * tiles_x + tx : -1;
}

// Origin of a tile and the pixel (ic, jc) its anchor is evaluated at.
struct TileFrame {
    int x0, y0, ic, jc;
    __host__ __device__ TileFrame(int tile, int tiles_x, int nx, int ny, int TW, int TH)
    {
        x0 = (tile % tiles_x) * TW;
        y0 = (tile / tiles_x) * TH;
        ic = min(x0 + TW / 2, nx - 1);
        jc = min(y0 + TH / 2, ny - 1);
    }
};

// Reference edge rule (src/bp_fan_2d.cc:72, src/bp_par_2d.cc:71): a sample
// at s = s_int + frac counts iff 0 <= s <= ns-1.
__device__ __forceinline__ bool on_detector(int s_int, float frac, int ns)
{
    return s_int >= 0 && (s_int < ns - 1 || (s_int == ns - 1 && frac == 0.0f));
}

}  // namespace ctr
