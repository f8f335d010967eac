// Pixel tiling shared by the pixel-driven BP kernels and their transposes,
// and the fixed-point scale used by the deterministic transposes.
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

// Deterministic transposes accumulate in 64-bit fixed point, with one scale
// 2^E per slice chosen so that every channel sum of that slice stays below
// 2^62: stage 1 (kFixBlocks x nbatch workgroups) writes per-block maxima of
// |x| for each slice, stage 2 (one 64-lane workgroup per slice) reduces them
// and picks E.  A slice of zeros gets the scale 1 (every product is an exact
// zero); a slice holding an Inf or a NaN gets the scale 0, which the
// transpose turns into NaN for every output of that slice (a fixed-point
// sum cannot carry a non-finite value).  No memset and no atomics, so the
// pair may sit inside a captured graph.
//
// Workspace layout (fix_workspace_bytes(nbatch) bytes, 8-byte aligned):
// nbatch doubles (the scales), then kFixBlocks x nbatch floats (the maxima).
constexpr int kFixBlocks = 64;

inline size_t fix_workspace_bytes(int nbatch)
{
    return static_cast<size_t>(nbatch) * (sizeof(double) + kFixBlocks * sizeof(float));
}

static __global__ void abs_max_partial_kernel(const float* __restrict__ x, size_t per_slice,
                                              float* __restrict__ partial)
{
    __shared__ float red[4];
    const float* xs = x + blockIdx.y * per_slice;
    float m = 0.0f;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < per_slice;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const float a = fabsf(xs[i]);
        // a NaN must survive the reduction (fmaxf would drop it)
        m = (a > m || a != a) ? a : m;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float b = __shfl_xor(m, o);
        m = (b > m || b != b) ? b : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = red[0];
        for (int w = 1; w < 4; ++w) r = (red[w] > r || red[w] != red[w]) ? red[w] : r;
        partial[blockIdx.y * kFixBlocks + blockIdx.x] = r;
    }
}

static __global__ void fix_scale_kernel(const float* __restrict__ partial, double bound_factor,
                                        double* __restrict__ scale)
{
    const float* ps = partial + blockIdx.x * kFixBlocks;
    float m = 0.0f;
    for (int i = threadIdx.x; i < kFixBlocks; i += 64) m = (ps[i] > m || ps[i] != ps[i]) ? ps[i] : m;
    for (int o = 32; o > 0; o >>= 1) {
        const float b = __shfl_xor(m, o);
        m = (b > m || b != b) ? b : m;
    }
    if (threadIdx.x == 0) {
        double s;
        if (!(m <= 3.4028234663852886e38f))  // Inf or NaN
            s = 0.0;
        else if (m == 0.0f)
            s = 1.0;
        else
            s = ldexp(1.0, 62 - static_cast<int>(ceil(log2(static_cast<double>(m) * bound_factor))));
        scale[blockIdx.x] = s;
    }
}

// Enqueue the two stages for nbatch slices of per_slice floats each; the
// scales land at the start of `ws` (fix_workspace_bytes(nbatch) bytes).
inline hipError_t launch_fix_scales(const float* x, int nbatch, size_t per_slice,
                                    double bound_factor, void* ws, hipStream_t st)
{
    double* scales = static_cast<double*>(ws);
    float* partial = reinterpret_cast<float*>(scales + nbatch);
    hipLaunchKernelGGL(abs_max_partial_kernel, dim3(kFixBlocks, nbatch), dim3(256), 0, st, x,
                       per_slice, partial);
    hipLaunchKernelGGL(fix_scale_kernel, dim3(nbatch), dim3(64), 0, st, partial, bound_factor,
                       scales);
    return hipGetLastError();
}

// Fixed-point image of one contribution; a non-finite or out-of-range
// product (only in a slice whose scale is 0 or for a bound exceeded by
// rounding) adds nothing instead of converting undefined values.
__device__ __forceinline__ unsigned long long to_fixed(double q)
{
    return static_cast<unsigned long long>(fabs(q) < 9.2e18 ? __double2ll_rn(q) : 0LL);
}

}  // namespace ctr
