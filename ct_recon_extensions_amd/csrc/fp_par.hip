// Ray-driven parallel-beam forward projection for gfx950 and its
// gather-form transpose.
//
// FP (fp_kernel): one thread per ray (view, channel) for NB slices; a
// workgroup covers 256 consecutive channels of one view, so the sinogram row
// is written coalesced.  The ray marches the rows of the image (x-branch) or
// its columns (y-branch); the branch, the start position and the step of
// every ray come from the reference's FP64 tables (src/fp_par_2d.cc:149-208)
// and the position advances by repeated FP64 addition exactly as the
// reference does (src/fp_par_2d.cc:236-273), so sample inclusion
// (0 <= x <= nx-1) and tap choice match it; the lerp and the sum are FP32
// (the reference also sums in T).  For y-branch views the kernel reads a
// transposed copy of the slice so both branches walk contiguous rows.
//
// Transpose (fpt_kernel): one thread per pixel for NB slices; per view the
// channels whose rays pass within one pixel step are gathered with the
// triangle weight 1 - |mid - is|/o1 of the reference's transpose
// (src/fp_par_2d.cc:355-385), the channel range clamped to the detector.
//
// Transpose on the BP engine (FptGeo): the pixel's centre projects to the
// channel mid = begin + i*o1 + j*o2 (linear in the pixel), and the channels
// within |o1| of it get 1 - |c - mid|/|o1|; with |o1| <= 1 (pixel pitch along
// the marched axis at most the channel pitch) those are at most floor(mid)
// and floor(mid)+1, i.e. two taps, so the engine's LDS-staged spans, slice
// batching and FP64 anchors apply unchanged.  The first and last pixel
// along the marched axis keep one side only (src/fp_par_2d.cc:360-368).
#include "bp_engine.hpp"
#include "ctr_common.hpp"
#include "fpt_geom.hpp"
#include "tile_common.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <type_traits>
#include <unordered_map>

namespace ctr {

struct FpProblem {
    int ns, na, nx, ny, nbatch;
    double dx, dy;
};

// Transposed copy src[b][iy][ix] -> dst[b][ix][iy] through a 32x33 LDS tile.
template <typename T>
__global__ void transpose_kernel(const T* __restrict__ src, T* __restrict__ dst, int ny, int nx)
{
    __shared__ T tile[32][33];
    const size_t base = static_cast<size_t>(blockIdx.z) * nx * ny;
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int y = y0 + r, x = x0 + threadIdx.x;
        if (y < ny && x < nx) tile[r][threadIdx.x] = src[base + static_cast<size_t>(y) * nx + x];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int x = x0 + r, y = y0 + threadIdx.x;
        if (y < ny && x < nx) dst[base + static_cast<size_t>(x) * ny + y] = tile[threadIdx.x][r];
    }
}

// img   : [b][ny][nx] (x-branch rows) ; imgT : [b][nx][ny] (y-branch rows)
template <typename T, int NB>
__global__ __launch_bounds__(256) void fp_kernel(const T* __restrict__ img,
                                                 const T* __restrict__ imgT,
                                                 T* __restrict__ proj,
                                                 const double* __restrict__ sincos,
                                                 const double* __restrict__ bo,
                                                 const int* __restrict__ usex, FpProblem P)
{
    using Acc = typename std::conditional<sizeof(T) == 8, double, float>::type;
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int a = blockIdx.y;
    const int b0 = blockIdx.z * NB;
    const int nb = min(NB, P.nbatch - b0);
    if (s >= P.ns) return;
    const bool xb = usex[a] != 0;
    // x-branch walks ny rows of length nx; y-branch walks nx rows of length ny
    const int nsteps = xb ? P.ny : P.nx;
    const int rowlen = xb ? P.nx : P.ny;
    const T* src = xb ? img : imgT;
    const double dw = xb ? fabs(P.dy / sincos[2 * a + 1]) : fabs(P.dx / sincos[2 * a]);
    double x = bo[2 * (static_cast<size_t>(a) * P.ns + s)];
    const double off = bo[2 * (static_cast<size_t>(a) * P.ns + s) + 1];
    const double hi = rowlen - 1;
    const size_t slice = static_cast<size_t>(P.nx) * P.ny;
    Acc sum[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) sum[b] = Acc(0);
    for (int k = 0; k < nsteps; ++k) {
        if (x >= 0 && x <= hi) {
            const double fl = floor(x);
            const int i1 = static_cast<int>(fl);
            const int i2 = min(i1 + 1, rowlen - 1);
            const Acc u = static_cast<Acc>(x - fl);
            const T* row = src + static_cast<size_t>(k) * rowlen;
            // every slice's taps loaded first (a slice past the batch
            // re-reads the last one and is not added)
            T t0[NB], t1[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                t0[b] = row[(b0 + min(b, nb - 1)) * slice + i1];
                t1[b] = row[(b0 + min(b, nb - 1)) * slice + i2];
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (b < nb) {
                    const Acc p0 = t0[b];
                    const Acc p1 = t1[b];
                    sum[b] += p0 + u * (p1 - p0);
                }
            }
        }
        x += off;
    }
#pragma unroll
    for (int b = 0; b < NB; ++b)
        if (b < nb)
            proj[(static_cast<size_t>(b0 + b) * P.na + a) * P.ns + s] =
                static_cast<T>(static_cast<double>(sum[b]) * dw);
}

// FP with LDS-staged image rows (float32).  A workgroup owns up to 256
// consecutive channels (one per thread) of V consecutive views and NB
// slices, and marches the rows of the image (x-branch views) or of its
// transpose (y-branch views) in chunks of kFpRows: per chunk it stages, for
// all NB slices, the span of columns its rays can touch (DMA from global
// straight into LDS, zero outside the row; slice rows kFpSpan words apart),
// then every thread advances its V rays through those rows.  Two chunk
// buffers alternate: the DMAs of chunk h + 1 are issued before the rays walk
// chunk h, so their latency hides behind that walk and one barrier per chunk
// separates the two.  A staged pixel is read by ~V rays instead of every view
// re-reading the image through L1/L2, and a ray's two taps for a pair of
// slices come from one ds_read2st64_b32.  Each ray keeps the reference's FP64
// position accumulation (x += offset per row) and inclusion test; the lerp
// and the sum are FP32.  Views of the other branch in the group take a second
// pass.
template <int NB, int V, int kFpRows, int kFpSpan>
__global__ __launch_bounds__(256) void fp_lds_kernel(const float* __restrict__ img,
                                                     const float* __restrict__ imgT,
                                                     float* __restrict__ proj,
                                                     const double* __restrict__ sincos,
                                                     const double* __restrict__ bo,
                                                     const int* __restrict__ usex, FpProblem P,
                                                     int cpb)
{
    static_assert(NB % 2 == 0 || NB == 1, "slices are read in pairs");
    constexpr int kChunk = kFpRows * NB * kFpSpan;  // floats of one chunk buffer
    __shared__ __align__(16) float rows[2 * kChunk];
    // the wave index is uniform: keeps the DMA's buffer descriptor in SGPRs
    // (no per-lane waterfall around buffer_load ... lds)
    const int tid = threadIdx.x, lane = tid & 63, wave = uniform_i(tid >> 6);
    const int s0 = blockIdx.x * cpb, s = s0 + tid;
    const int slast = min(s0 + cpb, P.ns) - 1;
    const bool has_ray = tid < cpb && s < P.ns;
    const int a0 = blockIdx.y * V;
    const int b0 = blockIdx.z * NB;
    const int nb = min(NB, P.nbatch - b0);
    const size_t slice = static_cast<size_t>(P.nx) * P.ny;
    // right tap through an offset the compiler cannot fold, so the slice
    // pairs load as ds_read2st64_b32 (rows kFpSpan apart); at x == hi its
    // weight u is 0 and it reads a staged zero
    int one = 1;
    asm volatile("" : "+v"(one));

    for (int pass = 0; pass < 2; ++pass) {
        const bool xb = pass == 0;
        bool act[V];
        int nact = 0;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int a = a0 + v;
            act[v] = a < P.na && (usex[a] != 0) == xb;
            nact += act[v] ? 1 : 0;
        }
        if (nact == 0) continue;  // uniform
        const int nsteps = xb ? P.ny : P.nx;
        const int rowlen = xb ? P.nx : P.ny;
        const float* src = xb ? img : imgT;
        const double hi = rowlen - 1;
        double x[V], off[V];
        // view v's rays cover columns [clo[v] + k0 * coff[v], chi[v] + k0 * coff[v]]
        // over rows k0 .. k0 + kFpRows - 1 (FP32 bounds: their error is far
        // below the column of margin each side)
        float clo[V], chi[V], coff[V];
        float sum[V][NB];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int a = min(a0 + v, P.na - 1);
            const double* r = bo + 2 * static_cast<size_t>(a) * P.ns;
            x[v] = has_ray ? r[2 * s] : 0.0;
            off[v] = r[1];
            const double blo = r[2 * s0], bhi = r[2 * slast];
            const double drift = (kFpRows - 1) * off[v];
            clo[v] = act[v] ? static_cast<float>(fmin(blo, bhi) + fmin(0.0, drift)) : 1e30f;
            chi[v] = act[v] ? static_cast<float>(fmax(blo, bhi) + fmax(0.0, drift)) : -1e30f;
            coff[v] = static_cast<float>(off[v]);
#pragma unroll
            for (int b = 0; b < NB; ++b) sum[v][b] = 0.0f;
        }
        // columns cmin .. cmin + 64 * nload - 1 of the chunk at row k0, or
        // false when no ray of the chunk is on the image
        auto span = [&](int k0, int& cmin, int& nload) -> bool {
            const float k = static_cast<float>(k0);
            float xmin = 1e30f, xmax = -1e30f;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                xmin = fminf(xmin, fmaf(k, coff[v], clo[v]));
                xmax = fmaxf(xmax, fmaf(k, coff[v], chi[v]));
            }
            if (xmax < -1.0f || xmin > static_cast<float>(hi) + 1.0f) return false;
            // one column of margin each side; the right tap of a sample at
            // x == nx-1 reads cmax + 1 (zero past the row); the host keeps
            // cmax - cmin + 2 <= kFpSpan
            cmin = max(0, static_cast<int>(floorf(xmin)) - 1);
            const int cmax = min(rowlen - 1, static_cast<int>(floorf(xmax)) + 2);
            nload = min((cmax - cmin + 65) / 64, kFpSpan / 64);
            return true;
        };
        // rows (k0 + kk, slice b) -> buf[(kk * NB + b) * kFpSpan + c - cmin]
        auto stage = [&](int k0, int cmin, int nload, float* buf) {
            const int nr = min(kFpRows, nsteps - k0);
            for (int e = wave; e < nr * NB; e += 4) {
                const int kk = e / NB, b = e - kk * NB;
                const float* rowp = src + static_cast<size_t>(b0 + b) * slice +
                                    static_cast<size_t>(k0 + kk) * rowlen;
                const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                    const_cast<float*>(rowp), 0,
                    b < nb ? static_cast<unsigned int>(rowlen) * 4u : 0u, 0x00020000);
                float* dst = buf + e * kFpSpan;
                for (int c = 0; c < nload; ++c)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                        rs, (__attribute__((address_space(3))) void*)(dst + 64 * c), 4,
                        (cmin + 64 * c + lane) * 4, 0, 0, 0);
            }
        };
        const int nchunks = (nsteps + kFpRows - 1) / kFpRows;
        int cm = 0, nl = 0;
        __syncthreads();  // the previous pass has finished reading both buffers
        bool on = span(0, cm, nl);
        if (on) stage(0, cm, nl, rows);
        for (int h = 0; h < nchunks; ++h) {
            const int k0 = h * kFpRows;
            const int nr = min(kFpRows, nsteps - k0);
            const float* cur = rows + (h & 1) * kChunk;
            // this wave's DMAs of chunk h have landed (none of chunk h + 1 is
            // issued yet); after the barrier every wave's have, and every
            // wave is done with chunk h - 1, whose buffer takes chunk h + 1
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            int cm_next = 0, nl_next = 0;
            bool on_next = false;
            if (h + 1 < nchunks) {
                on_next = span(k0 + kFpRows, cm_next, nl_next);
                if (on_next) stage(k0 + kFpRows, cm_next, nl_next, rows + ((h + 1) & 1) * kChunk);
            }
            if (on && has_ray) {
#pragma unroll 1
                for (int kk = 0; kk < nr; ++kk) {
                    const float* rk = cur + kk * NB * kFpSpan - cm;
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        if (act[v] && x[v] >= 0 && x[v] <= hi) {
                            // x >= 0: truncation is the floor, and the fraction is exact
                            const int i1 = static_cast<int>(x[v]);
                            const float u = static_cast<float>(__builtin_amdgcn_fract(x[v]));
                            const float w0 = 1.0f - u;
                            const float* q0 = rk + i1;
                            const float* q1 = q0 + one;
#pragma unroll
                            for (int b = 0; b < NB; ++b)
                                sum[v][b] = fmaf(u, q1[b * kFpSpan], fmaf(w0, q0[b * kFpSpan], sum[v][b]));
                        }
                        x[v] += off[v];
                    }
                }
            } else {
                // no ray of the chunk is on the image (or no ray here): only
                // advance them
#pragma unroll
                for (int v = 0; v < V; ++v)
                    for (int k = 0; k < nr; ++k) x[v] += off[v];
            }
            on = on_next;
            cm = cm_next;
        }
        if (has_ray) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                if (!act[v]) continue;
                const int a = a0 + v;
                const double dw = xb ? fabs(P.dy / sincos[2 * a + 1]) : fabs(P.dx / sincos[2 * a]);
#pragma unroll
                for (int b = 0; b < NB; ++b)
                    if (b < nb)
                        proj[(static_cast<size_t>(b0 + b) * P.na + a) * P.ns + s] =
                            static_cast<float>(static_cast<double>(sum[v][b]) * dw);
            }
        }
    }
}

// Pixels of the image border (first / last row and column), i < 2nx + 2(ny-2).
__device__ __forceinline__ void border_pixel(int i, int nx, int ny, int& ix, int& iy)
{
    if (i < nx) {
        ix = i;
        iy = 0;
    } else if (i < 2 * nx) {
        ix = i - nx;
        iy = ny - 1;
    } else {
        const int j = i - 2 * nx;
        ix = j < ny - 2 ? 0 : nx - 1;
        iy = 1 + (j < ny - 2 ? j : j - (ny - 2));
    }
}

// Contribution of view a to pixel (ix, iy) for slices b0 .. b0+nb-1,
// added to sum[]: the reference's gather (src/fp_par_2d.cc:355-385).
template <typename T, int NB>
__device__ __forceinline__ void fpt_view(const T* __restrict__ proj,
                                         const double* __restrict__ weights,
                                         const double* __restrict__ pos,
                                         const int* __restrict__ usex, const FpProblem& P,
                                         int stride, int a, int ix, int iy, int b0, int nb,
                                         double (&sum)[NB])
{
    using Acc = typename std::conditional<sizeof(T) == 8, double, float>::type;
    const size_t pslice = static_cast<size_t>(P.na) * P.ns;
    double mid, left, right, o1;
    // mid exactly as the reference rounds it (a product, then a sum): the
    // one-sided border rule below is decided on it
    if (usex[a]) {
        const double* pp = pos + 2 * (static_cast<size_t>(a) * stride + iy);
        o1 = pp[1];
        mid = __dadd_rn(pp[0], __dmul_rn(o1, static_cast<double>(ix)));
        left = ix == 0 ? mid : mid - o1;
        right = ix == P.nx - 1 ? mid : mid + o1;
    } else {
        const double* pp = pos + 2 * (static_cast<size_t>(a) * stride + ix);
        o1 = pp[1];
        mid = __dadd_rn(pp[0], __dmul_rn(o1, static_cast<double>(iy)));
        left = iy == 0 ? mid : mid - o1;
        right = iy == P.ny - 1 ? mid : mid + o1;
    }
    if (left > right) {
        const double t = left;
        left = right;
        right = t;
    }
    const int c0 = max(static_cast<int>(ceil(left)), 0);
    const int c1 = min(static_cast<int>(floor(right)), P.ns - 1);
    Acc tmp[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) tmp[b] = Acc(0);
    for (int c = c0; c <= c1; ++c) {
        const Acc w = static_cast<Acc>(1.0 - fabs((mid - c) / o1));
        const T* pr = proj + static_cast<size_t>(a) * P.ns + c;
        // every slice's value loaded first (a slice past the batch re-reads
        // the last one and is not added)
        T val[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) val[b] = pr[(b0 + min(b, nb - 1)) * pslice];
#pragma unroll
        for (int b = 0; b < NB; ++b)
            if (b < nb) tmp[b] += w * static_cast<Acc>(val[b]);
    }
    const double wa = weights[a];
#pragma unroll
    for (int b = 0; b < NB; ++b) sum[b] += static_cast<double>(tmp[b]) * wa;
}

// Every pixel, one thread per pixel for NB slices.
template <typename T, int NB>
__global__ __launch_bounds__(256) void fpt_kernel(const T* __restrict__ proj,
                                                  T* __restrict__ img,
                                                  const double* __restrict__ weights,
                                                  const double* __restrict__ pos,
                                                  const int* __restrict__ usex, FpProblem P,
                                                  int stride)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int b0 = blockIdx.y * NB;
    const int nb = min(NB, P.nbatch - b0);
    if (i >= P.nx * P.ny) return;
    const int ix = i % P.nx, iy = i / P.nx;
    double sum[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) sum[b] = 0.0;
    for (int a = 0; a < P.na; ++a) fpt_view<T, NB>(proj, weights, pos, usex, P, stride, a, ix, iy,
                                                   b0, nb, sum);
#pragma unroll
    for (int b = 0; b < NB; ++b)
        if (b < nb) img[(static_cast<size_t>(b0 + b) * P.ny + iy) * P.nx + ix] = static_cast<T>(sum[b]);
}

// The border pixels (nx, ny >= 3) after the engine pass, in two steps so
// that neighbouring threads read neighbouring channels: (1) thread = border
// pixel, a chunk of views per workgroup row, FP64 partial sums per (view
// chunk, slice, pixel); (2) the partials added in chunk order.  Chunks of
// kBorderViews views where the partials live in a caller's workspace (the
// allocation-free entry point's fixed scratch), of kBorderViewsAlloc views
// (four times the workgroups) where they are allocated per call.
constexpr int kBorderViews = 16;
constexpr int kBorderViewsAlloc = 4;

template <int NB>
__global__ __launch_bounds__(256) void fpt_border_partial_kernel(
    const float* __restrict__ proj, double* __restrict__ partial,
    const double* __restrict__ weights, const double* __restrict__ pos,
    const int* __restrict__ usex, FpProblem P, int stride, int nborder, int cviews)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int vc = blockIdx.y;
    const int b0 = blockIdx.z * NB;
    const int nb = min(NB, P.nbatch - b0);
    if (i >= nborder) return;
    int ix, iy;
    border_pixel(i, P.nx, P.ny, ix, iy);
    double sum[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) sum[b] = 0.0;
    const int a1 = min(P.na, (vc + 1) * cviews);
    for (int a = vc * cviews; a < a1; ++a)
        fpt_view<float, NB>(proj, weights, pos, usex, P, stride, a, ix, iy, b0, nb, sum);
#pragma unroll
    for (int b = 0; b < NB; ++b)
        if (b < nb) partial[(static_cast<size_t>(vc) * P.nbatch + b0 + b) * nborder + i] = sum[b];
}

__global__ void fpt_border_reduce_kernel(const double* __restrict__ partial,
                                         float* __restrict__ img, FpProblem P, int nborder,
                                         int nchunks)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= nborder) return;
    int ix, iy;
    border_pixel(i, P.nx, P.ny, ix, iy);
    // eight chunks' partials loaded together (chunks past the last re-read
    // it and are not added), then added in chunk order
    const double* q = partial + static_cast<size_t>(b) * nborder + i;
    const size_t cstride = static_cast<size_t>(P.nbatch) * nborder;
    double t = 0.0;
    for (int c = 0; c < nchunks; c += 8) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = q[static_cast<size_t>(min(c + k, nchunks - 1)) * cstride];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (c + k < nchunks) t += v[k];
    }
    img[(static_cast<size_t>(b) * P.ny + iy) * P.nx + ix] = static_cast<float>(t);
}

namespace {

template <typename T, int NB>
int launch_fp(const FpProblem& P, const T* img, const T* imgT, T* proj, const double* sc,
              const double* bo, const int* usex, hipStream_t st)
{
    dim3 grid(div_up(P.ns, 256), P.na, div_up(P.nbatch, NB));
    hipLaunchKernelGGL((fp_kernel<T, NB>), grid, dim3(256), 0, st, img, imgT, proj, sc, bo, usex, P);
    return check_launch("fp_kernel");
}

// Views per workgroup, at most kFpViews: 4 measured 0.78 ms on the
// 8 x 512^2, 720-view scan, 2 0.89 ms (each staged row chunk serves half as
// many rays) and, before the chunk pipeline, 8 1.26 ms (two waves per SIMD;
// 178 VGPRs with it).  Small scans take fewer (fp_views) so that the grid
// still fills the GPU.
constexpr int kFpViews = 4;
constexpr int kFpSlices = 4;
// Two buffers of 2-row chunks with 512-column spans (32 KB of LDS at 4
// slices; 97 VGPRs, four waves per SIMD): 0.775 ms on that scan, against
// 0.84 ms for one buffer of 4-row chunks (no overlap of the DMAs with the
// walk), 0.99 ms for two buffers of 4 rows (64 KB, two workgroups per CU),
// 0.79 ms for two of 1 row and 0.88 ms with the VGPRs held to 95 for five
// waves per SIMD.  Taking the taps of all V views branch-free (rays off the
// image reading a zero pad) so their LDS reads overlap measured 0.84 ms:
// whole waves then no longer skip the views they miss.  Staging rows of a
// slice-interleaved copy (a ray step's taps of four slices in two
// ds_read_b128 instead of four ds_read2st64_b32) measured 1.11 ms.
constexpr int kFpRowsDefault = 2, kFpSpanDefault = 512;

// Columns the rays of one fp_lds_kernel chunk can span: cpb channels at up
// to sqrt(2) ds / pixel per channel, kFpRows rows at up to one column per row
// (in pixel units of the marched axis), and the drift of a ray's column
// across `views` views (at most 4 R |orbit| per view for an image of
// half-diagonal R pixels), plus margins.
bool fp_lds_applies(const ctr_geom& g, int views, int rows = kFpRowsDefault,
                    int span_max = kFpSpanDefault)
{
    const double pmin = std::min(std::fabs(g.dx), std::fabs(g.dy));
    const double pmax = std::max(std::fabs(g.dx), std::fabs(g.dy));
    if (!(pmin > 0) || !(g.ds > 0)) return false;
    const double r = 0.5 * std::hypot(g.nx + 2 * std::fabs(g.offset_x),
                                      g.ny + 2 * std::fabs(g.offset_y)) * pmax / pmin;
    const int cpb = div_up(g.ns, div_up(g.ns, 256));  // channels per workgroup
    const double span = (cpb - 1) * std::sqrt(2.0) * g.ds / pmin + rows * pmax / pmin +
                        4.0 * r * views * std::fabs(g.orbit) + 8.0;
    return span + 2.0 <= span_max;
}

// Views per workgroup for nb slices per workgroup: the most (up to
// kFpViews) that leave at least 3 workgroups per CU (768 of the 1024 the
// chip holds at four waves per SIMD), else 1.  On 8 x 512^2 images with
// 729 channels, 640 to 680 views (960 to 1020 workgroups at 4 views) took
// 0.73 ms at 4 views against 0.76-0.84 ms at 2, when the bar was 1024.
int fp_views(const ctr_geom& g, int nbatch, int nb)
{
    const long long blocks = static_cast<long long>(div_up(g.ns, 256)) * div_up(nbatch, nb);
    for (int v = kFpViews; v > 1; v /= 2)
        if (blocks * div_up(g.na, v) >= 768) return v;
    return 1;
}

template <int NB, int V, int R = kFpRowsDefault, int SPAN = kFpSpanDefault>
int launch_fp_lds(const FpProblem& P, const float* img, const float* imgT, float* proj,
                  const double* sc, const double* bo, const int* usex, hipStream_t st)
{
    // channels split evenly over ceil(ns / 256) workgroups
    const int nblk = div_up(P.ns, 256);
    const int cpb = div_up(P.ns, nblk);
    dim3 grid(nblk, div_up(P.na, V), div_up(P.nbatch, NB));
    hipLaunchKernelGGL((fp_lds_kernel<NB, V, R, SPAN>), grid, dim3(256), 0, st, img, imgT, proj,
                       sc, bo, usex, P, cpb);
    return check_launch("fp_lds_kernel");
}

template <int NB>
int launch_fp_lds_v(int views, const FpProblem& P, const float* img, const float* imgT,
                    float* proj, const double* sc, const double* bo, const int* usex,
                    hipStream_t st)
{
    return views >= 4 ? launch_fp_lds<NB, 4>(P, img, imgT, proj, sc, bo, usex, st)
           : views >= 2 ? launch_fp_lds<NB, 2>(P, img, imgT, proj, sc, bo, usex, st)
                        : launch_fp_lds<NB, 1>(P, img, imgT, proj, sc, bo, usex, st);
}

template <typename T>
int fp_typed(const ctr_geom& g, const T* img, T* proj, int nbatch, const double* sc,
             const double* bo, const int* usex, hipStream_t st)
{
    FpProblem P{static_cast<int>(g.ns), static_cast<int>(g.na), static_cast<int>(g.nx),
                static_cast<int>(g.ny), nbatch, g.dx, g.dy};
    T* imgT = nullptr;
    const size_t n = static_cast<size_t>(nbatch) * g.nx * g.ny;
    CTR_HIP(hipMallocAsync(reinterpret_cast<void**>(&imgT), n * sizeof(T), st));
    hipLaunchKernelGGL(transpose_kernel<T>, dim3(div_up(g.nx, 32), div_up(g.ny, 32), nbatch),
                       dim3(32, 8), 0, st, img, imgT, static_cast<int>(g.ny),
                       static_cast<int>(g.nx));
    int rc = check_launch("transpose_kernel");
    if constexpr (std::is_same<T, float>::value) {
        const int nb = nbatch >= kFpSlices ? kFpSlices : nbatch >= 2 ? 2 : 1;
        const int views = fp_views(g, nbatch, nb);
        if (rc == CTR_OK && fp_lds_applies(g, views)) {
            rc = nb == kFpSlices ? launch_fp_lds_v<kFpSlices>(views, P, img, imgT, proj, sc, bo, usex, st)
                 : nb == 2       ? launch_fp_lds_v<2>(views, P, img, imgT, proj, sc, bo, usex, st)
                                 : launch_fp_lds_v<1>(views, P, img, imgT, proj, sc, bo, usex, st);
            CTR_HIP(hipFreeAsync(imgT, st));
            return rc;
        }
    }
    if (rc == CTR_OK) {
        const int nb = nbatch >= 8 ? 8 : nbatch >= 4 ? 4 : nbatch >= 2 ? 2 : 1;
        switch (nb) {
            case 1: rc = launch_fp<T, 1>(P, img, imgT, proj, sc, bo, usex, st); break;
            case 2: rc = launch_fp<T, 2>(P, img, imgT, proj, sc, bo, usex, st); break;
            case 4: rc = launch_fp<T, 4>(P, img, imgT, proj, sc, bo, usex, st); break;
            default: rc = launch_fp<T, 8>(P, img, imgT, proj, sc, bo, usex, st); break;
        }
    }
    CTR_HIP(hipFreeAsync(imgT, st));
    return rc;
}

template <typename T, int NB>
int launch_fpt(const FpProblem& P, const T* proj, T* img, const double* w, const double* pos,
               const int* usex, int stride, hipStream_t st)
{
    dim3 grid(div_up(static_cast<long long>(P.nx) * P.ny, 256), div_up(P.nbatch, NB));
    hipLaunchKernelGGL((fpt_kernel<T, NB>), grid, dim3(256), 0, st, proj, img, w, pos, usex, P,
                       stride);
    return check_launch("fpt_kernel");
}

// Bytes of FP64 partials the border pass needs (nchunks x nbatch x nborder)
// with chunks of cviews views.
size_t fpt_border_bytes(int nx, int ny, int na, int nbatch, int cviews = kBorderViews)
{
    const size_t nborder = 2 * static_cast<size_t>(nx) + 2 * (static_cast<size_t>(ny) - 2);
    return static_cast<size_t>(div_up(na, cviews)) * nbatch * nborder * sizeof(double);
}

// ws: caller workspace of at least fpt_border_bytes(...) bytes, or null for a
// stream-ordered allocation.  Every partial is written before it is read.
template <int NB>
int launch_fpt_border(const FpProblem& P, const float* proj, float* img, const double* w,
                      const double* pos, const int* usex, int stride, void* ws, hipStream_t st)
{
    const int nborder = 2 * P.nx + 2 * (P.ny - 2);
    const int cviews = ws ? kBorderViews : kBorderViewsAlloc;
    const int nchunks = div_up(P.na, cviews);
    double* partial = static_cast<double*>(ws);
    if (!ws)
        CTR_HIP(hipMallocAsync(reinterpret_cast<void**>(&partial),
                               fpt_border_bytes(P.nx, P.ny, P.na, P.nbatch, cviews), st));
    hipLaunchKernelGGL((fpt_border_partial_kernel<NB>),
                       dim3(div_up(nborder, 256), nchunks, div_up(P.nbatch, NB)), dim3(256), 0, st,
                       proj, partial, w, pos, usex, P, stride, nborder, cviews);
    int rc = check_launch("fpt_border_partial_kernel");
    if (rc == CTR_OK) {
        hipLaunchKernelGGL(fpt_border_reduce_kernel, dim3(div_up(nborder, 256), P.nbatch),
                           dim3(256), 0, st, partial, img, P, nborder, nchunks);
        rc = check_launch("fpt_border_reduce_kernel");
    }
    if (!ws) CTR_HIP(hipFreeAsync(partial, st));
    return rc;
}

template <int NB, int LROW>
int launch_fpt_engine(const ctr_geom& g, const float* proj, float* img, int nbatch,
                      const double* w, const double* pos, const int* usex, int stride,
                      hipStream_t st)
{
    constexpr int TW = 16, TH = 16;
    BpLaunch<FptGeo> L{};
    L.P = FptProblem{static_cast<int>(g.ns), static_cast<int>(g.na), static_cast<int>(g.nx),
                     static_cast<int>(g.ny)};
    L.nbatch = nbatch;
    L.tiles_x = div_up(g.nx, TW);
    L.tiles_y = div_up(g.ny, TH);
    L.lmax = LROW;
    L.vchunk = engine_vchunk<FptAnchor>(NB, LROW);
    L.scale = 1.0;
    L.proj = proj;
    L.img = img;
    const FptTables T{w, pos, usex, stride};
    const size_t lds = kAnchorChunk * sizeof(FptAnchor) +
                       2 * static_cast<size_t>(L.vchunk) * NB * LROW * sizeof(float);
    auto kern = bp_forward_kernel<FptGeo, TW, TH, NB, LROW>;
    if (lds > 64 * 1024)
        CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds)));
    hipLaunchKernelGGL(kern, dim3(super_grid(L.tiles_x, L.tiles_y), div_up(nbatch, NB)),
                       dim3(TileShape<TW, TH>::kBlock), lds, st, T, L);
    return check_launch("bp_forward_kernel<fp transpose>");
}

// The engine path needs |o1| <= 1 for every view (at most two channels per
// pixel-view) and a staged span that fits 64 channels per 16x16 tile:
// |o1| <= max(dx, dy)/ds and the span is at most 15 (|o1| + |o2|) + 3.
bool fpt_engine_applies(const ctr_geom& g)
{
    const double r = std::max(std::fabs(g.dx), std::fabs(g.dy)) / g.ds;
    return g.ds > 0 && r <= 1.0 && 15.0 * 2.0 * r + 4.0 <= 64.0;
}

template <typename T>
int fpt_typed(const ctr_geom& g, const T* proj, T* img, int nbatch, const double* w,
              const double* pos, const int* usex, unsigned int stride, void* ws, hipStream_t st)
{
    if constexpr (std::is_same<T, float>::value) {
        if (fpt_engine_applies(g) && g.nx >= 3 && g.ny >= 3) {
            // interior on the engine, then the border pixels exactly
            const int s = static_cast<int>(stride);
            int rc;
            const int nb = engine_nb(nbatch >= 8 ? 8 : nbatch >= 4 ? 4 : nbatch, g.na, g.ns);
            // the applicability bound keeps a tile-view's span within
            // 30 max(dx,dy)/ds + 5 <= 35 channels: 48-channel rows when the
            // slices are interleaved four to a word, 64 otherwise
            if (nb >= 8) rc = launch_fpt_engine<8, 48>(g, proj, img, nbatch, w, pos, usex, s, st);
            else if (nb >= 4) rc = launch_fpt_engine<4, 48>(g, proj, img, nbatch, w, pos, usex, s, st);
            else if (nb >= 2) rc = launch_fpt_engine<2, 64>(g, proj, img, nbatch, w, pos, usex, s, st);
            else rc = launch_fpt_engine<1, 64>(g, proj, img, nbatch, w, pos, usex, s, st);
            if (rc) return rc;
            FpProblem P{static_cast<int>(g.ns), static_cast<int>(g.na), static_cast<int>(g.nx),
                        static_cast<int>(g.ny), nbatch, g.dx, g.dy};
            return nbatch >= 8 ? launch_fpt_border<8>(P, proj, img, w, pos, usex, s, ws, st)
                               : launch_fpt_border<1>(P, proj, img, w, pos, usex, s, ws, st);
        }
    }
    FpProblem P{static_cast<int>(g.ns), static_cast<int>(g.na), static_cast<int>(g.nx),
                static_cast<int>(g.ny), nbatch, g.dx, g.dy};
    const int nb = nbatch >= 8 ? 8 : nbatch >= 4 ? 4 : nbatch >= 2 ? 2 : 1;
    const int s = static_cast<int>(stride);
    switch (nb) {
        case 1: return launch_fpt<T, 1>(P, proj, img, w, pos, usex, s, st);
        case 2: return launch_fpt<T, 2>(P, proj, img, w, pos, usex, s, st);
        case 4: return launch_fpt<T, 4>(P, proj, img, w, pos, usex, s, st);
        default: return launch_fpt<T, 8>(P, proj, img, w, pos, usex, s, st);
    }
}

}  // namespace

int fp_par(const ctr_geom& g, int dtype, const void* img, void* proj, int nbatch,
           const double* sc, const double* bo, const int* usex, hipStream_t st)
{
    CTR_REQUIRE(g.ns >= 1 && g.na >= 1 && g.nx >= 1 && g.ny >= 1, CTR_ERR_ARG,
                "ray-driven FP: sizes must be positive");
    if (dtype == CTR_F32)
        return fp_typed(g, static_cast<const float*>(img), static_cast<float*>(proj), nbatch, sc,
                        bo, usex, st);
    return fp_typed(g, static_cast<const double*>(img), static_cast<double*>(proj), nbatch, sc, bo,
                    usex, st);
}

int fp_par_t(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
             const double* w, const double* pos, const int* usex, unsigned int pos_stride,
             hipStream_t st)
{
    CTR_REQUIRE(g.ns >= 1 && g.na >= 1 && g.nx >= 1 && g.ny >= 1, CTR_ERR_ARG,
                "ray-driven FP transpose: sizes must be positive");
    if (dtype == CTR_F32)
        return fpt_typed(g, static_cast<const float*>(proj), static_cast<float*>(img), nbatch, w,
                         pos, usex, pos_stride, nullptr, st);
    return fpt_typed(g, static_cast<const double*>(proj), static_cast<double*>(img), nbatch, w,
                     pos, usex, pos_stride, nullptr, st);
}

// ---------------------------------------------- allocation-free FP transpose
// The only device memory the FP transpose uses beyond its inputs and output
// is the border pass's FP64 partials (float32 buffers, geometries on the BP
// engine).  ctr_fp_t_noalloc keeps them in this file-scope array instead of a
// stream-ordered allocation: kFptScratchBytes per device, held until the
// process ends.  Kernels take its device address, which hipGetSymbolAddress
// returns once per device (the name itself is a host-side stand-in).
constexpr size_t kFptScratchBytes = size_t{16} << 20;
static __device__ __align__(256) unsigned char scratch[kFptScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::unordered_map<int, void*> g_scratch_ptr;  // device -> address of `scratch` there

int scratch_address(void** out)
{
    int dev = 0;
    CTR_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void* p = nullptr;
        CTR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        it = g_scratch_ptr.emplace(dev, p).first;
    }
    *out = it->second;
    return CTR_OK;
}
}  // namespace

int fp_par_t_noalloc(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
                     const double* w, const double* pos, const int* usex,
                     unsigned int pos_stride, hipStream_t st)
{
    CTR_REQUIRE(g.ns >= 1 && g.na >= 1 && g.nx >= 1 && g.ny >= 1, CTR_ERR_ARG,
                "ray-driven FP transpose: sizes must be positive");
    if (dtype != CTR_F32)  // the float64 transpose uses no scratch
        return fpt_typed(g, static_cast<const double*>(proj), static_cast<double*>(img), nbatch,
                         w, pos, usex, pos_stride, nullptr, st);
    // only the engine path (with its border pass) uses the scratch
    const bool border = fpt_engine_applies(g) && g.nx >= 3 && g.ny >= 3;
    const size_t need = border ? fpt_border_bytes(static_cast<int>(g.nx), static_cast<int>(g.ny),
                                                  static_cast<int>(g.na), nbatch)
                               : 0;
    if (need > kFptScratchBytes) {
        set_error("ctr_fp_t_noalloc: the border partials need " + std::to_string(need) +
                  " bytes, more than the " + std::to_string(kFptScratchBytes) +
                  "-byte scratch (use ctr_fp_t, or fewer slices per call)");
        return CTR_ERR_SHAPE;
    }
    void* ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc) return rc;
    return fpt_typed(g, static_cast<const float*>(proj), static_cast<float*>(img), nbatch, w, pos,
                     usex, pos_stride, ws, st);
}

}  // namespace ctr
