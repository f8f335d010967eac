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
#include "ctr_common.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

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
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (b < nb) {
                    const Acc p0 = row[(b0 + b) * slice + i1];
                    const Acc p1 = row[(b0 + b) * slice + i2];
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

template <typename T, int NB>
__global__ __launch_bounds__(256) void fpt_kernel(const T* __restrict__ proj,
                                                  T* __restrict__ img,
                                                  const double* __restrict__ weights,
                                                  const double* __restrict__ pos,
                                                  const int* __restrict__ usex, FpProblem P,
                                                  int stride)
{
    using Acc = typename std::conditional<sizeof(T) == 8, double, float>::type;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int b0 = blockIdx.y * NB;
    const int nb = min(NB, P.nbatch - b0);
    if (i >= P.nx * P.ny) return;
    const int ix = i % P.nx, iy = i / P.nx;
    const size_t pslice = static_cast<size_t>(P.na) * P.ns;
    double sum[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) sum[b] = 0.0;
    for (int a = 0; a < P.na; ++a) {
        double mid, left, right, o1;
        if (usex[a]) {
            const double* pp = pos + 2 * (static_cast<size_t>(a) * stride + iy);
            o1 = pp[1];
            mid = pp[0] + o1 * ix;
            left = ix == 0 ? mid : mid - o1;
            right = ix == P.nx - 1 ? mid : mid + o1;
        } else {
            const double* pp = pos + 2 * (static_cast<size_t>(a) * stride + ix);
            o1 = pp[1];
            mid = pp[0] + o1 * iy;
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
#pragma unroll
            for (int b = 0; b < NB; ++b)
                if (b < nb) tmp[b] += w * static_cast<Acc>(pr[(b0 + b) * pslice]);
        }
        const double wa = weights[a];
#pragma unroll
        for (int b = 0; b < NB; ++b) sum[b] += static_cast<double>(tmp[b]) * wa;
    }
#pragma unroll
    for (int b = 0; b < NB; ++b)
        if (b < nb) img[(static_cast<size_t>(b0 + b) * P.ny + iy) * P.nx + ix] = static_cast<T>(sum[b]);
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

template <typename T>
int fpt_typed(const ctr_geom& g, const T* proj, T* img, int nbatch, const double* w,
              const double* pos, const int* usex, unsigned int stride, hipStream_t st)
{
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
                         pos, usex, pos_stride, st);
    return fpt_typed(g, static_cast<const double*>(proj), static_cast<double*>(img), nbatch, w,
                     pos, usex, pos_stride, st);
}

}  // namespace ctr
