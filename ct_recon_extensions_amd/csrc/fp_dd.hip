// Distance-driven parallel FP (src/fp_par_2d.cc:396-543; the TF FP op's
// method "disdriven") and the gradient the reference defines for it
// (src/fp_par_2d.cc:548-655), for gfx950.  Both are bit-identical to the
// reference's CPU path in float32 and float64: the reference's operation
// order and its float / double mix are kept, with contraction off, and the
// work is split only where the reference's sums are independent (one thread
// per (channel, view) for the FP, one per pixel for the gradient).
//
// FP, per channel and per row (x-branch; per column for the y-branch): the
// channel's footprint [left, right] (centre -+ half a channel width, in pixel
// units) averages the cell values (img[v] + img[v+1]) / 2 of the cells v it
// overlaps, weighted by the overlap; row averages are summed and scaled by
// |dy/cos| (|dx/sin|).  The reference's cell index is unsigned: a row whose
// footprint starts left of the first cell turns floor(left) < 0 into a huge
// index and adds nothing on the x86-64 build; that rule is kept.  The y-branch
// reads a transposed copy of the image so that a footprint's cells are
// contiguous.
//
// Gradient, per pixel and view: the span [mid - o1, mid + o1] of the pixel on
// the detector (one-sided at the first and last pixel along the branch's
// axis) and the channels whose cells it overlaps, each adding
// overlap * proj / 2, times the view's weight.  ceil(left - 1/2) < 0 adds
// nothing (the reference's unsigned index again); channels past ns - 1, which
// the reference reads beyond the row, are not read.  This pair is not an
// exact transpose (the reference's gradient approximates one); it is the
// pair the reference's ops use.
#include "ctr_common.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

#pragma clang fp contract(off)

namespace ctr {
namespace {

struct DdProblem {
    int ns, na, nx, ny, nbatch;
    double dx, dy;
};

template <typename T>
__global__ void dd_transpose_kernel(const T* __restrict__ src, T* __restrict__ dst, int ny, int nx)
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

// One thread per (channel, view) and NB slices, a wavefront per 64 channels
// of a view and 4 views per workgroup; data[b][u][v] is the image
// (x-branch: u = iy, v = ix) or its transpose (y-branch: u = ix, v = iy).
constexpr int kDdViews = 4;

template <typename T, int NB>
__global__ __launch_bounds__(64 * kDdViews) void fp_dd_kernel(const T* __restrict__ img,
                                                   const T* __restrict__ imgT,
                                                   T* __restrict__ proj,
                                                   const double* __restrict__ sc,
                                                   const double* __restrict__ bo,
                                                   const int* __restrict__ usex, DdProblem P)
{
    const int is = blockIdx.x * 64 + threadIdx.x;
    const int ia = blockIdx.y * kDdViews + threadIdx.y, b0 = blockIdx.z * NB;
    if (is >= P.ns || ia >= P.na) return;
    const int nb = min(NB, P.nbatch - b0);
    const bool bx = usex[ia] != 0;
    const double dw = bx ? fabs(P.dy / sc[2 * ia + 1]) : fabs(P.dx / sc[2 * ia]);
    const int nu = bx ? P.ny : P.nx, nv = bx ? P.nx : P.ny;
    const T* data = (bx ? img : imgT) + static_cast<size_t>(b0) * P.nx * P.ny;
    const size_t slice = static_cast<size_t>(P.nx) * P.ny;
    const double* b = bo + 3 * (static_cast<size_t>(ia) * P.ns + is);
    double left = b[0] - b[2], right = b[0] + b[2];
    const double off = b[1];
    if (left > right) {
        const double t = left;
        left = right;
        right = t;
    }
    T sum[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) sum[k] = T(0);
    for (int u = 0; u < nu; ++u) {
        T tsum[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) tsum[k] = T(0);
        T lsum = T(0);
        const double fl = floor(left);
        // cells past nv - 2 are skipped by the reference, so the loop stops
        // there (the same sums, and a bounded trip count for any geometry)
        if (fl >= 0 && fl < nv - 1) {
            const double vend = fmin(ceil(right), static_cast<double>(nv - 1));
            const T* row = data + static_cast<size_t>(u) * nv;
            for (unsigned int v = static_cast<unsigned int>(fl); static_cast<double>(v) < vend; ++v) {
                const double len = fmin(static_cast<double>(v + 1), right) -
                                   fmax(static_cast<double>(v), left);
                // every slice's cell pair loaded first (a slice past the
                // batch re-reads the last one and is not added)
                T c0[NB], c1[NB];
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const T* q = row + static_cast<size_t>(min(k, nb - 1)) * slice + v;
                    c0[k] = q[0];
                    c1[k] = q[1];
                }
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    if (k < nb) {
                        const T pair = c0[k] + c1[k];
                        tsum[k] = static_cast<T>(static_cast<double>(tsum[k]) +
                                                 len * static_cast<double>(pair) / 2);
                    }
                }
                lsum = static_cast<T>(static_cast<double>(lsum) + len);
            }
        }
        if (lsum > 0) {
            if constexpr (std::is_same<T, float>::value) {
                // The float quotients tsum / lsum, correctly rounded, from one
                // FP64 reciprocal: q = tsum * RN(1 / lsum) in FP64 is within
                // 2^-52 (relative) of the exact quotient, and a quotient of two
                // floats is never an exact float midpoint and lies at least
                // 2^-49 (relative) from every one (the midpoint has 25
                // significant bits and lsum 24, so tsum - m lsum is a nonzero
                // multiple of 2^-49 |m lsum|): rounding q to float gives the
                // float division's result.
                const double r = 1.0 / static_cast<double>(lsum);
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    sum[k] = sum[k] + static_cast<float>(static_cast<double>(tsum[k]) * r);
            } else {
#pragma unroll
                for (int k = 0; k < NB; ++k) sum[k] = sum[k] + tsum[k] / lsum;
            }
        }
        left += off;
        right += off;
    }
#pragma unroll
    for (int k = 0; k < NB; ++k)
        if (k < nb)
            proj[(static_cast<size_t>(b0 + k) * P.na + ia) * P.ns + is] =
                static_cast<T>(static_cast<double>(sum[k]) * dw);
}

// One thread per pixel and NB slices; pos[view][stride][2] as ctr_fp_grad_tables.
template <typename T, int NB>
__global__ __launch_bounds__(256) void fp_dd_grad_kernel(const T* __restrict__ proj,
                                                         T* __restrict__ img,
                                                         const double* __restrict__ weights,
                                                         const double* __restrict__ pos,
                                                         const int* __restrict__ usex, int stride,
                                                         DdProblem P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nx * P.ny) return;
    const int ix = i % P.nx, iy = i / P.nx;
    const int b0 = blockIdx.y * NB, nb = min(NB, P.nbatch - b0);
    const size_t sslice = static_cast<size_t>(P.na) * P.ns;
    const T* pr = proj + static_cast<size_t>(b0) * sslice;
    double sum[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) sum[k] = 0.0;
    for (int ia = 0; ia < P.na; ++ia) {
        const bool bx = usex[ia] != 0;
        const double* pp = pos + 2 * (static_cast<size_t>(ia) * stride + (bx ? iy : ix));
        const double begin = pp[0], off = pp[1];
        double mid, left, right;
        if (bx) {
            mid = begin + off * ix;
            left = ix == 0 ? mid : mid - off;
            right = ix == P.nx - 1 ? mid : mid + off;
        } else {
            mid = begin + off * iy;
            left = iy == 0 ? mid : mid - off;
            right = iy == P.ny - 1 ? mid : mid + off;
        }
        if (left > right) {
            const double t = left;
            left = right;
            right = t;
        }
        double tmp[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) tmp[k] = 0.0;
        const double c0 = ceil(left - 0.5);
        if (c0 >= 0 && c0 < P.ns) {
            const double cend = fmin(floor(right + 0.5), static_cast<double>(P.ns - 1));
            const T* row = pr + static_cast<size_t>(ia) * P.ns;
            for (unsigned int is = static_cast<unsigned int>(c0); static_cast<double>(is) <= cend;
                 ++is) {
                const double len = fmin(static_cast<double>(is) + 0.5, right) -
                                   fmax(static_cast<double>(is) - 0.5, left);
                // every slice's value loaded first (a slice past the batch
                // re-reads the last one and is not added)
                T val[NB];
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    val[k] = row[static_cast<size_t>(min(k, nb - 1)) * sslice + is];
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (k < nb) tmp[k] += len * static_cast<double>(val[k]) / 2;
            }
        }
        const double w = weights[ia];
#pragma unroll
        for (int k = 0; k < NB; ++k) sum[k] += tmp[k] * w;
    }
#pragma unroll
    for (int k = 0; k < NB; ++k)
        if (k < nb) img[static_cast<size_t>(b0 + k) * P.nx * P.ny + i] = static_cast<T>(sum[k]);
}

template <typename T, int NB>
int launch_fp_dd(const DdProblem& P, const T* img, const T* imgT, T* proj, const double* sc,
                 const double* bo, const int* usex, hipStream_t st)
{
    dim3 grid(div_up(P.ns, 64), div_up(P.na, kDdViews), div_up(P.nbatch, NB));
    hipLaunchKernelGGL((fp_dd_kernel<T, NB>), grid, dim3(64, kDdViews), 0, st, img, imgT, proj,
                       sc, bo, usex, P);
    return check_launch("fp_dd_kernel");
}

template <typename T>
int fp_dd_typed(const ctr_geom& g, const T* img, T* proj, int nbatch, const double* sc,
                const double* bo, const int* usex, hipStream_t st)
{
    const DdProblem P{static_cast<int>(g.ns), static_cast<int>(g.na), static_cast<int>(g.nx),
                      static_cast<int>(g.ny), nbatch, g.dx, g.dy};
    T* imgT = nullptr;
    const size_t n = static_cast<size_t>(nbatch) * g.nx * g.ny;
    CTR_HIP(hipMallocAsync(reinterpret_cast<void**>(&imgT), n * sizeof(T), st));
    hipLaunchKernelGGL(dd_transpose_kernel<T>, dim3(div_up(g.nx, 32), div_up(g.ny, 32), nbatch),
                       dim3(32, 8), 0, st, img, imgT, static_cast<int>(g.ny),
                       static_cast<int>(g.nx));
    int rc = check_launch("dd_transpose_kernel");
    if (rc == CTR_OK)
        rc = nbatch >= 4 ? launch_fp_dd<T, 4>(P, img, imgT, proj, sc, bo, usex, st)
             : nbatch >= 2 ? launch_fp_dd<T, 2>(P, img, imgT, proj, sc, bo, usex, st)
                           : launch_fp_dd<T, 1>(P, img, imgT, proj, sc, bo, usex, st);
    CTR_HIP(hipFreeAsync(imgT, st));
    return rc;
}

template <typename T, int NB>
int launch_fp_dd_grad(const DdProblem& P, const T* proj, T* img, const double* w,
                      const double* pos, const int* usex, int stride, hipStream_t st)
{
    dim3 grid(div_up(static_cast<long long>(P.nx) * P.ny, 256), div_up(P.nbatch, NB));
    hipLaunchKernelGGL((fp_dd_grad_kernel<T, NB>), grid, dim3(256), 0, st, proj, img, w, pos,
                       usex, stride, P);
    return check_launch("fp_dd_grad_kernel");
}

template <typename T>
int fp_dd_grad_typed(const ctr_geom& g, const T* proj, T* img, int nbatch, const double* w,
                     const double* pos, const int* usex, unsigned int stride, hipStream_t st)
{
    const DdProblem P{static_cast<int>(g.ns), static_cast<int>(g.na), static_cast<int>(g.nx),
                      static_cast<int>(g.ny), nbatch, g.dx, g.dy};
    const int s = static_cast<int>(stride);
    return nbatch >= 4 ? launch_fp_dd_grad<T, 4>(P, proj, img, w, pos, usex, s, st)
           : nbatch >= 2 ? launch_fp_dd_grad<T, 2>(P, proj, img, w, pos, usex, s, st)
                         : launch_fp_dd_grad<T, 1>(P, proj, img, w, pos, usex, s, st);
}

}  // namespace

int fp_dd(const ctr_geom& g, int dtype, const void* img, void* proj, int nbatch,
          const double* sc, const double* bo, const int* usex, hipStream_t st)
{
    CTR_REQUIRE(g.ns >= 1 && g.na >= 1 && g.nx >= 1 && g.ny >= 1, CTR_ERR_ARG,
                "distance-driven FP: sizes must be positive");
    if (dtype == CTR_F32)
        return fp_dd_typed(g, static_cast<const float*>(img), static_cast<float*>(proj), nbatch,
                           sc, bo, usex, st);
    return fp_dd_typed(g, static_cast<const double*>(img), static_cast<double*>(proj), nbatch, sc,
                       bo, usex, st);
}

int fp_dd_grad(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
               const double* w, const double* pos, const int* usex, unsigned int pos_stride,
               hipStream_t st)
{
    CTR_REQUIRE(g.ns >= 1 && g.na >= 1 && g.nx >= 1 && g.ny >= 1, CTR_ERR_ARG,
                "distance-driven FP gradient: sizes must be positive");
    if (dtype == CTR_F32)
        return fp_dd_grad_typed(g, static_cast<const float*>(proj), static_cast<float*>(img),
                                nbatch, w, pos, usex, pos_stride, st);
    return fp_dd_grad_typed(g, static_cast<const double*>(proj), static_cast<double*>(img),
                            nbatch, w, pos, usex, pos_stride, st);
}

}  // namespace ctr
