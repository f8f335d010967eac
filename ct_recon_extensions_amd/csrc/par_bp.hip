// Parallel-beam pixel-driven backprojection for gfx950 and its exact
// transpose.  Same structure as the fan kernels (tile x slice-group
// workgroups, LDS-staged detector spans, FP64 per-(tile, view) anchors, FP32
// pixel offsets); the channel position is linear in the pixel:
//     s = (x cos(theta) + y sin(theta)) / ds + cents    (src/bp_par_2d.cc:66-70)
// so per tile and view  s(i, j) = s_c + i*a + j*b  with the anchor s_c taken
// from the reference's own FP64 tables xcos / ysin and the slopes from their
// differences.  Output scale is the signed orbit (src/bp_par_2d.cc:83).
#include "ctr_common.hpp"
#include "tile_common.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace ctr {

struct alignas(16) ParAnchor {
    float a, b;  // channels per pixel step in x / in y
    float fc;    // anchor channel minus is0
    int is0;     // first staged channel
    int inside;  // whole tile strictly inside the detector
    int pad[3];
};

struct ParProblem {
    double ds, cents;
    int ns, na, nx, ny;
};

__device__ inline ParAnchor par_anchor(const ParProblem& P, const double* __restrict__ xcos,
                                       const double* __restrict__ ysin, int view, int ic, int jc,
                                       int ilo, int ihi, int jlo, int jhi)
{
    const size_t na = P.na;
    const double xc = xcos[ic * na + view], yc = ysin[jc * na + view];
    const double sc = (xc + yc) / P.ds + P.cents;
    const int ic2 = ic + 1 < P.nx ? ic + 1 : ic - 1;
    const int jc2 = jc + 1 < P.ny ? jc + 1 : jc - 1;
    const double a = P.nx > 1 ? (xcos[ic2 * na + view] - xc) / ((ic2 - ic) * P.ds) : 0.0;
    const double b = P.ny > 1 ? (ysin[jc2 * na + view] - yc) / ((jc2 - jc) * P.ds) : 0.0;
    const double lo = fmin(a * ilo, a * ihi) + fmin(b * jlo, b * jhi);
    const double hi = fmax(a * ilo, a * ihi) + fmax(b * jlo, b * jhi);
    const int is0 = static_cast<int>(floor(sc + lo)) - 1;
    ParAnchor A;
    A.a = static_cast<float>(a);
    A.b = static_cast<float>(b);
    A.fc = static_cast<float>(sc - is0);
    A.is0 = is0;
    A.inside = (sc + lo >= 0.25 && sc + hi <= P.ns - 1.25) ? 1 : 0;
    A.pad[0] = A.pad[1] = A.pad[2] = 0;
    return A;
}

struct ParLaunch {
    ParProblem P;
    int nbatch, tiles_x, tiles_y, lmax, vchunk;
    double scale;
    const double* fix_scale;
};

template <int TW, int TH, int NB>
__global__ __launch_bounds__((TileShape<TW, TH>::kBlock)) void par_bp_kernel(
    const float* __restrict__ proj, float* __restrict__ img, const double* __restrict__ xcos,
    const double* __restrict__ ysin, const ParLaunch L)
{
    using TS = TileShape<TW, TH>;
    constexpr int BLOCK = TS::kBlock, PPT = TS::kPPT;
    extern __shared__ __align__(16) unsigned char smem[];
    ParAnchor* anchors = reinterpret_cast<ParAnchor*>(smem);
    float* spans = reinterpret_cast<float*>(smem + kAnchorChunk * sizeof(ParAnchor));
    const ParProblem& P = L.P;
    const int tid = threadIdx.x;
    const TileFrame F(blockIdx.x, L.tiles_x, P.nx, P.ny, TW, TH);
    const int b0 = blockIdx.y * NB;
    const int nb = min(NB, L.nbatch - b0);
    const int lmax = L.lmax;
    const int ilo = F.x0 - F.ic, ihi = F.x0 + TW - 1 - F.ic;
    const int jlo = F.y0 - F.jc, jhi = F.y0 + TH - 1 - F.jc;

    float fi[PPT], fj[PPT];
    double acc64[PPT][NB];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int p = tid + k * BLOCK;
        fi[k] = static_cast<float>(F.x0 + p % TW - F.ic);
        fj[k] = static_cast<float>(F.y0 + p / TW - F.jc);
#pragma unroll
        for (int b = 0; b < NB; ++b) acc64[k][b] = 0.0;
    }

    for (int va = 0; va < P.na; va += kAnchorChunk) {
        const int nva = min(kAnchorChunk, P.na - va);
        __syncthreads();
        if (tid < nva)
            anchors[tid] = par_anchor(P, xcos, ysin, va + tid, F.ic, F.jc, ilo, ihi, jlo, jhi);
        float acc[PPT][NB];
#pragma unroll
        for (int k = 0; k < PPT; ++k)
#pragma unroll
            for (int b = 0; b < NB; ++b) acc[k][b] = 0.0f;
        for (int vc = 0; vc < nva; vc += L.vchunk) {
            const int nv = min(L.vchunk, nva - vc);
            __syncthreads();
            const int rows = nv * NB;
            for (int e = tid; e < rows * lmax; e += BLOCK) {
                const int row = e / lmax, c = e - row * lmax;
                const int t = row / NB, b = row - t * NB;
                const int ch = anchors[vc + t].is0 + c;
                float val = 0.0f;
                if (b < nb && ch >= 0 && ch < P.ns)
                    val = proj[(static_cast<size_t>(b0 + b) * P.na + va + vc + t) * P.ns + ch];
                spans[e] = val;
            }
            __syncthreads();
            for (int t = 0; t < nv; ++t) {
                const ParAnchor A = anchors[vc + t];
                const float* srow = spans + t * NB * lmax;
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    const float f = fmaf(A.a, fi[k], fmaf(A.b, fj[k], A.fc));
                    const float fl = floorf(f);
                    const float fr = f - fl;
                    int idx = static_cast<int>(fl);
                    float w = 1.0f;
                    if (!A.inside) w = on_detector(A.is0 + idx, fr, P.ns) ? 1.0f : 0.0f;
                    idx = min(max(idx, 0), lmax - 2);
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        const float p0 = srow[b * lmax + idx];
                        const float p1 = srow[b * lmax + idx + 1];
                        acc[k][b] = fmaf(w, fmaf(fr, p1 - p0, p0), acc[k][b]);
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < PPT; ++k)
#pragma unroll
            for (int b = 0; b < NB; ++b) acc64[k][b] += static_cast<double>(acc[k][b]);
    }
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int p = tid + k * BLOCK;
        const int ix = F.x0 + p % TW, iy = F.y0 + p / TW;
        if (ix < P.nx && iy < P.ny) {
#pragma unroll
            for (int b = 0; b < NB; ++b)
                if (b < nb)
                    img[(static_cast<size_t>(b0 + b) * P.ny + iy) * P.nx + ix] =
                        static_cast<float>(acc64[k][b] * L.scale);
        }
    }
}

// View-parallel scatter transpose (see fan_bp.hip); weights bit-identical to
// par_bp_kernel through the shared anchors.
template <int TW, int TH, int NB, bool DET>
__global__ __launch_bounds__(256) void par_bpt_kernel(const float* __restrict__ img,
                                                      float* __restrict__ proj,
                                                      const double* __restrict__ xcos,
                                                      const double* __restrict__ ysin,
                                                      const ParLaunch L)
{
    using TS = TileShape<TW, TH>;
    constexpr int BLOCK = TS::kBlock, PPT = TS::kPPT;
    static_assert(BLOCK == 256, "transpose assumes 256-thread tiles");
    using Acc = typename std::conditional<DET, unsigned long long, float>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    ParAnchor* anchors = reinterpret_cast<ParAnchor*>(smem);
    Acc* rows = reinterpret_cast<Acc*>(smem + kAnchorChunk * sizeof(ParAnchor));
    const ParProblem& P = L.P;
    const int tid = threadIdx.x;
    const int view = blockIdx.x;
    const int b0 = blockIdx.y * NB;
    const int nb = min(NB, L.nbatch - b0);
    const int rlen = P.ns + 1;
    double fixs = 1.0;
    if constexpr (DET) fixs = *L.fix_scale;
    for (int e = tid; e < NB * rlen; e += BLOCK) rows[e] = Acc(0);

    const int ntiles = L.tiles_x * L.tiles_y;
    for (int tc = 0; tc < ntiles; tc += kAnchorChunk) {
        const int ntc = min(kAnchorChunk, ntiles - tc);
        __syncthreads();
        if (tid < ntc) {
            const TileFrame F(tc + tid, L.tiles_x, P.nx, P.ny, TW, TH);
            anchors[tid] = par_anchor(P, xcos, ysin, view, F.ic, F.jc, F.x0 - F.ic,
                                      F.x0 + TW - 1 - F.ic, F.y0 - F.jc, F.y0 + TH - 1 - F.jc);
        }
        __syncthreads();
        for (int t = 0; t < ntc; ++t) {
            const ParAnchor A = anchors[t];
            const TileFrame F(tc + t, L.tiles_x, P.nx, P.ny, TW, TH);
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                const int p = tid + k * BLOCK;
                const int ix = F.x0 + p % TW, iy = F.y0 + p / TW;
                if (ix >= P.nx || iy >= P.ny) continue;
                const float f = fmaf(A.a, static_cast<float>(ix - F.ic),
                                     fmaf(A.b, static_cast<float>(iy - F.jc), A.fc));
                const float fl = floorf(f);
                const float fr = f - fl;
                const int s = A.is0 + static_cast<int>(fl);
                if (!on_detector(s, fr, P.ns)) continue;
                const float w0 = 1.0f - fr;
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    if (b >= nb) break;
                    const float v = img[(static_cast<size_t>(b0 + b) * P.ny + iy) * P.nx + ix];
                    if constexpr (DET) {
                        atomicAdd(&rows[b * rlen + s], static_cast<unsigned long long>(
                                      __double2ll_rn(static_cast<double>(w0 * v) * fixs)));
                        atomicAdd(&rows[b * rlen + s + 1], static_cast<unsigned long long>(
                                      __double2ll_rn(static_cast<double>(fr * v) * fixs)));
                    } else {
                        atomicAdd(&rows[b * rlen + s], w0 * v);
                        atomicAdd(&rows[b * rlen + s + 1], fr * v);
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < nb * P.ns; e += BLOCK) {
        const int b = e / P.ns, s = e - b * P.ns;
        double v;
        if constexpr (DET)
            v = static_cast<double>(static_cast<long long>(rows[b * rlen + s])) / fixs;
        else
            v = rows[b * rlen + s];
        proj[(static_cast<size_t>(b0 + b) * P.na + view) * P.ns + s] = static_cast<float>(v * L.scale);
    }
}

// FP64 buffers: one thread per pixel, the reference's FP64 arithmetic.
__global__ void par_bp_f64_kernel(const double* __restrict__ proj, double* __restrict__ img,
                                  const double* __restrict__ xcos,
                                  const double* __restrict__ ysin, ParProblem P, int nbatch,
                                  double scale)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= P.nx * P.ny || b >= nbatch) return;
    const int ix = i % P.nx, iy = i / P.nx;
    const double* pr = proj + static_cast<size_t>(b) * P.na * P.ns;
    double sum = 0.0;
    for (int ia = 0; ia < P.na; ++ia) {
        const double s = (xcos[static_cast<size_t>(ix) * P.na + ia] +
                          ysin[static_cast<size_t>(iy) * P.na + ia]) / P.ds + P.cents;
        if (s >= 0 && s <= P.ns - 1) {
            const double fl = floor(s);
            const double u = s - fl;
            sum += (1 - u) * pr[static_cast<int>(fl) + ia * P.ns] +
                   u * pr[static_cast<int>(ceil(s)) + ia * P.ns];
        }
    }
    img[static_cast<size_t>(b) * P.nx * P.ny + i] = sum * scale;
}

__global__ void par_bpt_f64_kernel(const double* __restrict__ img, double* __restrict__ proj,
                                   const double* __restrict__ xcos,
                                   const double* __restrict__ ysin, ParProblem P, int nbatch)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= P.nx * P.ny || b >= nbatch) return;
    const int ix = i % P.nx, iy = i / P.nx;
    const double v = img[static_cast<size_t>(b) * P.nx * P.ny + i];
    double* out = proj + static_cast<size_t>(b) * P.na * P.ns;
    for (int ia = 0; ia < P.na; ++ia) {
        const double s = (xcos[static_cast<size_t>(ix) * P.na + ia] +
                          ysin[static_cast<size_t>(iy) * P.na + ia]) / P.ds + P.cents;
        if (s >= 0 && s <= P.ns - 1) {
            const double fl = floor(s);
            const double u = s - fl;
            atomicAdd(&out[static_cast<int>(fl) + ia * P.ns], (1 - u) * v);
            atomicAdd(&out[static_cast<int>(ceil(s)) + ia * P.ns], u * v);
        }
    }
}

namespace {

int par_problem(const ctr_geom& g, int TW, int TH, ParProblem& P, int& lmax)
{
    CTR_REQUIRE(g.ns >= 2 && g.na >= 1 && g.nx >= 1 && g.ny >= 1, CTR_ERR_ARG,
                "parallel BP: sizes must be positive (ns >= 2)");
    CTR_REQUIRE(g.ds > 0, CTR_ERR_ARG, "parallel BP: ds must be positive");
    P.ds = g.ds;
    P.cents = cent_s(g);
    P.ns = static_cast<int>(g.ns);
    P.na = static_cast<int>(g.na);
    P.nx = static_cast<int>(g.nx);
    P.ny = static_cast<int>(g.ny);
    lmax = static_cast<int>(std::ceil((TW * std::fabs(g.dx) + TH * std::fabs(g.dy)) / g.ds)) + 4;
    lmax = (lmax + 1) & ~1;
    return CTR_OK;
}

template <int TW, int TH, int NB>
int launch_par_bp(const ParLaunch& L, const float* proj, float* img, const double* xcos,
                  const double* ysin, hipStream_t st)
{
    using TS = TileShape<TW, TH>;
    const size_t lds = kAnchorChunk * sizeof(ParAnchor) +
                       static_cast<size_t>(L.vchunk) * NB * L.lmax * sizeof(float);
    CTR_REQUIRE(lds <= 160 * 1024, CTR_ERR_SHAPE, "parallel BP: detector span too wide for LDS");
    auto kern = par_bp_kernel<TW, TH, NB>;
    if (lds > 64 * 1024)
        CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds)));
    dim3 grid(L.tiles_x * L.tiles_y, div_up(L.nbatch, NB));
    hipLaunchKernelGGL(kern, grid, dim3(TS::kBlock), lds, st, proj, img, xcos, ysin, L);
    return check_launch("par_bp_kernel");
}

}  // namespace

int par_bp(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
           const double* xcos, const double* ysin, hipStream_t st)
{
    constexpr int TW = 16, TH = 16;
    ParLaunch L{};
    int rc = par_problem(g, TW, TH, L.P, L.lmax);
    if (rc) return rc;
    L.nbatch = nbatch;
    L.tiles_x = div_up(g.nx, TW);
    L.tiles_y = div_up(g.ny, TH);
    L.vchunk = 16;
    L.scale = g.orbit;
    if (dtype == CTR_F64) {
        hipLaunchKernelGGL(par_bp_f64_kernel, dim3(div_up(g.nx * g.ny, 256), nbatch), dim3(256),
                           0, st, static_cast<const double*>(proj), static_cast<double*>(img), xcos,
                           ysin, L.P, nbatch, g.orbit);
        return check_launch("par_bp_f64_kernel");
    }
    auto p = static_cast<const float*>(proj);
    auto o = static_cast<float*>(img);
    const int nb = nbatch >= 16 ? 16 : nbatch >= 8 ? 8 : nbatch >= 4 ? 4 : nbatch >= 2 ? 2 : 1;
    switch (nb) {
        case 1: return launch_par_bp<TW, TH, 1>(L, p, o, xcos, ysin, st);
        case 2: return launch_par_bp<TW, TH, 2>(L, p, o, xcos, ysin, st);
        case 4: return launch_par_bp<TW, TH, 4>(L, p, o, xcos, ysin, st);
        case 8: return launch_par_bp<TW, TH, 8>(L, p, o, xcos, ysin, st);
        default: return launch_par_bp<TW, TH, 16>(L, p, o, xcos, ysin, st);
    }
}

int par_bpt(const ctr_geom& g, int dtype, const void* img, void* proj, int nbatch,
            const double* xcos, const double* ysin, bool det, hipStream_t st)
{
    constexpr int TW = 16, TH = 16;
    ParLaunch L{};
    int rc = par_problem(g, TW, TH, L.P, L.lmax);
    if (rc) return rc;
    L.nbatch = nbatch;
    L.tiles_x = div_up(g.nx, TW);
    L.tiles_y = div_up(g.ny, TH);
    L.scale = g.orbit;
    if (dtype == CTR_F64) {
        const size_t n = static_cast<size_t>(nbatch) * g.na * g.ns;
        CTR_HIP(hipMemsetAsync(proj, 0, n * sizeof(double), st));
        hipLaunchKernelGGL(par_bpt_f64_kernel, dim3(div_up(g.nx * g.ny, 256), nbatch), dim3(256),
                           0, st, static_cast<const double*>(img), static_cast<double*>(proj),
                           xcos, ysin, L.P, nbatch);
        rc = check_launch("par_bpt_f64_kernel");
        if (rc) return rc;
        hipLaunchKernelGGL(scale_f64_kernel, dim3(std::min<size_t>(4096, (n + 255) / 256)),
                           dim3(256), 0, st, static_cast<double*>(proj), n, g.orbit);
        return check_launch("scale_f64_kernel");
    }
    auto im = static_cast<const float*>(img);
    auto pr = static_cast<float*>(proj);
    const size_t rlen = g.ns + 1;
    if (det) {
        float* partial = nullptr;
        double* fs = nullptr;
        CTR_HIP(hipMallocAsync(reinterpret_cast<void**>(&partial), kMaxBlocks * sizeof(float), st));
        CTR_HIP(hipMallocAsync(reinterpret_cast<void**>(&fs), sizeof(double), st));
        const size_t n = static_cast<size_t>(nbatch) * g.nx * g.ny;
        const int nblk = static_cast<int>(std::min<size_t>(kMaxBlocks, (n + 255) / 256));
        hipLaunchKernelGGL(abs_max_partial_kernel, dim3(nblk), dim3(256), 0, st, im, n, partial);
        hipLaunchKernelGGL(fix_scale_kernel, dim3(1), dim3(64), 0, st, partial, nblk,
                           2.0 * g.nx * g.ny, fs);
        L.fix_scale = fs;
        constexpr int NB = 4;
        const size_t lds = kAnchorChunk * sizeof(ParAnchor) + NB * rlen * 8;
        CTR_REQUIRE(lds <= 160 * 1024, CTR_ERR_SHAPE, "parallel BP transpose: ns too large");
        auto kern = par_bpt_kernel<TW, TH, NB, true>;
        if (lds > 64 * 1024)
            CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
        hipLaunchKernelGGL(kern, dim3(g.na, div_up(nbatch, NB)), dim3(256), lds, st, im, pr, xcos,
                           ysin, L);
        rc = check_launch("par_bpt_kernel<det>");
        CTR_HIP(hipFreeAsync(partial, st));
        CTR_HIP(hipFreeAsync(fs, st));
        return rc;
    }
    const int nb = nbatch >= 8 ? 8 : nbatch >= 4 ? 4 : nbatch >= 2 ? 2 : 1;
    const size_t lds = kAnchorChunk * sizeof(ParAnchor) + static_cast<size_t>(nb) * rlen * 4;
    CTR_REQUIRE(lds <= 160 * 1024, CTR_ERR_SHAPE, "parallel BP transpose: ns too large");
    dim3 grid(g.na, div_up(nbatch, nb));
#define CTR_PART(NB_)                                                                       \
    {                                                                                       \
        auto kern = par_bpt_kernel<TW, TH, NB_, false>;                                     \
        if (lds > 64 * 1024)                                                                \
            CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),                \
                                        hipFuncAttributeMaxDynamicSharedMemorySize,         \
                                        static_cast<int>(lds)));                            \
        hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, im, pr, xcos, ysin, L);          \
        return check_launch("par_bpt_kernel");                                              \
    }
    switch (nb) {
        case 1: CTR_PART(1)
        case 2: CTR_PART(2)
        case 4: CTR_PART(4)
        default: CTR_PART(8)
    }
#undef CTR_PART
}

}  // namespace ctr
