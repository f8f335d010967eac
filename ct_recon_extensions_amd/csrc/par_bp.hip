// Parallel-beam pixel-driven backprojection for gfx950 and its exact
// transpose: the BP engine (bp_engine.hpp) with a parallel geometry policy
// (tile x slice-group workgroups, LDS-staged detector spans, FP64
// per-(tile, view) anchors, FP32 pixel offsets); the channel position is
// linear in the pixel:
//     s = (x cos(theta) + y sin(theta)) / ds + cents    (src/bp_par_2d.cc:66-70)
// so per tile and view  s(i, j) = s_c + i*a + j*b  with the anchor s_c taken
// from the reference's own FP64 tables xcos / ysin and the slopes from their
// differences.  Output scale is the signed orbit (src/bp_par_2d.cc:83).
#include "bp_engine.hpp"
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
    int nch;     // channels from is0 the tile's taps can touch (one of margin)
    int pad[2];
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
    A.nch = static_cast<int>(floor(sc + hi)) + 3 - is0;
    A.pad[0] = A.pad[1] = 0;
    return A;
}

struct ParTables {
    const double* xcos;  // [nx][na]  (src/bp_par_2d.cc:33-39)
    const double* ysin;  // [ny][na]
};

// Policy for the BP engine: w = 1 and s linear in the pixel offsets.
struct ParGeo {
    using Problem = ParProblem;
    using Tables = ParTables;
    using Anchor = ParAnchor;
    __device__ static Anchor anchor(const Problem& P, const Tables& T, const TileFrame& F,
                                    int view, int ilo, int ihi, int jlo, int jhi)
    {
        return par_anchor(P, T.xcos, T.ysin, view, F.ic, F.jc, ilo, ihi, jlo, jhi);
    }
    __device__ __forceinline__ static void pixel(const Problem&, const Anchor& A, float fi,
                                                 float fj, float& f, float& w)
    {
        f = fmaf(A.a, fi, fmaf(A.b, fj, A.fc));
        w = 1.0f;
    }
};

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
    // span of a tile: |a|(TW-1) + |b|(TH-1) <= hypot(TW dx, TH dy) / ds channels
    lmax = static_cast<int>(std::ceil(std::hypot(TW * g.dx, TH * g.dy) / g.ds)) + 5;
    lmax = (lmax + 1) & ~1;
    return CTR_OK;
}

template <int TW, int TH, int NB, int LROW>
int launch_par_bp_l(BpLaunch<ParGeo> L, const ParTables& T, hipStream_t st)
{
    using TS = TileShape<TW, TH>;
    if (L.vchunk <= 0) L.vchunk = engine_vchunk<ParAnchor>(NB, LROW);
    L.vchunk = std::max(1, std::min(L.vchunk, (150 * 1024 / 8) / (NB * LROW)));
    const size_t lds = kAnchorChunk * sizeof(ParAnchor) +
                       2 * static_cast<size_t>(L.vchunk) * NB * LROW * sizeof(float);
    CTR_REQUIRE(lds <= 160 * 1024, CTR_ERR_SHAPE, "parallel BP: span staging exceeds LDS");
    auto kern = bp_forward_kernel<ParGeo, TW, TH, NB, LROW>;
    if (lds > 64 * 1024)
        CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds)));
    dim3 grid(super_grid(L.tiles_x, L.tiles_y), div_up(L.nbatch, NB));
    hipLaunchKernelGGL(kern, grid, dim3(TS::kBlock), lds, st, T, L);
    return check_launch("bp_forward_kernel<par>");
}

template <int TW, int TH, int NB>
int launch_par_bp(const BpLaunch<ParGeo>& L, const ParTables& T, hipStream_t st)
{
    if constexpr (NB % 4 == 0)
        if (L.lmax <= 48) return launch_par_bp_l<TW, TH, NB, 48>(L, T, st);
    if (L.lmax <= 64) return launch_par_bp_l<TW, TH, NB, 64>(L, T, st);
    if (L.lmax <= 128) return launch_par_bp_l<TW, TH, NB, 128>(L, T, st);
    if (L.lmax <= 256) return launch_par_bp_l<TW, TH, NB, 256>(L, T, st);
    if (L.lmax <= 512) return launch_par_bp_l<TW, TH, NB, 512>(L, T, st);
    set_error("parallel BP: a tile spans more than 512 detector channels (ds << pixel size)");
    return CTR_ERR_SHAPE;
}

template <int NB, bool DET>
int launch_par_bpt(const BptLaunch<ParGeo>& L, int ns, int na, const float* img, float* proj,
                   const ParTables& T, hipStream_t st)
{
    const size_t lds = kAnchorChunk * sizeof(ParAnchor) + static_cast<size_t>(NB) * (ns + 1) *
                       (DET ? 8 : 4);
    CTR_REQUIRE(lds <= 160 * 1024, CTR_ERR_SHAPE, "parallel BP transpose: ns too large for LDS");
    auto kern = bp_transpose_kernel<ParGeo, 16, 16, NB, DET>;
    if (lds > 64 * 1024)
        CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds)));
    hipLaunchKernelGGL(kern, dim3(na, div_up(L.nbatch, NB)), dim3(256), lds, st, img, proj, T, L);
    return check_launch("bp_transpose_kernel<par>");
}

}  // namespace

int par_bp(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
           const double* xcos, const double* ysin, hipStream_t st)
{
    constexpr int TW = 16, TH = 16;
    BpLaunch<ParGeo> L{};
    int rc = par_problem(g, TW, TH, L.P, L.lmax);
    if (rc) return rc;
    L.nbatch = nbatch;
    L.tiles_x = div_up(g.nx, TW);
    L.tiles_y = div_up(g.ny, TH);
    L.vchunk = 0;  // engine_vchunk
    L.scale = g.orbit;
    if (dtype == CTR_F64) {
        hipLaunchKernelGGL(par_bp_f64_kernel, dim3(div_up(g.nx * g.ny, 256), nbatch), dim3(256),
                           0, st, static_cast<const double*>(proj), static_cast<double*>(img), xcos,
                           ysin, L.P, nbatch, g.orbit);
        return check_launch("par_bp_f64_kernel");
    }
    const ParTables T{xcos, ysin};
    L.proj = static_cast<const float*>(proj);
    L.img = static_cast<float*>(img);
    const int nb = engine_nb(
        nbatch >= 16 ? 16 : nbatch >= 8 ? 8 : nbatch >= 4 ? 4 : nbatch >= 2 ? 2 : 1, g.na, g.ns);
    switch (nb) {
        case 1: return launch_par_bp<TW, TH, 1>(L, T, st);
        case 2: return launch_par_bp<TW, TH, 2>(L, T, st);
        case 4: return launch_par_bp<TW, TH, 4>(L, T, st);
        case 8: return launch_par_bp<TW, TH, 8>(L, T, st);
        default: return launch_par_bp<TW, TH, 16>(L, T, st);
    }
}

int par_bpt(const ctr_geom& g, int dtype, const void* img, void* proj, int nbatch,
            const double* xcos, const double* ysin, bool det, hipStream_t st)
{
    constexpr int TW = 16, TH = 16;
    BptLaunch<ParGeo> L{};
    int lmax = 0;
    int rc = par_problem(g, TW, TH, L.P, lmax);
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
    const ParTables T{xcos, ysin};
    auto im = static_cast<const float*>(img);
    auto pr = static_cast<float*>(proj);
    const int ns = static_cast<int>(g.ns), na = static_cast<int>(g.na);
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
        rc = check_launch("fix_scale_kernel");
        if (rc == CTR_OK) rc = launch_par_bpt<4, true>(L, ns, na, im, pr, T, st);
        CTR_HIP(hipFreeAsync(partial, st));
        CTR_HIP(hipFreeAsync(fs, st));
        return rc;
    }
    const int nb = nbatch >= 8 ? 8 : nbatch >= 4 ? 4 : nbatch >= 2 ? 2 : 1;
    switch (nb) {
        case 1: return launch_par_bpt<1, false>(L, ns, na, im, pr, T, st);
        case 2: return launch_par_bpt<2, false>(L, ns, na, im, pr, T, st);
        case 4: return launch_par_bpt<4, false>(L, ns, na, im, pr, T, st);
        default: return launch_par_bpt<8, false>(L, ns, na, im, pr, T, st);
    }
}

}  // namespace ctr
