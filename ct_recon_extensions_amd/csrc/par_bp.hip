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
#include "bp_gather.hpp"
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
    // Level line s = s_star in pixel-index coordinates: the tables are linear
    // in the index (xcos[ix][view] = x0 cos + ix dx cos), so
    // L(ix, iy) = pi ix + qj iy + rr = ds (s - s_star) >= 0 where s >= s_star.
    __device__ static void boundary(const Problem& P, const Tables& T, int view, double s_star,
                                    double& pi, double& qj, double& rr)
    {
        const size_t na = P.na;
        const double xc0 = T.xcos[view], ys0 = T.ysin[view];
        pi = P.nx > 1 ? (T.xcos[(P.nx - 1) * na + view] - xc0) / (P.nx - 1) : 0.0;
        qj = P.ny > 1 ? (T.ysin[(P.ny - 1) * na + view] - ys0) / (P.ny - 1) : 0.0;
        rr = xc0 + ys0 - (s_star - P.cents) * P.ds;
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

// FP64 transpose in gather form (float64 buffers, both modes): one thread
// per (channel, view) and up to four slices sums, row by row and in
// increasing ix, the pixels whose position s falls in [c-1, c+1].  Along a
// row s is affine in ix (xcos[ix][view] ~ x0 + ix * slope), so the
// candidates form an interval, found with 1/64 channel and one pixel of
// margin and then tested with the forward kernel's FP64 expressions.  Fixed
// summation order: bitwise repeatable without atomics or a cleared output.
constexpr int kParGatherSlices = 4;

__global__ void par_bpt_f64_gather_kernel(const double* __restrict__ img,
                                          double* __restrict__ proj,
                                          const double* __restrict__ xcos,
                                          const double* __restrict__ ysin, ParProblem P,
                                          int nbatch, double scale)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int ia = blockIdx.y, b0 = blockIdx.z * kParGatherSlices;
    if (c >= P.ns) return;
    const int nb = min(kParGatherSlices, nbatch - b0);
    const size_t na = P.na;
    const double x0 = xcos[ia];
    const double slope = P.nx > 1 ? (xcos[(P.nx - 1) * na + ia] - x0) / (P.nx - 1) : 0.0;
    double acc[kParGatherSlices] = {0.0, 0.0, 0.0, 0.0};
    for (int iy = 0; iy < P.ny; ++iy) {
        const double ys = ysin[iy * na + ia];
        // (c -+ (1 + 1/64) - cents) ds - ysin <= xcos <= ...
        const double lo_v = (c - 1.015625 - P.cents) * P.ds - ys;
        const double hi_v = (c + 1.015625 - P.cents) * P.ds - ys;
        double lo = -1.0, hi = P.nx;
        if (slope > 0.0) {
            lo = fmax(lo, (lo_v - x0) / slope);
            hi = fmin(hi, (hi_v - x0) / slope);
        } else if (slope < 0.0) {
            lo = fmax(lo, (hi_v - x0) / slope);
            hi = fmin(hi, (lo_v - x0) / slope);
        } else if (!(x0 >= lo_v && x0 <= hi_v)) {
            continue;
        }
        if (!(lo <= hi)) continue;
        const int ixl = max(0, static_cast<int>(floor(fmax(lo, -2.0))) - 1);
        const int ixh = min(P.nx - 1, static_cast<int>(ceil(fmin(hi, P.nx + 2.0))) + 1);
        for (int ix = ixl; ix <= ixh; ++ix) {
            const double s = (xcos[ix * na + ia] + ys) / P.ds + P.cents;
            if (!(s >= 0 && s <= P.ns - 1)) continue;
            const double fl = floor(s);
            const double u = s - fl;
            const int is1 = static_cast<int>(fl), is2 = static_cast<int>(ceil(s));
            const double wc = (is1 == c ? 1 - u : 0.0) + (is2 == c ? u : 0.0);
            if (wc == 0.0) continue;
            const size_t pix = static_cast<size_t>(iy) * P.nx + ix;
            // every slice's pixel loaded first (a slice past the batch
            // re-reads the last one and is not added)
            double pv[kParGatherSlices];
#pragma unroll
            for (int b = 0; b < kParGatherSlices; ++b)
                pv[b] = img[static_cast<size_t>(b0 + min(b, nb - 1)) * P.nx * P.ny + pix];
#pragma unroll
            for (int b = 0; b < kParGatherSlices; ++b)
                if (b < nb) acc[b] += wc * pv[b];
        }
    }
#pragma unroll
    for (int b = 0; b < kParGatherSlices; ++b)
        if (b < nb) proj[(static_cast<size_t>(b0 + b) * P.na + ia) * P.ns + c] = acc[b] * scale;
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
            const double* xcos, const double* ysin, bool det, void* ws, hipStream_t st)
{
    ParProblem P{};
    int lmax = 0;
    int rc = par_problem(g, 16, 16, P, lmax);
    if (rc) return rc;
    // Both forms below sum in a fixed order, so `det` needs nothing extra.
    (void)det;
    if (dtype == CTR_F64) {
        hipLaunchKernelGGL(par_bpt_f64_gather_kernel,
                           dim3(div_up(g.ns, 64), g.na, div_up(nbatch, kParGatherSlices)),
                           dim3(64), 0, st, static_cast<const double*>(img),
                           static_cast<double*>(proj), xcos, ysin, P, nbatch, g.orbit);
        return check_launch("par_bpt_f64_gather_kernel");
    }
    const ParTables T{xcos, ysin};
    return launch_bp_gather_t<ParGeo>(P, T, static_cast<const float*>(img),
                                      static_cast<float*>(proj), nbatch, g.orbit, ws, st);
}

}  // namespace ctr
