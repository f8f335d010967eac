// Fan-beam pixel-driven backprojection for gfx950 (bp_fan_2d and
// bp_fan_2d_angle), its exact transpose, and the reference-structured
// baseline.  The forward kernel and the transpose are the BP engine
// (bp_engine.hpp) with the fan geometry policy (fan_geom.hpp): workgroups of
// one pixel tile x NB slices sharing the geometry, LDS-staged detector spans,
// FP64 per-(tile, view) anchors and FP32 pixel offsets; the transpose gathers
// each (view, channel) output in registers from the pixels that reach it
// (bp_gather.hpp), with the forward's anchors and pixel geometry.
//
// Edge rule (src/bp_fan_2d.cc:72-80): a pixel-view contributes iff
// 0 <= s <= ns-1, with taps floor(s) and ceil(s).
#include "ctr_common.hpp"
#include "bp_engine.hpp"
#include "bp_gather.hpp"
#include "fan_geom.hpp"
#include "tile_common.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace ctr {

// ------------------------------------------------- reference-structured path
// One thread per pixel on a grid-stride loop (CudaGridRangeX,
// cuda/cuda_common.h:83-91), FP64 tables read from global memory, FP64
// atan2 and accumulation per view (cuda/bp_fan_2d.cu:72-110).  The launcher
// chooses the grid: the reference's (at most 32 x 512 threads) or a full one.
template <typename T>
__global__ void plain_fan_bp_kernel(const T* __restrict__ proj, T* __restrict__ img,
                                    const double* __restrict__ xpos,
                                    const double* __restrict__ ypos,
                                    const double* __restrict__ sincos, FanProblem P,
                                    double scale)
{
    const int npix = P.nx * P.ny;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
        const int ix = i % P.nx, iy = i / P.nx;
        const double x = xpos[ix], y = ypos[iy];
        double sum = 0.0;
        for (int ia = 0; ia < P.na; ++ia) {
            const double sb = sincos[2 * ia], cb = sincos[2 * ia + 1];
            const double d = P.dso + x * sb - y * cb;
            const double r = x * cb + y * sb;
            const double s = atan2(r, d) * P.dsd / P.ds + P.cents;
            const double w = P.dsd * P.dsd / (d * d + r * r);
            if (s >= 0 && s <= P.ns - 1) {
                const double fl = floor(s);
                const int is1 = static_cast<int>(fl), is2 = static_cast<int>(ceil(s));
                const double u = s - fl;
                sum += ((1 - u) * proj[is1 + ia * P.ns] + u * proj[is2 + ia * P.ns]) * w;
            }
        }
        img[i] = static_cast<T>(sum * scale);
    }
}

// Interval of the (real) index t with p t + q >= 0, intersected into
// [lo, hi]; p = 0 keeps everything or nothing by the sign of q.
__device__ __forceinline__ void clip_halfline(double p, double q, double& lo, double& hi)
{
    if (p > 0.0)
        lo = fmax(lo, -q / p);
    else if (p < 0.0)
        hi = fmin(hi, -q / p);
    else if (q < 0.0)
        hi = -1e300;
}

// FP64 transpose in gather form (float64 buffers, both modes): one thread
// per (channel, view) and up to four slices sums, row by row and in
// increasing ix, the contributions of the pixels whose position s falls in
// [c-1, c+1].  Along an image row s is monotone (atan of a ratio of two
// functions linear in x, with d > 0 over the image), so those pixels form an
// interval, bounded here by the rays at c -+ (1 + 1/64) channels plus one
// pixel of margin; every candidate is then evaluated with the forward
// kernel's FP64 expressions.  The order of the sums is fixed: the result is
// bitwise repeatable without atomics or a cleared output.
constexpr int kGatherSlices = 4;

__global__ void fan_bpt_f64_gather_kernel(const double* __restrict__ img,
                                          double* __restrict__ proj,
                                          const double* __restrict__ xpos,
                                          const double* __restrict__ ypos,
                                          const double* __restrict__ sincos, FanProblem P,
                                          int nbatch, double scale)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int ia = blockIdx.y, b0 = blockIdx.z * kGatherSlices;
    if (c >= P.ns) return;
    const int nb = min(kGatherSlices, nbatch - b0);
    const double sb = sincos[2 * ia], cb = sincos[2 * ia + 1];
    const double k = P.dsd / P.ds;
    const double glo = (c - 1.015625 - P.cents) / k, ghi = (c + 1.015625 - P.cents) / k;
    const double half_pi = 1.5707963267948966;
    double acc[kGatherSlices] = {0.0, 0.0, 0.0, 0.0};
    // |gamma| < pi/2 everywhere on the image: channels beyond are never hit
    if (glo < half_pi && ghi > -half_pi) {
        const bool has_lo = glo > -half_pi, has_hi = ghi < half_pi;
        const double tlo = has_lo ? tan(glo) : 0.0, thi = has_hi ? tan(ghi) : 0.0;
        // x(ix) ~ x0 + ix * xs (the table accumulates dx; the margin covers it)
        const double x0 = xpos[0];
        const double xs = P.nx > 1 ? (xpos[P.nx - 1] - x0) / (P.nx - 1) : 0.0;
        for (int iy = 0; iy < P.ny; ++iy) {
            const double y = ypos[iy];
            double lo = -1.0, hi = P.nx;
            // gamma >= glo  <=>  r - tan(glo) d >= 0, linear in x (d > 0)
            if (has_lo) {
                const double A = cb - tlo * sb, B = y * sb - tlo * (P.dso - y * cb);
                clip_halfline(A * xs, A * x0 + B, lo, hi);
            }
            // gamma <= ghi  <=>  tan(ghi) d - r >= 0
            if (has_hi) {
                const double A = cb - thi * sb, B = y * sb - thi * (P.dso - y * cb);
                clip_halfline(-A * xs, -(A * x0 + B), lo, hi);
            }
            if (!(lo <= hi)) continue;
            const int ixl = max(0, static_cast<int>(floor(fmax(lo, -2.0))) - 1);
            const int ixh = min(P.nx - 1, static_cast<int>(ceil(fmin(hi, P.nx + 2.0))) + 1);
            for (int ix = ixl; ix <= ixh; ++ix) {
                const double x = xpos[ix];
                const double d = P.dso + x * sb - y * cb;
                const double r = x * cb + y * sb;
                const double s = atan2(r, d) * P.dsd / P.ds + P.cents;
                if (!(s >= 0 && s <= P.ns - 1)) continue;
                const double w = P.dsd * P.dsd / (d * d + r * r);
                const double fl = floor(s);
                const double u = s - fl;
                const int is1 = static_cast<int>(fl), is2 = static_cast<int>(ceil(s));
                const double wc = (is1 == c ? w * (1 - u) : 0.0) + (is2 == c ? w * u : 0.0);
                if (wc == 0.0) continue;
                const size_t pix = static_cast<size_t>(iy) * P.nx + ix;
                // every slice's pixel loaded first (a slice past the batch
                // re-reads the last one and is not added)
                double pv[kGatherSlices];
#pragma unroll
                for (int b = 0; b < kGatherSlices; ++b)
                    pv[b] = img[static_cast<size_t>(b0 + min(b, nb - 1)) * P.nx * P.ny + pix];
#pragma unroll
                for (int b = 0; b < kGatherSlices; ++b)
                    if (b < nb) acc[b] += pv[b] * wc;
            }
        }
    }
#pragma unroll
    for (int b = 0; b < kGatherSlices; ++b)
        if (b < nb)
            proj[(static_cast<size_t>(b0 + b) * P.na + ia) * P.ns + c] = acc[b] * scale;
}

// ---------------------------------------------------------------- host side

// Smallest source-to-pixel distance over the tile-padded pixel grid (every
// source position sits at distance dso from the rotation centre).
double fan_min_distance(const ctr_geom& g, int TW, int TH)
{
    const double cx = cent_x(g), cy = cent_y(g);
    const double wx = std::max(std::fabs(cx), std::fabs(div_up(g.nx, TW) * TW - 1 - cx)) *
                      std::fabs(g.dx);
    const double wy = std::max(std::fabs(cy), std::fabs(div_up(g.ny, TH) * TH - 1 - cy)) *
                      std::fabs(g.dy);
    return g.dso - std::sqrt(wx * wx + wy * wy);
}

// Geometry checks and the LDS span / series-accuracy bounds for a tile size.
int fan_problem(const ctr_geom& g, int TW, int TH, FanProblem& P, int& lmax, bool& exact)
{
    CTR_REQUIRE(g.ns >= 2 && g.na >= 1 && g.nx >= 1 && g.ny >= 1, CTR_ERR_ARG,
                "fan BP: sizes must be positive (ns >= 2)");
    CTR_REQUIRE(g.ds > 0 && g.dsd > 0 && g.dso > 0, CTR_ERR_ARG,
                "fan BP: ds, dso, dsd must be positive");
    P.dso = g.dso;
    P.dsd = g.dsd;
    P.ds = g.ds;
    P.cents = cent_s(g);
    P.dx = g.dx;
    P.dy = g.dy;
    P.ns = static_cast<int>(g.ns);
    P.na = static_cast<int>(g.na);
    P.nx = static_cast<int>(g.nx);
    P.ny = static_cast<int>(g.ny);
    const double k = g.dsd / g.ds;
    P.k = static_cast<float>(k);
    P.p1 = static_cast<float>(-1.0 / (3 * k * k));
    P.p2 = static_cast<float>(1.0 / (5 * k * k * k * k));
    P.g1 = static_cast<float>(-1.0 / (k * k));
    P.g2 = static_cast<float>(1.0 / (k * k * k * k));
    const double lmin = fan_min_distance(g, TW, TH);
    CTR_REQUIRE(lmin > 0.05 * g.dso, CTR_ERR_SHAPE,
                "fan BP: the source comes too close to the image (dso <= image radius)");
    // A tile of diagonal D seen from distance >= lmin subtends at most
    // 2 asin(D / (2 lmin)); the staged span needs that many channels plus
    // the floor / right-tap / rounding margin.
    const double diag = std::hypot(TW * std::fabs(g.dx), TH * std::fabs(g.dy));
    const double ang = 2.0 * std::asin(std::min(1.0, diag / (2.0 * lmin)));
    lmax = static_cast<int>(std::ceil(k * ang)) + 5;
    lmax = (lmax + 1) & ~1;
    // |u| = tan of the angle between a pixel and its anchor (at most half the
    // tile's angle); the series drop terms of order k u^7 / 7 and u^6.
    const double umax = std::tan(0.5 * ang);
    exact = (k * std::pow(umax, 7) / 7 > 2e-7) || (std::pow(umax, 6) > 5e-8) || umax > 0.25;
    return CTR_OK;
}

namespace {

template <int TW, int TH, int NB, bool EXACT, int LROW>
int launch_fan_bp_l(BpLaunch<FanGeo<EXACT>> L, const FanTables& T, hipStream_t st)
{
    using TS = TileShape<TW, TH>;
    // two span buffers of vchunk x NB rows of LROW floats must fit in LDS
    if (L.vchunk <= 0) L.vchunk = engine_vchunk<FanAnchor>(NB, LROW);
    L.vchunk = std::max(1, std::min(L.vchunk, (150 * 1024 / 8) / (NB * LROW)));
    const size_t lds = kAnchorChunk * sizeof(FanAnchor) +
                       2 * static_cast<size_t>(L.vchunk) * NB * LROW * sizeof(float);
    CTR_REQUIRE(lds <= 160 * 1024, CTR_ERR_SHAPE,
                "fan BP: view chunk x slices x span exceeds LDS; lower the view chunk");
    auto kern = bp_forward_kernel<FanGeo<EXACT>, TW, TH, NB, LROW>;
    if (lds > 64 * 1024)
        CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds)));
    dim3 grid(super_grid(L.tiles_x, L.tiles_y), div_up(L.nbatch, NB));
    hipLaunchKernelGGL(kern, grid, dim3(TS::kBlock), lds, st, T, L);
    return check_launch("bp_forward_kernel<fan>");
}

template <int TW, int TH, int NB, bool EXACT>
int launch_fan_bp(const BpLaunch<FanGeo<EXACT>>& L, const FanTables& T, hipStream_t st)
{
    // slice-interleaved spans come in 16-channel rows: a 48-channel row leaves
    // LDS for more views per round than 64
    if constexpr (NB % 4 == 0)
        if (L.lmax <= 48) return launch_fan_bp_l<TW, TH, NB, EXACT, 48>(L, T, st);
    if (L.lmax <= 64) return launch_fan_bp_l<TW, TH, NB, EXACT, 64>(L, T, st);
    if (L.lmax <= 128) return launch_fan_bp_l<TW, TH, NB, EXACT, 128>(L, T, st);
    if (L.lmax <= 256) return launch_fan_bp_l<TW, TH, NB, EXACT, 256>(L, T, st);
    if (L.lmax <= 512) return launch_fan_bp_l<TW, TH, NB, EXACT, 512>(L, T, st);
    set_error("fan BP: a pixel tile spans more than 512 detector channels");
    return CTR_ERR_SHAPE;
}

template <int TW, int TH, int NB>
int launch_fan_bp_nb(const ctr_geom& g, const FanProblem& P, int lmax, bool exact, int nbatch,
                     int vchunk, double scale, const float* proj, float* img, const FanTables& T,
                     hipStream_t st)
{
    auto fill = [&](auto& L) {
        L.P = P;
        L.nbatch = nbatch;
        L.tiles_x = div_up(g.nx, TW);
        L.tiles_y = div_up(g.ny, TH);
        L.lmax = lmax;
        L.vchunk = vchunk;
        L.scale = scale;
        L.proj = proj;
        L.img = img;
    };
    if (exact) {
        BpLaunch<FanGeo<true>> L{};
        fill(L);
        return launch_fan_bp<TW, TH, NB, true>(L, T, st);
    }
    BpLaunch<FanGeo<false>> L{};
    fill(L);
    return launch_fan_bp<TW, TH, NB, false>(L, T, st);
}

}  // namespace

double fan_bp_scale(const ctr_geom& g)
{
    const double factor = std::round(std::fabs(g.na * g.orbit) / M_PI);
    return std::fabs(g.orbit) / factor;
}

static int auto_nb(int nbatch)
{
    return nbatch >= 16 ? 16 : nbatch >= 8 ? 8 : nbatch >= 4 ? 4 : nbatch >= 2 ? 2 : 1;
}

// Launch shape of the forward kernel, per call: tile 0 = 16x16, 1 = 8x8,
// 2 = 32x8; views per LDS round (0 = engine_vchunk: on the 16-slice 512^2
// fan scan 6 views of 48-channel rows, 0.418 ms against 0.453 for 4 views of
// 64); slices per workgroup (0 = from the batch).  The caller validated it.
int fan_bp_f32(const ctr_geom& g, const float* proj, float* img, int nbatch, const double* xpos,
               const double* ypos, const double* sc, double scale, ctr_fan_bp_launch cfg,
               hipStream_t st)
{
    static const int tw_of[3] = {16, 8, 32}, th_of[3] = {16, 8, 8};
    FanProblem P{};
    int lmax = 0;
    bool exact = false;
    int rc = fan_problem(g, tw_of[cfg.tile], th_of[cfg.tile], P, lmax, exact);
    if (rc) return rc;
    if (lmax > 256 && cfg.tile != 1) {
        // coarse pixels or a close source: smaller tiles keep the span short
        cfg.tile = 1;
        rc = fan_problem(g, 8, 8, P, lmax, exact);
        if (rc) return rc;
    }
    const FanTables T{xpos, ypos, sc};
    const int nb = engine_nb(cfg.nb ? cfg.nb : auto_nb(nbatch), g.na, g.ns);
#define CTR_FAN_NB(TW_, TH_)                                                                   \
    switch (nb) {                                                                              \
        case 1: return launch_fan_bp_nb<TW_, TH_, 1>(g, P, lmax, exact, nbatch, cfg.vchunk,     \
                                                     scale, proj, img, T, st);                 \
        case 2: return launch_fan_bp_nb<TW_, TH_, 2>(g, P, lmax, exact, nbatch, cfg.vchunk,     \
                                                     scale, proj, img, T, st);                 \
        case 4: return launch_fan_bp_nb<TW_, TH_, 4>(g, P, lmax, exact, nbatch, cfg.vchunk,     \
                                                     scale, proj, img, T, st);                 \
        case 8: return launch_fan_bp_nb<TW_, TH_, 8>(g, P, lmax, exact, nbatch, cfg.vchunk,     \
                                                     scale, proj, img, T, st);                 \
        default: return launch_fan_bp_nb<TW_, TH_, 16>(g, P, lmax, exact, nbatch, cfg.vchunk,   \
                                                       scale, proj, img, T, st);               \
    }
    if (cfg.tile == 1) { CTR_FAN_NB(8, 8) }
    if (cfg.tile == 2) { CTR_FAN_NB(32, 8) }
    CTR_FAN_NB(16, 16)
#undef CTR_FAN_NB
}

// float32 transpose: the gather form (bp_gather.hpp), deterministic in both
// modes.  ws: caller workspace of gather_workspace_bytes<FanAnchor>(...)
// bytes, or null for a stream-ordered allocation.
int fan_bpt_f32(const ctr_geom& g, const float* img, float* proj, int nbatch, const double* xpos,
                const double* ypos, const double* sc, double scale, bool det, void* ws,
                hipStream_t st)
{
    FanProblem P{};
    int lmax = 0;
    bool exact = false;
    int rc = fan_problem(g, 16, 16, P, lmax, exact);
    if (rc) return rc;
    (void)det;
    const FanTables T{xpos, ypos, sc};
    return exact ? launch_bp_gather_t<FanGeo<true>>(P, T, img, proj, nbatch, scale, ws, st)
                 : launch_bp_gather_t<FanGeo<false>>(P, T, img, proj, nbatch, scale, ws, st);
}

int fan_bp_f64(const ctr_geom& g, const double* proj, double* img, int nbatch,
               const double* xpos, const double* ypos, const double* sc, double scale,
               hipStream_t st)
{
    FanProblem P{};
    int lmax;
    bool exact;
    int rc = fan_problem(g, 16, 16, P, lmax, exact);
    if (rc) return rc;
    const int npix = g.nx * g.ny;
    for (int b = 0; b < nbatch; ++b) {
        hipLaunchKernelGGL(plain_fan_bp_kernel<double>, dim3(div_up(npix, 256)), dim3(256), 0, st,
                           proj + static_cast<size_t>(b) * g.na * g.ns,
                           img + static_cast<size_t>(b) * npix, xpos, ypos, sc, P, scale);
        rc = check_launch("plain_fan_bp_kernel<double>");
        if (rc) return rc;
    }
    return CTR_OK;
}

// float64: the gather form, deterministic in both modes.
int fan_bpt_f64(const ctr_geom& g, const double* img, double* proj, int nbatch,
                const double* xpos, const double* ypos, const double* sc, double scale,
                hipStream_t st)
{
    FanProblem P{};
    int lmax;
    bool exact;
    int rc = fan_problem(g, 16, 16, P, lmax, exact);
    if (rc) return rc;
    hipLaunchKernelGGL(fan_bpt_f64_gather_kernel,
                       dim3(div_up(g.ns, 64), g.na, div_up(nbatch, kGatherSlices)), dim3(64), 0,
                       st, img, proj, xpos, ypos, sc, P, nbatch, scale);
    return check_launch("fan_bpt_f64_gather_kernel");
}

// max_blocks = 0: a full grid of 256-thread workgroups; otherwise the
// reference's GetCudaLaunchConfig (cuda/cuda_common.h:30-38) with the given
// caps.  One launch per slice and a stream synchronisation after each.
int plain_fan_bp(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
                 const double* xpos, const double* ypos, const double* sc, double scale,
                 int max_blocks, int max_threads, hipStream_t st)
{
    FanProblem P{};
    int lmax;
    bool exact;
    int rc = fan_problem(g, 16, 16, P, lmax, exact);
    if (rc) return rc;
    const int npix = g.nx * g.ny;
    int threads = 256, blocks = div_up(npix, 256);
    if (max_blocks > 0) {
        threads = std::min(max_threads, npix);
        blocks = std::min(div_up(npix, threads), max_blocks);
    }
    for (int b = 0; b < nbatch; ++b) {
        if (dtype == CTR_F32)
            hipLaunchKernelGGL(plain_fan_bp_kernel<float>, dim3(blocks), dim3(threads), 0, st,
                               static_cast<const float*>(proj) + static_cast<size_t>(b) * g.na * g.ns,
                               static_cast<float*>(img) + static_cast<size_t>(b) * npix, xpos, ypos,
                               sc, P, scale);
        else
            hipLaunchKernelGGL(plain_fan_bp_kernel<double>, dim3(blocks), dim3(threads), 0, st,
                               static_cast<const double*>(proj) + static_cast<size_t>(b) * g.na * g.ns,
                               static_cast<double*>(img) + static_cast<size_t>(b) * npix, xpos, ypos,
                               sc, P, scale);
        rc = check_launch("plain_fan_bp_kernel");
        if (rc) return rc;
        CTR_HIP(hipStreamSynchronize(st));
    }
    return CTR_OK;
}

}  // namespace ctr
