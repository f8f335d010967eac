// Fan-beam pixel-driven backprojection for gfx950, its exact transpose, and
// the reference-structured baseline.
//
// Forward (fan_bp_kernel): one workgroup per (pixel tile, group of NB
// slices).  All slices of a batch share the geometry, so the per
// pixel-view geometry (channel position and weight, ~20 VALU ops) is
// evaluated once and reused for NB slices, whose cost is then an LDS read
// pair, a lerp and an FMA.  Per chunk of views the detector span each tile
// touches is staged into LDS for every slice (zero outside [0, ns-1]);
// per-(tile, view) anchors are computed in FP64 inside the kernel
// (fan_geom.hpp).  Partial sums are kept in FP32 within a chunk of views and
// carried across chunks in FP64.
//
// Transpose (fan_bpt_kernel): view-parallel scatter.  One workgroup per
// (view, slice group) owns that view's detector rows in LDS, walks all
// pixel tiles with the same tiling and anchors as the forward kernel (so the
// weights are bit-identical), scatters with LDS atomics and writes the rows
// once, coalesced: no global atomics and no zero-fill pass.  Deterministic
// mode accumulates in 64-bit fixed point (integer LDS atomics), which is
// associative, so the result is independent of the order of contributions.
//
// Edge rule (src/bp_fan_2d.cc:72-80): a pixel-view contributes iff
// 0 <= s <= ns-1, with taps floor(s) and ceil(s).  With s = is0 + f the
// test is  s_int >= 0  and  (s_int < ns-1  or  (s_int == ns-1 and frac == 0)).
#include "ctr_common.hpp"
#include "fan_geom.hpp"
#include "tile_common.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace ctr {

struct FanLaunch {
    FanProblem P;
    int nbatch;
    int tiles_x, tiles_y;
    int lmax;      // staged channels per (view, slice)
    int vchunk;    // views staged per LDS round
    double scale;  // output scale
};

template <int TW, int TH, int NB, bool EXACT>
__global__ __launch_bounds__((TileShape<TW, TH>::kBlock)) void fan_bp_kernel(
    const float* __restrict__ proj, float* __restrict__ img, const double* __restrict__ xpos,
    const double* __restrict__ ypos, const double* __restrict__ sincos, const FanLaunch L)
{
    using TS = TileShape<TW, TH>;
    constexpr int BLOCK = TS::kBlock, PPT = TS::kPPT;
    extern __shared__ __align__(16) unsigned char smem[];
    FanAnchor* anchors = reinterpret_cast<FanAnchor*>(smem);
    float* spans = reinterpret_cast<float*>(smem + kAnchorChunk * sizeof(FanAnchor));

    const FanProblem& P = L.P;
    const int tid = threadIdx.x;
    const TileFrame F(blockIdx.x, L.tiles_x, P.nx, P.ny, TW, TH);
    const int b0 = blockIdx.y * NB;
    const int nb = min(NB, L.nbatch - b0);
    const int lmax = L.lmax;
    const double xc = xpos[F.ic], yc = ypos[F.jc];

    float fi[PPT], fj[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int p = tid + k * BLOCK;
        fi[k] = static_cast<float>(F.x0 + p % TW - F.ic);
        fj[k] = static_cast<float>(F.y0 + p / TW - F.jc);
    }
    double acc64[PPT][NB];
#pragma unroll
    for (int k = 0; k < PPT; ++k)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc64[k][b] = 0.0;

    const int ilo = F.x0 - F.ic, ihi = F.x0 + TW - 1 - F.ic;
    const int jlo = F.y0 - F.jc, jhi = F.y0 + TH - 1 - F.jc;

    for (int va = 0; va < P.na; va += kAnchorChunk) {
        const int nva = min(kAnchorChunk, P.na - va);
        __syncthreads();  // previous anchors and spans are no longer read
        if (tid < nva)
            anchors[tid] = fan_anchor(P, xc, yc, sincos[2 * (va + tid)],
                                      sincos[2 * (va + tid) + 1], ilo, ihi, jlo, jhi);
        float acc[PPT][NB];
#pragma unroll
        for (int k = 0; k < PPT; ++k)
#pragma unroll
            for (int b = 0; b < NB; ++b) acc[k][b] = 0.0f;

        for (int vc = 0; vc < nva; vc += L.vchunk) {
            const int nv = min(L.vchunk, nva - vc);
            __syncthreads();  // anchors written / previous spans consumed
            // Stage spans: row (t, b) holds channels is0 .. is0+lmax-1.
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
                const FanAnchor A = anchors[vc + t];
                const float* srow = spans + t * NB * lmax;
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    float f, w;
                    fan_pixel<EXACT>(P, A, fi[k], fj[k], f, w);
                    const float fl = floorf(f);
                    const float fr = f - fl;
                    int idx = static_cast<int>(fl);
                    if (!A.inside) {
                        const int s = A.is0 + idx;
                        const bool ok = on_detector(s, fr, P.ns);
                        w = ok ? w : 0.0f;
                    }
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

// ---------------------------------------------------------------- transpose
struct FanTLaunch {
    FanProblem P;
    int nbatch;
    int tiles_x, tiles_y;
    double scale;
    const double* fix_scale;  // deterministic mode: device pointer to 2^E
};

template <int TW, int TH, int NB, bool EXACT, bool DET>
__global__ __launch_bounds__(256) void fan_bpt_kernel(
    const float* __restrict__ img, float* __restrict__ proj, const double* __restrict__ xpos,
    const double* __restrict__ ypos, const double* __restrict__ sincos, const FanTLaunch L)
{
    using TS = TileShape<TW, TH>;
    constexpr int BLOCK = TS::kBlock, PPT = TS::kPPT;
    static_assert(BLOCK == 256, "transpose assumes 256-thread tiles");
    using Acc = typename std::conditional<DET, unsigned long long, float>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    FanAnchor* anchors = reinterpret_cast<FanAnchor*>(smem);
    Acc* rows = reinterpret_cast<Acc*>(smem + kAnchorChunk * sizeof(FanAnchor));

    const FanProblem& P = L.P;
    const int tid = threadIdx.x;
    const int view = blockIdx.x;
    const int b0 = blockIdx.y * NB;
    const int nb = min(NB, L.nbatch - b0);
    const int rlen = P.ns + 1;  // one pad slot for the ceil tap at s == ns-1
    const double sb = sincos[2 * view], cb = sincos[2 * view + 1];
    double fixs = 1.0;
    if constexpr (DET) fixs = *L.fix_scale;

    for (int e = tid; e < NB * rlen; e += BLOCK) rows[e] = Acc(0);

    const int ntiles = L.tiles_x * L.tiles_y;
    for (int tc = 0; tc < ntiles; tc += kAnchorChunk) {
        const int ntc = min(kAnchorChunk, ntiles - tc);
        __syncthreads();
        if (tid < ntc) {
            const TileFrame F(tc + tid, L.tiles_x, P.nx, P.ny, TW, TH);
            anchors[tid] = fan_anchor(P, xpos[F.ic], ypos[F.jc], sb, cb, F.x0 - F.ic,
                                      F.x0 + TW - 1 - F.ic, F.y0 - F.jc, F.y0 + TH - 1 - F.jc);
        }
        __syncthreads();
        for (int t = 0; t < ntc; ++t) {
            const FanAnchor A = anchors[t];
            const TileFrame F(tc + t, L.tiles_x, P.nx, P.ny, TW, TH);
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                const int p = tid + k * BLOCK;
                const int ix = F.x0 + p % TW, iy = F.y0 + p / TW;
                if (ix >= P.nx || iy >= P.ny) continue;
                float f, w;
                fan_pixel<EXACT>(P, A, static_cast<float>(ix - F.ic),
                                 static_cast<float>(iy - F.jc), f, w);
                const float fl = floorf(f);
                const float fr = f - fl;
                const int s = A.is0 + static_cast<int>(fl);
                const bool ok = on_detector(s, fr, P.ns);
                if (!ok) continue;
                const float w1 = w * fr, w0 = w - w1;
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    if (b >= nb) break;
                    const float v = img[(static_cast<size_t>(b0 + b) * P.ny + iy) * P.nx + ix];
                    if constexpr (DET) {
                        const long long q0 = __double2ll_rn(static_cast<double>(w0 * v) * fixs);
                        const long long q1 = __double2ll_rn(static_cast<double>(w1 * v) * fixs);
                        atomicAdd(&rows[b * rlen + s], static_cast<unsigned long long>(q0));
                        atomicAdd(&rows[b * rlen + s + 1], static_cast<unsigned long long>(q1));
                    } else {
                        atomicAdd(&rows[b * rlen + s], w0 * v);
                        atomicAdd(&rows[b * rlen + s + 1], w1 * v);
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

// ------------------------------------------------- reference-structured path
template <typename T>
__global__ void plain_fan_bp_kernel(const T* __restrict__ proj, T* __restrict__ img,
                                    const double* __restrict__ xpos,
                                    const double* __restrict__ ypos,
                                    const double* __restrict__ sincos, FanProblem P,
                                    double scale)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nx * P.ny) return;
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

// FP64 transpose for float64 buffers: one thread per pixel scattering with
// FP64 global atomics (double is the reference's high-precision mode).
__global__ void fan_bpt_f64_kernel(const double* __restrict__ img, double* __restrict__ proj,
                                   const double* __restrict__ xpos,
                                   const double* __restrict__ ypos,
                                   const double* __restrict__ sincos, FanProblem P, int nbatch)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= P.nx * P.ny || b >= nbatch) return;
    const int ix = i % P.nx, iy = i / P.nx;
    const double x = xpos[ix], y = ypos[iy];
    const double v = img[static_cast<size_t>(b) * P.nx * P.ny + i];
    double* out = proj + static_cast<size_t>(b) * P.na * P.ns;
    for (int ia = 0; ia < P.na; ++ia) {
        const double sb = sincos[2 * ia], cb = sincos[2 * ia + 1];
        const double d = P.dso + x * sb - y * cb;
        const double r = x * cb + y * sb;
        const double w = P.dsd * P.dsd / (d * d + r * r);
        const double s = atan2(r, d) * P.dsd / P.ds + P.cents;
        if (s >= 0 && s <= P.ns - 1) {
            const double fl = floor(s);
            const double u = s - fl;
            const int is1 = static_cast<int>(fl), is2 = static_cast<int>(ceil(s));
            atomicAdd(&out[is1 + ia * P.ns], v * (w * (1 - u)));
            atomicAdd(&out[is2 + ia * P.ns], v * (w * u));
        }
    }
}

// ---------------------------------------------------------------- host side
namespace {

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
    // Farthest pixel (tile-padded grid) from the rotation centre.
    const double cx = cent_x(g), cy = cent_y(g);
    const double wx = std::max(std::fabs(-cx), std::fabs(div_up(g.nx, TW) * TW - 1 - cx)) *
                      std::fabs(g.dx);
    const double wy = std::max(std::fabs(-cy), std::fabs(div_up(g.ny, TH) * TH - 1 - cy)) *
                      std::fabs(g.dy);
    const double rimg = std::sqrt(wx * wx + wy * wy);
    const double lmin = g.dso - rimg;
    CTR_REQUIRE(lmin > 0.05 * g.dso, CTR_ERR_SHAPE,
                "fan BP: the source comes too close to the image (dso <= image radius)");
    // channels per pixel step <= k * pitch / distance
    const double ext = TW * std::fabs(g.dx) + TH * std::fabs(g.dy);
    lmax = static_cast<int>(std::ceil(k * ext / lmin)) + 4;
    lmax = (lmax + 1) & ~1;
    const double umax = std::tan(0.5 * ext / lmin);
    exact = (k * std::pow(umax, 7) / 7 > 2e-7) || (std::pow(umax, 6) > 5e-8) || umax > 0.25;
    return CTR_OK;
}

template <int TW, int TH, int NB, bool EXACT>
int launch_fan_bp(const FanLaunch& L, const float* proj, float* img, const double* xpos,
                  const double* ypos, const double* sc, hipStream_t st)
{
    using TS = TileShape<TW, TH>;
    const size_t lds = kAnchorChunk * sizeof(FanAnchor) +
                       static_cast<size_t>(L.vchunk) * NB * L.lmax * sizeof(float);
    CTR_REQUIRE(lds <= 160 * 1024, CTR_ERR_SHAPE, "fan BP: detector span too wide for LDS");
    auto kern = fan_bp_kernel<TW, TH, NB, EXACT>;
    if (lds > 64 * 1024)
        CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds)));
    dim3 grid(L.tiles_x * L.tiles_y, div_up(L.nbatch, NB));
    hipLaunchKernelGGL(kern, grid, dim3(TS::kBlock), lds, st, proj, img, xpos, ypos, sc, L);
    return check_launch("fan_bp_kernel");
}

template <int TW, int TH, int NB>
int launch_fan_bp_nb(const FanLaunch& L, bool exact, const float* proj, float* img,
                     const double* xpos, const double* ypos, const double* sc, hipStream_t st)
{
    return exact ? launch_fan_bp<TW, TH, NB, true>(L, proj, img, xpos, ypos, sc, st)
                 : launch_fan_bp<TW, TH, NB, false>(L, proj, img, xpos, ypos, sc, st);
}

}  // namespace

double fan_bp_scale(const ctr_geom& g)
{
    const double factor = std::round(std::fabs(g.na * g.orbit) / M_PI);
    return std::fabs(g.orbit) / factor;
}

// Tunables of the forward kernel (the autotuner writes them).
struct FanBpConfig {
    int tile = 0;    // 0: 16x16, 1: 8x8, 2: 32x8
    int vchunk = 16;
    int nb = 0;      // 0: automatic
};
static thread_local FanBpConfig g_fan_cfg;

int fan_bp_f32(const ctr_geom& g, const float* proj, float* img, int nbatch, const double* xpos,
               const double* ypos, const double* sc, double scale, hipStream_t st)
{
    static const int tw_of[3] = {16, 8, 32}, th_of[3] = {16, 8, 8};
    const int tile = g_fan_cfg.tile;
    const int TW = tw_of[tile], TH = th_of[tile];
    FanLaunch L{};
    bool exact = false;
    int rc = fan_problem(g, TW, TH, L.P, L.lmax, exact);
    if (rc) return rc;
    L.nbatch = nbatch;
    L.tiles_x = div_up(g.nx, TW);
    L.tiles_y = div_up(g.ny, TH);
    L.vchunk = std::max(1, std::min(g_fan_cfg.vchunk, kAnchorChunk));
    L.scale = scale;
    int nb = g_fan_cfg.nb;
    if (nb == 0) nb = nbatch >= 16 ? 16 : nbatch >= 8 ? 8 : nbatch >= 4 ? 4 : nbatch >= 2 ? 2 : 1;
#define CTR_FAN_NB(TW_, TH_)                                                                  \
    switch (nb) {                                                                             \
        case 1: return launch_fan_bp_nb<TW_, TH_, 1>(L, exact, proj, img, xpos, ypos, sc, st); \
        case 2: return launch_fan_bp_nb<TW_, TH_, 2>(L, exact, proj, img, xpos, ypos, sc, st); \
        case 4: return launch_fan_bp_nb<TW_, TH_, 4>(L, exact, proj, img, xpos, ypos, sc, st); \
        case 8: return launch_fan_bp_nb<TW_, TH_, 8>(L, exact, proj, img, xpos, ypos, sc, st); \
        default: return launch_fan_bp_nb<TW_, TH_, 16>(L, exact, proj, img, xpos, ypos, sc, st); \
    }
    if (tile == 1) { CTR_FAN_NB(8, 8) }
    if (tile == 2) { CTR_FAN_NB(32, 8) }
    CTR_FAN_NB(16, 16)
#undef CTR_FAN_NB
}

void fan_bp_set_config(int tile, int vchunk, int nb)
{
    g_fan_cfg.tile = (tile >= 0 && tile <= 2) ? tile : 0;
    g_fan_cfg.vchunk = vchunk > 0 ? vchunk : 16;
    g_fan_cfg.nb = nb;
}

int fan_bpt_f32(const ctr_geom& g, const float* img, float* proj, int nbatch, const double* xpos,
                const double* ypos, const double* sc, double scale, bool det, hipStream_t st)
{
    constexpr int TW = 16, TH = 16;
    FanTLaunch L{};
    int lmax = 0;
    bool exact = false;
    int rc = fan_problem(g, TW, TH, L.P, lmax, exact);
    if (rc) return rc;
    L.nbatch = nbatch;
    L.tiles_x = div_up(g.nx, TW);
    L.tiles_y = div_up(g.ny, TH);
    L.scale = scale;
    const size_t rlen = g.ns + 1;
    if (det) {
        // Fixed-point scale from max|img| * max weight * contributions per channel.
        float* partial = nullptr;
        double* fs = nullptr;
        CTR_HIP(hipMallocAsync(reinterpret_cast<void**>(&partial), kMaxBlocks * sizeof(float), st));
        CTR_HIP(hipMallocAsync(reinterpret_cast<void**>(&fs), sizeof(double), st));
        const size_t n = static_cast<size_t>(nbatch) * g.nx * g.ny;
        const int nblk = static_cast<int>(std::min<size_t>(kMaxBlocks, (n + 255) / 256));
        hipLaunchKernelGGL(abs_max_partial_kernel, dim3(nblk), dim3(256), 0, st, img, n, partial);
        const double rimg = std::sqrt(std::pow(std::fabs(cent_x(g)) * g.dx + g.nx * g.dx, 2) +
                                      std::pow(std::fabs(cent_y(g)) * g.dy + g.ny * g.dy, 2));
        const double wmax = g.dsd / std::max(g.dso - rimg, 1e-3 * g.dso);
        const double bound = 2.0 * g.nx * g.ny * std::max(1.0, wmax * wmax);
        hipLaunchKernelGGL(fix_scale_kernel, dim3(1), dim3(64), 0, st, partial, nblk, bound, fs);
        L.fix_scale = fs;
        constexpr int NB = 4;
        const size_t lds = kAnchorChunk * sizeof(FanAnchor) + NB * rlen * 8;
        CTR_REQUIRE(lds <= 160 * 1024, CTR_ERR_SHAPE, "fan BP transpose: ns too large");
        dim3 grid(g.na, div_up(nbatch, NB));
        auto kern = exact ? fan_bpt_kernel<TW, TH, NB, true, true>
                          : fan_bpt_kernel<TW, TH, NB, false, true>;
        if (lds > 64 * 1024)
            CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
        hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, img, proj, xpos, ypos, sc, L);
        rc = check_launch("fan_bpt_kernel<det>");
        CTR_HIP(hipFreeAsync(partial, st));
        CTR_HIP(hipFreeAsync(fs, st));
        return rc;
    }
    const int nb = nbatch >= 8 ? 8 : nbatch >= 4 ? 4 : nbatch >= 2 ? 2 : 1;
    const size_t lds = kAnchorChunk * sizeof(FanAnchor) + static_cast<size_t>(nb) * rlen * 4;
    CTR_REQUIRE(lds <= 160 * 1024, CTR_ERR_SHAPE, "fan BP transpose: ns too large");
    dim3 grid(g.na, div_up(nbatch, nb));
#define CTR_FANT(NB_)                                                                       \
    {                                                                                       \
        auto kern = exact ? fan_bpt_kernel<TW, TH, NB_, true, false>                        \
                          : fan_bpt_kernel<TW, TH, NB_, false, false>;                      \
        if (lds > 64 * 1024)                                                                \
            CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),                \
                                        hipFuncAttributeMaxDynamicSharedMemorySize,         \
                                        static_cast<int>(lds)));                            \
        hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, img, proj, xpos, ypos, sc, L);   \
        return check_launch("fan_bpt_kernel");                                              \
    }
    switch (nb) {
        case 1: CTR_FANT(1)
        case 2: CTR_FANT(2)
        case 4: CTR_FANT(4)
        default: CTR_FANT(8)
    }
#undef CTR_FANT
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

int fan_bpt_f64(const ctr_geom& g, const double* img, double* proj, int nbatch,
                const double* xpos, const double* ypos, const double* sc, double scale,
                hipStream_t st)
{
    FanProblem P{};
    int lmax;
    bool exact;
    int rc = fan_problem(g, 16, 16, P, lmax, exact);
    if (rc) return rc;
    const size_t n = static_cast<size_t>(nbatch) * g.na * g.ns;
    CTR_HIP(hipMemsetAsync(proj, 0, n * sizeof(double), st));
    hipLaunchKernelGGL(fan_bpt_f64_kernel, dim3(div_up(g.nx * g.ny, 256), nbatch), dim3(256), 0,
                       st, img, proj, xpos, ypos, sc, P, nbatch);
    rc = check_launch("fan_bpt_f64_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(scale_f64_kernel, dim3(std::min<size_t>(4096, (n + 255) / 256)), dim3(256),
                       0, st, proj, n, scale);
    return check_launch("scale_f64_kernel");
}

int plain_fan_bp(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
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
        if (dtype == CTR_F32)
            hipLaunchKernelGGL(plain_fan_bp_kernel<float>, dim3(div_up(npix, 256)), dim3(256), 0,
                               st, static_cast<const float*>(proj) + static_cast<size_t>(b) * g.na * g.ns,
                               static_cast<float*>(img) + static_cast<size_t>(b) * npix, xpos, ypos,
                               sc, P, scale);
        else
            hipLaunchKernelGGL(plain_fan_bp_kernel<double>, dim3(div_up(npix, 256)), dim3(256), 0,
                               st, static_cast<const double*>(proj) + static_cast<size_t>(b) * g.na * g.ns,
                               static_cast<double*>(img) + static_cast<size_t>(b) * npix, xpos, ypos,
                               sc, P, scale);
        rc = check_launch("plain_fan_bp_kernel");
        if (rc) return rc;
        CTR_HIP(hipStreamSynchronize(st));
    }
    return CTR_OK;
}

}  // namespace ctr
