// Gather-form exact transpose of the pixel-driven BP engine, for float32.
//
// The BP sums, for every pixel and view, w * lerp(proj[view], s) with
// s = s(pixel, view) and the taps floor(s), floor(s)+1.  Its transpose
// therefore gives channel c of a view the sum of img * w * (1 - frac) over
// the pixels with floor(s) = c, plus img * w * frac over those with
// floor(s) = c - 1.  Here one lane owns one (view, channel interval
// [i, i+1)): it finds the pixels with floor(s) = i and sums, in registers,
// both P_i = sum img * w * (1 - frac) and Q_i = sum img * w * frac; channel c
// is then P_c + Q_{c-1}, Q taken from the lane on the left.  Every pixel is
// a candidate of one lane (two, when the channel on the left is wanted too)
// and all sums run in a fixed order (image rows in increasing order, pixels
// along a row in increasing order): no atomics, no cleared output, and the
// result is bitwise repeatable.
//
// Finding the pixels: the level line s = s* of both geometries is a
// straight line in pixel-index space (a ray: r - tan(gamma*) d = 0 for the
// fan, x cos + y sin = const for parallel beams), written L(ix, iy) =
// pi ix + qj iy + rr with L >= 0 exactly where s >= s* (Geo::boundary).  On
// one image row the pixels with s in [i-eps, i+1+eps] are then the
// intersection of two half-lines, computed per row from two linear bounds;
// each candidate is evaluated with the forward kernel's own per-tile
// anchors and FP32 pixel geometry (Geo::pixel), so the weights are those of
// the forward.  The margin eps (1/32 channel) keeps every pixel of the
// support inside the candidate interval despite FP32 rounding of the bounds.
//
// Work layout (gfx950): a workgroup = 4 views (one wave each) x 64 channel
// intervals (one lane each: 63 output channels, lane 0 only supplies the Q
// of the interval left of them) x NB slices.  It walks the image in bands of R rows; per
// band the union of its lanes' candidate intervals is staged into LDS in
// chunks of WMAX pixels (slice-interleaved: [row][slice group][pixel][4
// slices], one ds_read_b128 per 4 slices), together with the anchors of the
// 16x16 tiles the chunk covers (from a table built once per call).  Views
// whose rays run closer to the image rows than to the columns walk the
// columns instead ("frame 1"), on a transposed copy of the image, so that a
// row's candidates are always a short interval.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "bp_engine.hpp"
#include "ctr_common.hpp"
#include "tile_common.hpp"

namespace ctr {

constexpr int kGatherTile = 16;     // anchors of 16 x 16 tiles (the forward's default)
constexpr int kGatherViews = 4;     // views per workgroup (one wave each)
constexpr int kGatherLanes = 64;    // channel intervals per workgroup (one lane each)
constexpr int kGatherOut = kGatherLanes - 1;  // output channels per workgroup
constexpr double kGatherEps = 1.0 / 32;

// anchors[view][tile] of 16 x 16 tiles: the FP64 per-(tile, view) geometry,
// evaluated once per call.
template <class Geo>
__global__ void bp_anchor_table_kernel(const typename Geo::Problem P, const typename Geo::Tables T,
                                       int tiles_x, int ntiles, typename Geo::Anchor* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y;
    if (t >= ntiles) return;
    const TileFrame F(t, tiles_x, P.nx, P.ny, kGatherTile, kGatherTile);
    out[static_cast<size_t>(v) * ntiles + t] =
        Geo::anchor(P, T, F, v, F.x0 - F.ic, F.x0 + kGatherTile - 1 - F.ic, F.y0 - F.jc,
                    F.y0 + kGatherTile - 1 - F.jc);
}

// dst[b][x][y] = src[b][y][x] (the image seen column by column).
static __global__ void image_transpose_kernel(const float* __restrict__ src,
                                              float* __restrict__ dst, int ny, int nx)
{
    __shared__ float tile[32][33];
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

template <class Geo>
struct GatherLaunch {
    typename Geo::Problem P;
    typename Geo::Tables T;
    const typename Geo::Anchor* anchors;  // [na][tiles_x * tiles_y]
    const float* img;                     // [nbatch][ny][nx]
    const float* imgT;                    // [nbatch][nx][ny]
    float* proj;                          // [nbatch][na][ns]
    int nbatch, tiles_x, tiles_y;
    double scale;
};

// x of lane `lane` (uniform) of the wave, through a scalar register.
__device__ __forceinline__ float lane_value(float x, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}

// One linear constraint of a lane's candidates along a row: w >= A + B u
// (lower) or w <= A + B u (upper); none = no constraint from this line.
struct GatherBound {
    double A, B;
    bool lower, none;
};

// Constraint sign * L >= 0 in frame coordinates (u walked, w along the row):
// L = cw w + cu u + rr.
__device__ inline GatherBound gather_bound(double cw, double cu, double rr, double sign)
{
    cw *= sign;
    cu *= sign;
    rr *= sign;
    GatherBound b;
    const double mag = fabs(cw) + fabs(cu);
    if (!(fabs(cw) > 1e-12 * mag) || !(mag > 0)) {
        // parallel to the rows (or a degenerate one-pixel axis): no bound
        b.A = b.B = 0.0;
        b.lower = true;
        b.none = true;
        return b;
    }
    b.A = -rr / cw;
    b.B = -cu / cw;
    b.lower = cw > 0;
    b.none = false;
    return b;
}

template <class Geo, int NB, int R, int WMAX>
struct GatherShape {
    static constexpr bool kQuad = NB % 4 == 0;
    static constexpr int kGroups = kQuad ? NB / 4 : 1;
    static constexpr int kTiles = WMAX / kGatherTile;  // tiles along w per chunk
    // one float4 of padding per staged row: lanes of a wave on different
    // rows then fall on different LDS banks
    static constexpr int kRowF4 = (kQuad ? kGroups * WMAX : WMAX / 4) + 1;
    static constexpr int kImgFloats = R * kRowF4 * 4;
    // anchors are staged as their first 32 bytes (all the gather reads)
    static constexpr int kAnchorDwords = kGatherViews * kTiles * 8;
    static constexpr size_t kLds = kImgFloats * sizeof(float) + kAnchorDwords * 4 + 64;
    static_assert(NB == 1 || kQuad, "slices per workgroup: 1 or a multiple of 4");
    static_assert(WMAX % (kQuad ? 16 : 64) == 0, "chunk width: whole DMA rows");
    static_assert(kGatherTile % R == 0, "bands must not straddle a tile row");
};

// LDS (about 34 KB at 16 slices) keeps four workgroups on a CU, i.e. four
// waves per SIMD: the registers of that occupancy (128) are available.
template <class Geo, int NB, int R, int WMAX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4)))
void bp_gather_t_kernel(const GatherLaunch<Geo> L)
{
    using S = GatherShape<Geo, NB, R, WMAX>;
    using Anchor = typename Geo::Anchor;
    constexpr int V = kGatherViews, K = S::kTiles, NG = S::kGroups;
    static_assert(sizeof(Anchor) % 16 == 0 && sizeof(Anchor) >= 32 &&
                      offsetof(Anchor, inside) < 32 && offsetof(Anchor, is0) < 32,
                  "the gather reads an anchor's first 32 bytes");
    static_assert(kGatherViews * S::kTiles <= 256, "one thread per staged anchor");
    extern __shared__ __align__(16) unsigned char smem[];
    float* simg = reinterpret_cast<float*>(smem);
    unsigned int* sanc = reinterpret_cast<unsigned int*>(smem + S::kImgFloats * sizeof(float));
    int* sred = reinterpret_cast<int*>(smem + S::kImgFloats * sizeof(float) + S::kAnchorDwords * 4);

    const auto& P = L.P;
    typename Geo::Problem Pg = P;
    if constexpr (HasPin<Geo>::value) Geo::pin(Pg);
    const int tid = threadIdx.x, lane = tid & 63, wave = uniform_i(tid >> 6);
    // lane: interval [ci, ci + 1); lanes 1..63 write channels c0 .. c0 + 62
    const int c0 = blockIdx.x * kGatherOut, v0 = blockIdx.y * V, b0 = blockIdx.z * NB;
    const int ci = c0 - 1 + lane, view = v0 + wave;
    const int nb = min(NB, L.nbatch - b0);
    const bool active = ci < P.ns && view < P.na;

    // Frame of the workgroup, from the level line through its middle view
    // and channel: walk rows (0) when that line is closer to the columns.
    int frame;
    {
        const int vm = v0 + (min(V, P.na - v0) - 1) / 2;
        const int cm = c0 + (min(kGatherOut, P.ns - c0) - 1) / 2;
        double pi, qj, rr;
        Geo::boundary(P, L.T, vm, static_cast<double>(cm), pi, qj, rr);
        frame = uniform_i(fabs(pi) >= fabs(qj) ? 0 : 1);
    }
    const int nu = frame ? P.nx : P.ny, nw = frame ? P.ny : P.nx;
    const float* src = frame ? L.imgT : L.img;
    const unsigned int slice_bytes = static_cast<unsigned int>(nu) * nw * 4u;

    // This lane's two constraints, s >= ci - eps and s <= ci + 1 + eps, as
    // one lower and one upper line along the row: w >= la + lb u and
    // w <= ua + ub u.  A constraint parallel to the rows bounds nothing, and
    // should both be lower (or upper) lines, one of them is kept: either way
    // the candidates are a superset (the exact test in the loop decides).
    double la = -1e30, lb = 0.0, ua = 1e30, ub = 0.0;
    bool open = false;    // a side without a bound
    bool lower0 = false;  // the first constraint is the lower line
    {
        const int vv = active ? view : min(view, P.na - 1);
        const double sl[2] = {ci - kGatherEps, ci + 1 + kGatherEps};
        bool has_l = false, has_u = false;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            double pi, qj, rr;
            Geo::boundary(P, L.T, vv, sl[k], pi, qj, rr);
            const double cw = frame ? qj : pi, cu = frame ? pi : qj;
            const GatherBound b = gather_bound(cw, cu, rr, k == 0 ? 1.0 : -1.0);
            if (k == 0) lower0 = b.lower;
            if (b.none) continue;
            if (b.lower && !has_l) {
                la = b.A;
                lb = b.B;
                has_l = true;
            } else if (!b.lower && !has_u) {
                ua = b.A;
                ub = b.B;
                has_u = true;
            }
        }
        open = active && !(has_l && has_u);
    }
    // A lane's level lines turn monotonically with its channel (by less
    // than pi over the wave), so while no line of the wave has turned
    // through the row direction -- every lane has both bounds, all with the
    // same orientation -- the bounds along a row move monotonically with the
    // lane and the union of the wave's intervals is the hull of its first
    // and last lanes'.  Otherwise every lane takes part in a reduction.
    const bool hull = __ballot(open) == 0 && (__ballot(active && lower0) == 0 ||
                                              __ballot(active && !lower0) == 0);
    const int last = min(kGatherLanes - 1, P.ns - c0);  // last lane with ci < ns
    // the image rows of the frame, per 4-slice group
    const float* gsrc[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g)
        gsrc[g] = src + static_cast<size_t>(min(b0 + 4 * g, L.nbatch - 1)) * nu * nw;

    float accp[NB], accq[NB];  // P and Q of this lane's interval
#pragma unroll
    for (int b = 0; b < NB; ++b) accp[b] = accq[b] = 0.0f;

    const int nbands = (nu + R - 1) / R;
    for (int band = 0; band < nbands; ++band) {
        const int u0 = band * R;
        const int rmax = min(R, nu - u0);
        // Bounds re-anchored at the band's first row (FP64), then FP32 per row.
        const float fla = static_cast<float>(la + lb * u0), flb = static_cast<float>(lb);
        const float fua = static_cast<float>(ua + ub * u0), fub = static_cast<float>(ub);
        // This lane's candidate range over the band's rows (a superset): the
        // bounds are linear in the row, so their extremes are at end rows.
        float lo = 3.0e38f, hi = -3.0e38f;
        if (active) {
            const float rl = static_cast<float>(rmax - 1);
            lo = fminf(fla, fmaf(flb, rl, fla));
            hi = fmaxf(fua, fmaf(fub, rl, fua));
        }
        int ilo, ihi;
        if (hull) {
            // one pixel of margin on each side covers the rounding of the
            // lanes' FP32 bounds about the hull
            const float l0 = lane_value(lo, 0), l1 = lane_value(lo, last);
            const float h0 = lane_value(hi, 0), h1 = lane_value(hi, last);
            ilo = static_cast<int>(floorf(fmaxf(fminf(l0, l1), -2.0f))) - 1;
            ihi = static_cast<int>(ceilf(fminf(fmaxf(h0, h1), static_cast<float>(nw + 1)))) + 1;
            ilo = max(ilo, 0);
            ihi = min(ihi, nw - 1);
            if (!active || ilo > ihi) {
                ilo = 1 << 30;
                ihi = -(1 << 30);
            }
        } else {
            ilo = static_cast<int>(floorf(fmaxf(lo, -1.0f)));
            ihi = static_cast<int>(ceilf(fminf(hi, static_cast<float>(nw))));
            ilo = max(ilo, 0);
            ihi = min(ihi, nw - 1);
            if (ilo > ihi) {
                ilo = 1 << 30;
                ihi = -(1 << 30);
            }
            for (int o = 32; o > 0; o >>= 1) {
                ilo = min(ilo, __shfl_xor(ilo, o));
                ihi = max(ihi, __shfl_xor(ihi, o));
            }
        }
        // previous band's candidates are done with the LDS before this
        // barrier; the reduction slots alternate between bands
        int* red = sred + (band & 1) * 2 * V;
        if (lane == 0) {
            red[2 * wave] = ilo;
            red[2 * wave + 1] = ihi;
        }
        __syncthreads();
        int W0 = red[0], W1 = red[1];
#pragma unroll
        for (int w = 1; w < V; ++w) {
            W0 = min(W0, red[2 * w]);
            W1 = max(W1, red[2 * w + 1]);
        }
        W0 = uniform_i(W0);
        W1 = uniform_i(W1);
        if (W0 > W1) continue;

        for (int cw = W0 & ~(kGatherTile - 1); cw <= W1; cw += WMAX) {
            if (cw != (W0 & ~(kGatherTile - 1))) __syncthreads();  // previous chunk done
            // ---- stage the image: rows u0 .. u0+rmax-1, pixels cw .. cw+WMAX-1.
            // A wave stages whole rows: one buffer descriptor per (row, slice
            // group), then K DMAs that differ only in the pixel block.
            // Only the 16-pixel blocks up to the band's last candidate W1 are
            // staged (no candidate reads past it).
            const int kcnt = min(K, ((min(W1, cw + WMAX - 1) - cw) >> 4) + 1);
            if constexpr (S::kQuad) {
                const int lane_px = lane >> 2;
                const unsigned int lane_off = (lane & 3) * slice_bytes;
                for (int r = wave; r < rmax; r += V) {
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        const int ng = nb - 4 * g;
                        if (ng <= 0) break;
                        const float* base = gsrc[g] + static_cast<size_t>(u0 + r) * nw;
                        const unsigned int range = ng >= 4 ? 4u * slice_bytes : ng * slice_bytes;
                        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                            const_cast<float*>(base), 0, range, 0x00020000);
                        float* dst = simg + (r * S::kRowF4 + g * WMAX) * 4;
                        for (int k = 0; k < kcnt; ++k) {
                            const int px = cw + 16 * k + lane_px;
                            const unsigned int voff =
                                static_cast<unsigned int>(px) < static_cast<unsigned int>(nw)
                                    ? lane_off + static_cast<unsigned int>(px) * 4u
                                    : 0x80000000u;
                            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                                rs, (__attribute__((address_space(3))) void*)(dst + 64 * k), 4,
                                voff, 0, 0, 0);
                        }
                    }
                }
            } else {
                for (int r = wave; r < rmax; r += V) {
                    const float* base = src + (static_cast<size_t>(b0) * nu + u0 + r) * nw;
                    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                        const_cast<float*>(base), 0, static_cast<unsigned int>(nw) * 4u, 0x00020000);
#pragma unroll
                    for (int k = 0; k < WMAX / 64; ++k)
                        __builtin_amdgcn_raw_ptr_buffer_load_lds(
                            rs, (__attribute__((address_space(3))) void*)(simg + r * S::kRowF4 * 4 + 64 * k),
                            4, static_cast<unsigned int>(cw + 64 * k + lane) * 4u, 0, 0, 0);
                }
            }
            // ---- the anchors of the chunk's tiles, for the workgroup's views
            if (tid < V * K) {
                const int v = tid / K, k = tid - v * K;
                const int tu = u0 / kGatherTile, tw = cw / kGatherTile + k;
                const int tw_max = frame ? L.tiles_y : L.tiles_x;
                uint4 q0 = make_uint4(0u, 0u, 0u, 0u), q1 = q0;
                if (v0 + v < P.na && tw < tw_max) {
                    const int tile = frame ? tw * L.tiles_x + tu : tu * L.tiles_x + tw;
                    const uint4* a = reinterpret_cast<const uint4*>(
                        L.anchors + static_cast<size_t>(v0 + v) * L.tiles_x * L.tiles_y + tile);
                    q0 = a[0];
                    q1 = a[1];
                }
                uint4* d = reinterpret_cast<uint4*>(sanc) + 2 * tid;
                d[0] = q0;
                d[1] = q1;
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();

            // ---- this lane's candidates: the row intervals first, then one
            // counted loop over all of them (candidate k -> row r by the
            // prefix counts), so lanes with different rows do not diverge
            const int wlo = max(cw, 0), whi = min(cw + WMAX - 1, nw - 1);
            int dlt[R], pre[R + 1];  // row r: candidates pre[r] .. pre[r+1]-1 at w = kc + dlt[r]
            pre[0] = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float lbr = fmaf(flb, static_cast<float>(r), fla);
                const float ubr = fmaf(fub, static_cast<float>(r), fua);
                const int w0 = max(static_cast<int>(ceilf(fmaxf(lbr, -2.0f))), wlo);
                const int w1 = min(static_cast<int>(floorf(fminf(ubr, static_cast<float>(nw + 1)))), whi);
                const int n = (active && r < rmax) ? max(w1 - w0 + 1, 0) : 0;
                dlt[r] = w0 - pre[r];
                pre[r + 1] = pre[r] + n;
            }
            const int total = pre[R];
            // the band's tile centre across the rows (bands do not straddle
            // a tile row) and along them
            const int uc = min((u0 & ~(kGatherTile - 1)) + kGatherTile / 2, (frame ? P.nx : P.ny) - 1);
            const int wcl = (frame ? P.ny : P.nx) - 1;
            for (int kc = 0; kc < total; ++kc) {
                int r = 0, d = dlt[0];
#pragma unroll
                for (int q = 1; q < R; ++q) {
                    const bool past = kc >= pre[q];
                    r = past ? q : r;
                    d = past ? dlt[q] : d;
                }
                const int w = kc + d;
                // the candidate's image values first: their LDS reads are in
                // flight while the geometry below is computed
                float4 pv[NG];
                float p1 = 0.0f;
                if constexpr (S::kQuad) {
                    const float4* q =
                        reinterpret_cast<const float4*>(simg) + r * S::kRowF4 + (w - cw);
#pragma unroll
                    for (int g = 0; g < NG; ++g) pv[g] = q[g * WMAX];
                } else {
                    p1 = simg[r * S::kRowF4 * 4 + (w - cw)];
                }
                Anchor A{};
                __builtin_memcpy(&A, sanc + 8 * (wave * K + ((w - cw) >> 4)), 32);
                const float fu = static_cast<float>(u0 + r - uc);
                const float fw = static_cast<float>(
                    w - min((w & ~(kGatherTile - 1)) + kGatherTile / 2, wcl));
                float f, wt;
                Geo::pixel(Pg, A, frame ? fu : fw, frame ? fw : fu, f, wt);
                const float fr = __builtin_amdgcn_fractf(f);
                const int s = A.is0 + static_cast<int>(f);
                // the forward's edge rule, only for tiles that reach past the
                // detector (few, and the same for most lanes of a wave)
                if (!A.inside) wt = on_detector(s, fr, P.ns) ? wt : 0.0f;
                const float wm = s == ci ? wt : 0.0f;
                const float w1 = wm * fr, w0 = wm - w1;
                if constexpr (S::kQuad) {
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        accp[4 * g] = fmaf(w0, pv[g].x, accp[4 * g]);
                        accp[4 * g + 1] = fmaf(w0, pv[g].y, accp[4 * g + 1]);
                        accp[4 * g + 2] = fmaf(w0, pv[g].z, accp[4 * g + 2]);
                        accp[4 * g + 3] = fmaf(w0, pv[g].w, accp[4 * g + 3]);
                        accq[4 * g] = fmaf(w1, pv[g].x, accq[4 * g]);
                        accq[4 * g + 1] = fmaf(w1, pv[g].y, accq[4 * g + 1]);
                        accq[4 * g + 2] = fmaf(w1, pv[g].z, accq[4 * g + 2]);
                        accq[4 * g + 3] = fmaf(w1, pv[g].w, accq[4 * g + 3]);
                    }
                } else {
                    accp[0] = fmaf(w0, p1, accp[0]);
                    accq[0] = fmaf(w1, p1, accq[0]);
                }
            }
        }
    }

    // channel ci = P_ci + Q_{ci-1}, the latter from the lane on the left
    // (every lane takes part in the exchange)
    const bool out = lane > 0 && ci >= 0 && active;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const float ql = __shfl_up(accq[b], 1);
        if (out && b < nb)
            L.proj[(static_cast<size_t>(b0 + b) * P.na + view) * P.ns + ci] =
                static_cast<float>(static_cast<double>(accp[b] + ql) * L.scale);
    }
}

// Host side: workspace bytes and the launch sequence (anchor table, the
// transposed image, the gather).  ws: caller-provided workspace of
// gather_workspace_bytes(...) bytes, or null to allocate it on the stream.
template <class Anchor>
inline size_t gather_workspace_bytes(int na, int nx, int ny, int nbatch)
{
    const size_t ntiles = static_cast<size_t>((nx + kGatherTile - 1) / kGatherTile) *
                          ((ny + kGatherTile - 1) / kGatherTile);
    const size_t anc = (na * ntiles * sizeof(Anchor) + 255) & ~static_cast<size_t>(255);
    return anc + static_cast<size_t>(nbatch) * nx * ny * sizeof(float);
}

template <class Geo, int NB, int R, int WMAX>
int launch_gather_nb(GatherLaunch<Geo> L, hipStream_t st)
{
    using S = GatherShape<Geo, NB, R, WMAX>;
    auto kern = bp_gather_t_kernel<Geo, NB, R, WMAX>;
    if (S::kLds > 64 * 1024)
        CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(S::kLds)));
    dim3 grid((L.P.ns + kGatherOut - 1) / kGatherOut, (L.P.na + kGatherViews - 1) / kGatherViews,
              (L.nbatch + NB - 1) / NB);
    hipLaunchKernelGGL(kern, grid, dim3(256), S::kLds, st, L);
    return check_launch("bp_gather_t_kernel");
}

template <class Geo>
int launch_bp_gather_t(const typename Geo::Problem& P, const typename Geo::Tables& T,
                       const float* img, float* proj, int nbatch, double scale, void* ws,
                       hipStream_t st)
{
    using Anchor = typename Geo::Anchor;
    GatherLaunch<Geo> L{};
    L.P = P;
    L.T = T;
    L.tiles_x = (P.nx + kGatherTile - 1) / kGatherTile;
    L.tiles_y = (P.ny + kGatherTile - 1) / kGatherTile;
    L.nbatch = nbatch;
    L.scale = scale;
    L.img = img;
    L.proj = proj;
    const int ntiles = L.tiles_x * L.tiles_y;
    const size_t bytes = gather_workspace_bytes<Anchor>(P.na, P.nx, P.ny, nbatch);
    void* own = nullptr;
    if (!ws) {
        CTR_HIP(hipMallocAsync(&own, bytes, st));
        ws = own;
    }
    auto* table = static_cast<Anchor*>(ws);
    const size_t anc = (static_cast<size_t>(P.na) * ntiles * sizeof(Anchor) + 255) &
                       ~static_cast<size_t>(255);
    float* imgT = reinterpret_cast<float*>(static_cast<unsigned char*>(ws) + anc);
    L.anchors = table;
    L.imgT = imgT;
    hipLaunchKernelGGL(bp_anchor_table_kernel<Geo>, dim3((ntiles + 63) / 64, P.na), dim3(64), 0,
                       st, P, T, L.tiles_x, ntiles, table);
    int rc = check_launch("bp_anchor_table_kernel");
    if (rc == CTR_OK) {
        hipLaunchKernelGGL(image_transpose_kernel, dim3((P.nx + 31) / 32, (P.ny + 31) / 32, nbatch),
                           dim3(32, 8), 0, st, img, imgT, P.ny, P.nx);
        rc = check_launch("image_transpose_kernel");
    }
    if (rc == CTR_OK) {
        // four slices of one image must fit a 32-bit buffer range
        const bool quad_ok = 4ull * P.nx * P.ny * 4 < (1ull << 31);
        if (nbatch >= 16 && quad_ok)
            rc = launch_gather_nb<Geo, 16, 4, 128>(L, st);
        else if (nbatch >= 5 && quad_ok)
            rc = launch_gather_nb<Geo, 8, 4, 128>(L, st);
        else if (nbatch >= 2 && quad_ok)
            rc = launch_gather_nb<Geo, 4, 4, 128>(L, st);
        else
            rc = launch_gather_nb<Geo, 1, 4, 128>(L, st);
    }
    if (own) CTR_HIP(hipFreeAsync(own, st));
    return rc;
}

}  // namespace ctr
