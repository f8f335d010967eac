// Pixel-driven backprojection engine shared by the fan- and parallel-beam
// operators (and the ragged batch): forward kernel and view-parallel scatter
// transpose, parameterised by a geometry policy Geo that supplies
//   Geo::Problem, Geo::Tables, Geo::Anchor (16-byte multiple, field `is0`
//   = first staged channel, `nch` = channels from is0 the tile's taps can
//   touch, `inside` = tile strictly inside the detector),
//   Geo::anchor(P, T, frame, view, ilo, ihi, jlo, jhi)   FP64 -> FP32 anchor
//   Geo::pixel(P, A, fi, fj, f, w)   channel offset f (from is0) and weight.
//
// Forward layout choices (gfx950):
//  * one workgroup = one pixel tile x NB slices of a batch that share the
//    geometry; the geometry of a pixel-view is evaluated once for all NB;
//  * per chunk of views the detector span each tile touches (nch channels,
//    LROW at most) is staged into LDS as spans[view][slice][LROW channels]
//    (or slice-interleaved, below) straight from global memory
//    with buffer loads that write LDS directly (no VGPRs, no per-row wait;
//    the buffer range check returns 0 for channels off the detector, which
//    is the zero padding the edge rule needs), double-buffered so the next
//    chunk's loads are in flight while the current chunk is computed;
//  * a pixel reads its two taps for two slices at a time with
//    ds_read2st64_b32 (rows are LROW = 64k words apart), and a tap pair
//    contributes w0*p[idx] + w1*p[idx+1] per slice (w0 = w(1-frac),
//    w1 = w frac, shared by the slices): two FMAs;
//  * sums are two-level FP32: within a chunk of 64 views, then across
//    chunks (on the fan C2 data this adds ~1e-7 relative error, against
//    ~9e-8 for FP64 carries, and frees the registers for a fourth wave).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "tile_common.hpp"

namespace ctr {

__device__ __forceinline__ float uniform_f(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ int uniform_i(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Views per LDS round for the forward kernel: the most (up to one anchor
// chunk) whose two span buffers and the anchors fit a quarter of a CU's
// 160 KB, i.e. four workgroups per CU (fewer views per round means a barrier
// more often; more LDS per workgroup, fewer workgroups to hide latency).
template <class Anchor>
inline int engine_vchunk(int nb, int lrow)
{
    const int avail = 160 * 1024 / 4 - kAnchorChunk * static_cast<int>(sizeof(Anchor));
    const int v = avail / (2 * nb * lrow * static_cast<int>(sizeof(float)));
    return v < 1 ? 1 : v > kAnchorChunk ? kAnchorChunk : v;
}

template <class Geo>
struct BpLaunch {
    typename Geo::Problem P;
    int nbatch, tiles_x, tiles_y;
    int lmax;    // channels a (view, slice) span needs (<= LROW)
    int vchunk;  // views per LDS buffer (0: engine_vchunk)
    double scale;
    const float* proj;
    float* img;
};

// Issue the DMA loads of the spans of views [v, v+nv) into buf:
// row (t, b) <- proj[b0+b][v+t][is0(t) .. is0(t)+LROW).  Each wave stages
// whole views: one anchor read per view, then NB row loads whose descriptors
// differ by a uniform stride.
template <class Geo, int NB, int LROW, int NWAVES>
__device__ __forceinline__ void stage_spans(const BpLaunch<Geo>& L,
                                            const typename Geo::Anchor* anchors, float* buf,
                                            int v, int vrel, int nv, int b0, int nb, int wave,
                                            int lane)
{
    const auto& P = L.P;
    const unsigned int row_bytes = static_cast<unsigned int>(P.ns) * 4u;
    const size_t slice_stride = static_cast<size_t>(P.na) * P.ns;
    for (int t = wave; t < nv; t += NWAVES) {
        const int is0 = uniform_i(anchors[vrel + t].is0);
        // only the 64-channel rows this view's taps can reach
        const int nrow = min(LROW / 64, (uniform_i(anchors[vrel + t].nch) + 63) / 64);
        const int voff = (is0 + lane) * 4;
        const float* src = L.proj + (static_cast<size_t>(b0) * P.na + v + t) * P.ns;
        float* dst = buf + t * NB * LROW;
#pragma unroll 2
        for (int b = 0; b < NB; ++b) {
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<float*>(src), 0, b < nb ? row_bytes : 0u, 0x00020000);
            for (int c = 0; c < nrow; ++c)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rs, (__attribute__((address_space(3))) void*)(dst + 64 * c), 4,
                    voff + c * 256, 0, 0, 0);
            src += slice_stride;
            dst += LROW;
        }
    }
}

// Slice-interleaved staging, used when NB is a multiple of 4: a view's rows
// are stored as [NB/4 groups][LROW channels][4 slices], so one 16-byte LDS
// read returns a tap for four slices (gfx950 serves ds_read_b128 at ~1.6x
// the bytes per clock of ds_read2st64_b32).  The DMA still writes LDS
// lane-linearly (lane i -> word i of a 64-word block), so lane i fetches
// slice i & 3 of channel i >> 2: per-lane global offsets from a descriptor
// based at the group's first slice.  Zero fill: a slice past the batch has
// an offset beyond the group's range; a channel off the detector gets an
// offset of 2^31, beyond every range (the host keeps 4 slices < 2^31 bytes).
template <class Geo, int NB, int LROW, int NWAVES>
__device__ __forceinline__ void stage_spans_quad(const BpLaunch<Geo>& L,
                                                 const typename Geo::Anchor* anchors, float* buf,
                                                 int v, int vrel, int nv, int b0, int nb,
                                                 int wave, int lane)
{
    static_assert(NB % 4 == 0 && LROW % 16 == 0, "four slices x 16 channels per DMA row");
    constexpr int NC = LROW / 16;
    const auto& P = L.P;
    const unsigned int slice_bytes = static_cast<unsigned int>(P.na) * P.ns * 4u;
    const size_t group_stride = 4 * static_cast<size_t>(P.na) * P.ns;
    const unsigned int lane_off = (lane & 3) * slice_bytes;
    const int lane_ch = lane >> 2;
    for (int t = wave; t < nv; t += NWAVES) {
        const int is0 = uniform_i(anchors[vrel + t].is0);
        // only the 16-channel rows this view's taps can reach (about half of
        // LROW on typical fan and parallel geometries)
        const int nrow = min(NC, (uniform_i(anchors[vrel + t].nch) + 15) / 16);
        const bool span_in = is0 >= 0 && is0 + 16 * nrow <= P.ns;
        unsigned int voff[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int ch = is0 + 16 * c + lane_ch;
            voff[c] = (span_in || static_cast<unsigned int>(ch) < static_cast<unsigned int>(P.ns))
                          ? lane_off + static_cast<unsigned int>(ch) * 4u
                          : 0x80000000u;
        }
        const float* src = L.proj + (static_cast<size_t>(b0) * P.na + v + t) * P.ns;
        float* dst = buf + t * NB * LROW;
#pragma unroll
        for (int g = 0; g < NB / 4; ++g) {
            const int ng = nb - 4 * g;
            const unsigned int range = ng <= 0 ? 0u : ng >= 4 ? 4u * slice_bytes : ng * slice_bytes;
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<float*>(src), 0, range, 0x00020000);
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (c < nrow)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                        rs, (__attribute__((address_space(3))) void*)(dst + 64 * c), 4, voff[c],
                        0, 0, 0);
            src += group_stride;
            dst += 4 * LROW;
        }
    }
}

template <class Geo, int NB, int LROW, int NWAVES>
__device__ __forceinline__ void stage_views(const BpLaunch<Geo>& L,
                                            const typename Geo::Anchor* anchors, float* buf,
                                            int v, int vrel, int nv, int b0, int nb, int wave,
                                            int lane)
{
    if constexpr (NB % 4 == 0)
        stage_spans_quad<Geo, NB, LROW, NWAVES>(L, anchors, buf, v, vrel, nv, b0, nb, wave, lane);
    else
        stage_spans<Geo, NB, LROW, NWAVES>(L, anchors, buf, v, vrel, nv, b0, nb, wave, lane);
}

// Geometry of one pixel-view in the form the accumulation needs: LDS word
// index of the left tap and the two tap weights.  Interpolating policies
// (the default) give a channel position and a weight and get the linear
// interpolation taps, zeroed off the detector; a policy with
// kCustomTaps = true computes its own two taps (Geo::taps) and relies on the
// zero padding of the staged span outside [0, ns).
template <class Geo>
struct HasCustomTaps {
    template <class G>
    static constexpr bool test(decltype(G::kCustomTaps)*) { return G::kCustomTaps; }
    template <class G>
    static constexpr bool test(...) { return false; }
    static constexpr bool value = test<Geo>(nullptr);
};

// Optional Geo::pin(Problem&): moves the per-pixel constants the policy
// combines with another uniform operand into VGPRs once per workgroup (a
// VALU instruction reads at most one SGPR, so otherwise each use costs a
// v_mov per pixel-view).
template <class Geo>
struct HasPin {
    template <class G>
    static constexpr bool test(decltype(&G::pin)) { return true; }
    template <class G>
    static constexpr bool test(...) { return false; }
    static constexpr bool value = test<Geo>(nullptr);
};

// EDGE: the view's tile may reach past the detector (anchor field inside
// = 0), so the weight of a tap off [0, ns-1] is zeroed; chunks whose views
// all have inside = 1 skip the test.
template <class Geo, int LROW, bool EDGE>
__device__ __forceinline__ void tap_weights(const typename Geo::Problem& P,
                                            const typename Geo::Anchor& A, float fi, float fj,
                                            int& idx, float& w0, float& w1)
{
    if constexpr (HasCustomTaps<Geo>::value) {
        Geo::taps(P, A, fi, fj, idx, w0, w1);
        idx = min(max(idx, 0), LROW - 2);
    } else {
        float f, w;
        Geo::pixel(P, A, fi, fj, f, w);
        // f >= 1 for every pixel of the tile (is0 sits a channel below the
        // tile's lowest position): truncation is the floor, and v_fract_f32
        // gives f - floor(f) exactly
        const float fr = __builtin_amdgcn_fractf(f);
        idx = static_cast<int>(f);
        if constexpr (EDGE) {
            const int inside = uniform_i(A.inside), is0 = uniform_i(A.is0);
            if (!inside) w = on_detector(is0 + idx, fr, P.ns) ? w : 0.0f;
            // keeps the reads inside the span for any geometry
            idx = min(max(idx, 0), LROW - 2);
        }
        // Without the edge test every anchor of the chunk is finite with
        // inside = 1, and the tile's positions lie within ~1e-5 channel of
        // [smin, smax]: f >= 1 - eps and f < nch - 2 + eps with
        // nch = floor(smax) + 3 - is0 <= lmax <= LROW (the launchers' span
        // bound), so 0 <= idx <= LROW - 2 needs no clamp.
        w1 = w * fr;
        w0 = w - w1;
    }
}

// Taps and weights of one pixel-view: LDS word index of the left tap and
// the two weights, for each of a thread's PPT pixels.
template <int PPT>
struct ViewTaps {
    int idx[PPT];
    float w0[PPT], w1[PPT];
};

// One view of the software pipeline: read the taps of this view (cur,
// computed one step earlier) from sview, compute the next view's taps into
// nxt while the reads are in flight (when has_next), then accumulate.  The
// caller alternates two ViewTaps, so nothing is copied between steps.
template <class Geo, int NB, int LROW, int PPT, bool EDGE>
__device__ __forceinline__ void bp_view(const typename Geo::Problem& Pg,
                                        const typename Geo::Anchor* an_next, bool has_next,
                                        const float* sview, const float (&fi)[PPT],
                                        const float (&fj)[PPT], const ViewTaps<PPT>& cur,
                                        ViewTaps<PPT>& nxt, float (&acc)[PPT][NB])
{
    using Anchor = typename Geo::Anchor;
    // LDS returns in order: the next view's anchor is read before this
    // view's taps (the empty asm keeps the compiler from sinking the read
    // below them), so the next geometry need not wait for the taps.
    const Anchor An = *an_next;
    asm volatile("" ::: "memory");
    float p0[PPT][NB], p1[PPT][NB];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        if constexpr (NB % 4 == 0) {
            // [group][channel][4 slices]: one ds_read_b128 per tap
            // and group, all at immediate offsets from one address
            const float4* q = reinterpret_cast<const float4*>(sview) + cur.idx[k];
#pragma unroll
            for (int g = 0; g < NB / 4; ++g) {
                const float4 a = q[g * LROW], c = q[g * LROW + 1];
                p0[k][4 * g] = a.x;
                p0[k][4 * g + 1] = a.y;
                p0[k][4 * g + 2] = a.z;
                p0[k][4 * g + 3] = a.w;
                p1[k][4 * g] = c.x;
                p1[k][4 * g + 1] = c.y;
                p1[k][4 * g + 2] = c.z;
                p1[k][4 * g + 3] = c.w;
            }
        } else {
            // Right taps come through a second base whose +1 the
            // compiler cannot see, so the loads pair up per tap
            // across slices (rows LROW words apart:
            // ds_read2st64_b32, no per-slice address arithmetic).
            int one = 1;
            asm volatile("" : "+v"(one));
            const float* q0 = sview + cur.idx[k];
            const float* q1 = sview + cur.idx[k] + one;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                p0[k][b] = q0[b * LROW];
                p1[k][b] = q1[b * LROW];
            }
        }
    }
    if (has_next) {
#pragma unroll
        for (int k = 0; k < PPT; ++k)
            tap_weights<Geo, LROW, EDGE>(Pg, An, fi[k], fj[k], nxt.idx[k], nxt.w0[k],
                                         nxt.w1[k]);
    }
#pragma unroll
    for (int k = 0; k < PPT; ++k)
#pragma unroll
        for (int b = 0; b < NB; ++b)
            acc[k][b] = fmaf(cur.w1[k], p1[k][b], fmaf(cur.w0[k], p0[k][b], acc[k][b]));
}

// One chunk of views of a tile: issue the DMA of the next chunk into `next`,
// then accumulate the views of this chunk (staged in `sbuf`) into acc.  The
// three LDS regions are disjoint, and the restrict qualifiers say so: the
// compiler then knows the span reads cannot depend on the DMA still in
// flight into `next` and inserts no wait for it.
template <class Geo, int NB, int LROW, int NWAVES, int PPT, bool EDGE>
__device__ __forceinline__ void bp_chunk(const BpLaunch<Geo>& L, const typename Geo::Problem& Pg,
                                         const typename Geo::Anchor* __restrict__ anchors,
                                         float* __restrict__ next, const float* __restrict__ sbuf,
                                         bool prefetch, int va, int vc, int nv, int nv_next,
                                         int b0, int nb, int wave, int lane,
                                         const float (&fi)[PPT], const float (&fj)[PPT],
                                         float (&acc)[PPT][NB])
{
    if (prefetch)
        stage_views<Geo, NB, LROW, NWAVES>(L, anchors, next, va + vc + L.vchunk, vc + L.vchunk,
                                           nv_next, b0, nb, wave, lane);
    // Software pipeline over the views of the chunk, two views per trip:
    // the taps of view t are read while the geometry of view t+1 is computed.
    ViewTaps<PPT> ta, tb;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        tap_weights<Geo, LROW, EDGE>(Pg, anchors[vc], fi[k], fj[k], ta.idx[k], ta.w0[k],
                                     ta.w1[k]);
        tb.idx[k] = 0;
        tb.w0[k] = tb.w1[k] = 0.0f;
    }
    // The anchor after a chunk's last view is read but not used: index
    // vc + nv <= kAnchorChunk, at worst the first words of the span buffers.
    const typename Geo::Anchor* an = anchors + vc;
    for (int t = 0; t < nv; t += 2) {
        bp_view<Geo, NB, LROW, PPT, EDGE>(Pg, an + t + 1, t + 1 < nv, sbuf + t * NB * LROW, fi,
                                          fj, ta, tb, acc);
        if (t + 1 < nv)
            bp_view<Geo, NB, LROW, PPT, EDGE>(Pg, an + t + 2, t + 2 < nv,
                                              sbuf + (t + 1) * NB * LROW, fi, fj, tb, ta, acc);
    }
}

// The views of one anchor chunk [va, va + nva), staged chunk by chunk
// through the two span buffers (the first chunk's loads already issued).
template <class Geo, int NB, int LROW, int NWAVES, int PPT, bool EDGE>
__device__ __forceinline__ void bp_anchor_chunk(const BpLaunch<Geo>& L,
                                                const typename Geo::Problem& Pg,
                                                const typename Geo::Anchor* anchors, float* spans,
                                                int va, int nva, int b0, int nb, int wave,
                                                int lane, const float (&fi)[PPT],
                                                const float (&fj)[PPT], float (&acc)[PPT][NB])
{
    const int bufsz = L.vchunk * NB * LROW;  // floats per buffer
    int cur = 0;
    for (int vc = 0; vc < nva; vc += L.vchunk) {
        const int nv = min(L.vchunk, nva - vc);
        __builtin_amdgcn_s_waitcnt(0);  // this wave's span loads have landed
        __syncthreads();                // ... and every other wave's
        // prefetch the next chunk into the other buffer, accumulate this one
        bp_chunk<Geo, NB, LROW, NWAVES, PPT, EDGE>(
            L, Pg, anchors, spans + (cur ^ 1) * bufsz, spans + cur * bufsz, vc + L.vchunk < nva,
            va, vc, nv, min(L.vchunk, nva - vc - L.vchunk), b0, nb, wave, lane, fi, fj, acc);
        cur ^= 1;
    }
}

// Pixel (px, py) of a TW-wide tile handled by thread index p.  A wave's 64
// lanes cover TW x (64 / TW) pixels, and each 16-lane quarter (the lanes one
// 16-byte LDS read serves per pass) takes a compact block: 4 x 4 pixels
// (8 x 2 for TW = 32) rather than a 16-pixel row, so the channels its taps
// touch stay within 16 of each other and do not collide on LDS banks.
template <int TW>
__device__ __forceinline__ void tile_pixel(int p, int& px, int& py)
{
    constexpr int kRowsPerWave = 64 / TW;  // TW in {8, 16, 32}
    constexpr int kBw = kRowsPerWave >= 4 ? 4 : 8, kBh = 16 / kBw;
    const int wave = p >> 6, l = p & 63, q = l >> 4, w = l & 15;
    constexpr int kBlocksX = TW / kBw;
    px = (q % kBlocksX) * kBw + (w % kBw);
    py = wave * kRowsPerWave + (q / kBlocksX) * kBh + w / kBw;
}

// One pixel tile x NB slices (starting at slice b0) over all views: the
// whole forward computation of one workgroup.  smem holds kAnchorChunk
// anchors followed by two span buffers of vchunk x NB x LROW floats.  The
// first LDS write comes after a barrier, so a persistent caller may run
// tiles back to back.
template <class Geo, int TW, int TH, int NB, int LROW>
__device__ __forceinline__ void bp_tile(const typename Geo::Tables& T, const BpLaunch<Geo>& L,
                                        int tile, int b0, unsigned char* smem)
{
    using TS = TileShape<TW, TH>;
    using Anchor = typename Geo::Anchor;
    constexpr int BLOCK = TS::kBlock, PPT = TS::kPPT, NWAVES = BLOCK / 64;
    static_assert(NB % 4 == 0 ? LROW % 16 == 0 : LROW % 64 == 0,
                  "span rows are whole DMA rows (16 channels x 4 slices, or 64 channels)");
    Anchor* anchors = reinterpret_cast<Anchor*>(smem);
    float* spans = reinterpret_cast<float*>(smem + kAnchorChunk * sizeof(Anchor));

    const auto& P = L.P;
    const int tid = threadIdx.x, lane = tid & 63, wave = uniform_i(tid >> 6);
    const TileFrame F(tile, L.tiles_x, P.nx, P.ny, TW, TH);
    const int nb = min(NB, L.nbatch - b0);
    const int ilo = F.x0 - F.ic, ihi = F.x0 + TW - 1 - F.ic;
    const int jlo = F.y0 - F.jc, jhi = F.y0 + TH - 1 - F.jc;

    typename Geo::Problem Pg = P;
    if constexpr (HasPin<Geo>::value) Geo::pin(Pg);
    float fi[PPT], fj[PPT];
    float total[PPT][NB];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        int px, py;
        tile_pixel<TW>(tid + k * BLOCK, px, py);
        fi[k] = static_cast<float>(F.x0 + px - F.ic);
        fj[k] = static_cast<float>(F.y0 + py - F.jc);
#pragma unroll
        for (int b = 0; b < NB; ++b) total[k][b] = 0.0f;
    }

    for (int va = 0; va < P.na; va += kAnchorChunk) {
        const int nva = min(kAnchorChunk, P.na - va);
        __syncthreads();  // previous anchors and spans are no longer read
        if (tid < nva) anchors[tid] = Geo::anchor(P, T, F, va + tid, ilo, ihi, jlo, jhi);
        __syncthreads();
        stage_views<Geo, NB, LROW, NWAVES>(L, anchors, spans, va, 0, min(L.vchunk, nva), b0, nb,
                                           wave, lane);
        float acc[PPT][NB];
#pragma unroll
        for (int k = 0; k < PPT; ++k)
#pragma unroll
            for (int b = 0; b < NB; ++b) acc[k][b] = 0.0f;

        // chunks whose views all keep the tile strictly inside the detector
        // (most of them) skip the per-tap edge test
        if (!HasCustomTaps<Geo>::value && __ballot(lane < nva && !anchors[lane].inside) != 0)
            bp_anchor_chunk<Geo, NB, LROW, NWAVES, PPT, true>(L, Pg, anchors, spans, va, nva, b0,
                                                              nb, wave, lane, fi, fj, acc);
        else
            bp_anchor_chunk<Geo, NB, LROW, NWAVES, PPT, false>(L, Pg, anchors, spans, va, nva,
                                                               b0, nb, wave, lane, fi, fj, acc);
#pragma unroll
        for (int k = 0; k < PPT; ++k)
#pragma unroll
            for (int b = 0; b < NB; ++b) total[k][b] += acc[k][b];
    }

#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        int px, py;
        tile_pixel<TW>(tid + k * BLOCK, px, py);
        const int ix = F.x0 + px, iy = F.y0 + py;
        if (ix < P.nx && iy < P.ny) {
#pragma unroll
            for (int b = 0; b < NB; ++b)
                if (b < nb)
                    L.img[(static_cast<size_t>(b0 + b) * P.ny + iy) * P.nx + ix] =
                        static_cast<float>(static_cast<double>(total[k][b]) * L.scale);
        }
    }
}

template <class Geo, int TW, int TH, int NB, int LROW>
__global__ __launch_bounds__((TileShape<TW, TH>::kBlock))
__attribute__((amdgpu_waves_per_eu(4, 4))) void bp_forward_kernel(
    const typename Geo::Tables T, const BpLaunch<Geo> L)
{
    extern __shared__ __align__(16) unsigned char smem[];
    // grid: super_grid(tiles_x, tiles_y) x slice groups (supertile order)
    int group;
    const int tile = super_tile(L.tiles_x, L.tiles_y, group);
    if (tile < 0) return;  // padding of the supertile grid: the whole workgroup
    bp_tile<Geo, TW, TH, NB, LROW>(T, L, tile, group * NB, smem);
}

}  // namespace ctr
