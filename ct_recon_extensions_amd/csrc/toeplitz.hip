// Ramp filter as a Toeplitz GEMM on the gfx950 matrix cores.
//
// The filter of src/filter.cc:67-91,  out[r][n] = ds * sum_k x[r][k] h[n-k],
// has h[m] = 0 for every even m != 0 (src/filter.cc:26-30, :45-49).  Split
// the channels by parity: with ho[q] = h[2q+1] and pi the parity of n,
//     out[r][2j+pi] = ds * ( h0 x[r][2j+pi]  +  sum_i x[r][2i+1-pi] ho[j-i+pi-1] ),
// i.e. for each parity one GEMM  C = A * G  with A the opposite-parity
// columns of the rows (M = rows, K ~ ns/2) and G[i][j] = ho[j-i+pi-1] a
// Toeplitz matrix (N ~ ns/2): half the MACs of the dense Toeplitz product,
// and G is never stored: a lane's B operand is one entry of the tap vector
// kept in LDS.  Fan / flat weighting (src/fan_weighting.cc:19-35) is fused
// into the staging of A.
//
// FP32 mode (ramp_mfma_kernel<false>): v_mfma_f32_16x16x4_f32, FP32 inputs,
// FP32 accumulation.
// BF16x3 mode (ramp_mfma_kernel<true>): v_mfma_f32_16x16x32_bf16 with the
// operands split as x = hi + lo (hi = bf16(x), lo = bf16(x - hi)), and the
// compensated product hi*hi + hi*lo + lo*hi accumulated in FP32 (the lo*lo
// term, ~2^-16 relative, is dropped).
//
// Near field in FP64.  The output is a small difference of large terms: the
// centre tap h0 and the first odd taps carry most of the filter's weight and
// nearly cancel on smooth rows, so FP32 accumulation of them (inside the
// MFMA) would cost ~2e-5 relative error on CT sinograms.  The odd taps
// |m| <= 2 kNear - 1 are therefore zeroed in the GEMM's tap vector and
// evaluated in the epilogue in FP64 (FP64 taps times the same FP32 weighted
// inputs, from an LDS window of the tile's rows), and the three parts are
// combined in FP64; the matrix cores only carry the far field, whose FP32 /
// BF16-split rounding is correspondingly small.
//
// Workgroup: 4 waves, a 64-row x 64-output tile of one parity; each wave a
// 32 x 32 block (2 x 2 MFMA tiles of 16 x 16).  A is staged in LDS in
// K-chunks of 64 (64 x 64 words, padded); the tap vector covers the whole
// row.
#include "ctr_common.hpp"
#include "tile_common.hpp"

#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>

namespace ctr {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kTM = 64, kTN = 64, kKC = 64, kAPad = 4;
// Near-field taps m = 2q + 1, q in [-kNear, kNear - 1] (|m| <= 2 kNear - 1),
// and the per-row window of opposite-parity inputs they read for a tile.
// With the centre, the taps |m| <= 7 carry most of the filter's weight and
// nearly cancel on smooth sinograms, so they are summed in FP64; the rest go
// to the matrix cores.  Worst per-slice relative L2 of the BF16x3 filter on
// 1440 x 1472 Shepp-Logan-type sinograms, against the FP64 filter, by the
// largest FP64 tap |m|: 15 -> 3.0e-6, 7 -> 3.4e-6, 5 -> 5.3e-6, 3 -> 9.8e-6,
// 1 -> 1.5e-5; the 1472-channel filter takes 0.42, 0.39, 0.38, 0.37 and
// 0.37 ms.
constexpr int kNear = 4;
constexpr int kWin = kTN + 2 * kNear;

struct ToeplitzParams {
    const float* x;      // [nrow][ns]
    float* out;          // [nrow][ns]
    const double* taps64;  // FP64 h[m] at taps64[ns + m] (centre and near field)
    const float* taps;   // FP32 h[m] at taps[ns + m], m in [-ns, ns]
    const float* wts;    // FP32 per-channel weight, or nullptr
    const __bf16* bcopies;  // BF16x3: far-field tap copies (bf16_tap_copies_kernel)
    int ns, nrow;
    int n_even, n_odd;   // outputs of each parity
    int tiles_even;      // column tiles of parity 0 (the rest are parity 1)
    int col_tiles;       // column tiles of both parities
    int row_tiles;
    int grid_cols;       // workgroups per row block (tiles, or groups of tiles)
    int wtype;           // fused weighting: 0 none, 1 fan, 2 flat
    double ds, dso, dsd, cents;
};

// Per-channel cosine weight dso |cos(gamma_k)| / dsd of FanWeighting<T>
// (src/fan_weighting.cc:19-35): gamma = s / dsd (fan) or atan2(s, dsd)
// (flat), s = ds (k - cents).
__device__ __forceinline__ float cos_weight(int k, int wtype, double ds, double dso, double dsd,
                                            double cents)
{
    const double s = ds * (static_cast<double>(k) - cents);
    const double c = wtype == 2 ? cos(atan2(s, dsd)) : cos(s / dsd);
    return static_cast<float>(dso * fabs(c) / dsd);
}

// Entry jc of the reversed far-field tap vector: ho[q] = h[2q+1] at
// q = nq - 1 - jc + qmin, zero for the near taps (|q + 1/2| < kNear, carried
// in FP64 by the epilogues) and past the filter's ends.
__device__ __forceinline__ float far_tap_entry(const double* __restrict__ h, int ns, int nq,
                                               int qmin, int jc)
{
    const int tj = nq - 1 - jc;
    if (tj < 0) return 0.0f;
    const int q = tj + qmin, m = 2 * q + 1;
    const bool near = q >= -kNear && q < kNear;
    return (m >= -ns && m <= ns && !near) ? static_cast<float>(h[ns + m]) : 0.0f;
}

// Workgroup -> (row tile bx, column tile by).  The grid is one-dimensional
// and workgroups go to the 8 XCDs round-robin; xcd_tile hands every XCD a
// contiguous run of logical indices in which the column tiles of one row
// block are consecutive, so the block's input rows come from HBM once and
// are re-read from that XCD's L2 by its other column tiles.
__device__ __forceinline__ void toeplitz_tile(const ToeplitzParams& P, int& bx, int& by)
{
    const int l = xcd_tile(blockIdx.x, P.row_tiles * P.grid_cols);
    bx = l / P.grid_cols;
    by = l - bx * P.grid_cols;
}

__device__ __forceinline__ float wt(const ToeplitzParams& P, int k)
{
    return P.wts ? P.wts[k] : 1.0f;
}

// Epilogue shared by the FP32 and BF16x3 kernels: stages the near-field
// window of the tile's rows into win (after a barrier: the caller's operand
// buffers are dead), then writes
//   out[2j+pi] = ds (h0 xw[2j+pi] + sum_q h[2q+1] xw[2(j+pi-1-q)+kpar] + far)
// with the centre and near taps in FP64 and far = the GEMM's FP32 result.
__device__ __forceinline__ void near_field_epilogue(const ToeplitzParams& P, const float* wl,
                                                    float* win, const f32x4 (&acc)[2][2], int r0,
                                                    int j0, int pi, int kpar, int kdim, int nout,
                                                    int wr, int wc, int tid, int lane)
{
    // Near-field window: win[rr][w] = weighted input of parity 1-pi at
    // i = j0 + pi - kNear + w (the A operand's values), zero off the row.
    __syncthreads();  // last chunk consumed: the window reuses operand LDS
    const int iw0 = j0 + pi - kNear;
    for (int e = tid; e < kTM * kWin; e += 256) {
        const int rr = e / kWin, w = e % kWin;
        const int r = r0 + rr, i = iw0 + w;
        float val = 0.0f;
        if (r < P.nrow && i >= 0 && i < kdim) {
            const int k = 2 * i + kpar;
            val = P.x[static_cast<size_t>(r) * P.ns + k] * wl[k];
        }
        win[e] = val;
    }
    __syncthreads();

    // D layout: lane holds D[4*(lane/16) + t][lane%16], t = 0..3.
    // out[2j+pi] = ds (h0 xw[2j+pi] + sum_q h[2q+1] xw[2(j+pi-1-q)+kpar] + far)
    const double h0 = P.taps64[P.ns];
    double hn[2 * kNear];
#pragma unroll
    for (int q = -kNear; q < kNear; ++q) {
        const int m = 2 * q + 1;
        hn[q + kNear] = (m >= -P.ns && m <= P.ns) ? P.taps64[P.ns + m] : 0.0;
    }
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int rr = wr + 16 * ta + 4 * (lane >> 4) + t;
                const int jj = wc + 16 * tb + (lane & 15);
                const int r = r0 + rr, j = j0 + jj;
                if (r < P.nrow && j < nout) {
                    const int n = 2 * j + pi;
                    const size_t o = static_cast<size_t>(r) * P.ns + n;
                    const float xw = P.x[o] * wl[n];
                    // input i = j + pi - 1 - q sits at window slot jj + kNear - 1 - q
                    const float* wrow = win + rr * kWin + jj + kNear - 1;
                    double near = h0 * static_cast<double>(xw);
#pragma unroll
                    for (int q = -kNear; q < kNear; ++q)
                        near = fma(hn[q + kNear], static_cast<double>(wrow[-q]), near);
                    P.out[o] = static_cast<float>(
                        P.ds * (near + static_cast<double>(acc[ta][tb][t])));
                }
            }
}

template <bool BF16X3>
__global__ __launch_bounds__(256) void ramp_mfma_kernel(const ToeplitzParams P)
{
    // taps ho[q] at tap[q - qmin] for every q a tile can touch (zero where
    // |2q+1| > ns): q = j - i + pi - 1 with j < nout + kTN, i < kdim + kKC
    extern __shared__ __align__(16) unsigned char smem[];
    const int qmin = -(P.ns / 2) - kKC - 4;
    const int nq = P.ns + kKC + kTN + 10;
    // LDS: channel weights [ns] | taps [nq] | A chunk [kTM][kKC + kAPad]
    // (the near-field window reuses the A chunk region and what follows)
    float* wl = reinterpret_cast<float*>(smem);
    float* tap = wl + P.ns;
    float* a_lds = tap + ((nq + 3) & ~3);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int bx, by;
    toeplitz_tile(P, bx, by);
    const int r0 = bx * kTM;
    const int pi = by < P.tiles_even ? 0 : 1;
    const int j0 = (pi == 0 ? by : by - P.tiles_even) * kTN;
    const int nout = pi == 0 ? P.n_even : P.n_odd;
    const int kdim = pi == 0 ? P.ns / 2 : (P.ns + 1) / 2;  // inputs of parity 1-pi
    const int kpar = 1 - pi;

    for (int e = tid; e < nq; e += 256) {
        const int q = e + qmin;
        const int m = 2 * q + 1;
        const bool near = q >= -kNear && q < kNear;  // evaluated in the epilogue
        tap[e] = (m >= -P.ns && m <= P.ns && !near) ? P.taps[P.ns + m] : 0.0f;
    }
    for (int e = tid; e < P.ns; e += 256) wl[e] = wt(P, e);

    // A chunk loads (unweighted), issued one chunk ahead of the MFMA work
    constexpr int kPer = kTM * kKC / 256;
    float v[kPer];
    auto load_chunk = [&](int i0) {
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = tid + u * 256;
            const int rr = e / kKC, ii = e % kKC;
            const int r = r0 + rr, i = i0 + ii;
            v[u] = (r < P.nrow && i < kdim)
                       ? P.x[static_cast<size_t>(r) * P.ns + 2 * i + kpar] : 0.0f;
        }
    };
    load_chunk(0);

    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;  // wave block in the tile
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int i0 = 0; i0 < kdim; i0 += kKC) {
        __syncthreads();  // taps / weights written, previous chunk consumed
        // Stage A: rows r0.., inputs i0.. (x[r][2i + kpar] * w[2i + kpar]).
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = tid + u * 256;
            const int ii = e % kKC, i = i0 + ii;
            const float w = i < kdim ? wl[2 * i + kpar] : 0.0f;
            a_lds[(e / kKC) * (kKC + kAPad) + ii] = v[u] * w;
        }
        __syncthreads();
        if (i0 + kKC < kdim) load_chunk(i0 + kKC);  // in flight during the MFMAs
        const int tapbase = j0 + wc - i0 + pi - 1 - qmin;  // G[i][j] = ho[j - i + pi - 1]
        if constexpr (!BF16X3) {
#pragma unroll 4
            for (int ik = 0; ik < kKC; ik += 4) {
                // lane: A[m = lane%16][k = lane/16], B[k = lane/16][n = lane%16]
                const int m = lane & 15, kk = ik + (lane >> 4);
                float a[2], b[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) a[t] = a_lds[(wr + 16 * t + m) * (kKC + kAPad) + kk];
#pragma unroll
                for (int t = 0; t < 2; ++t) b[t] = tap[tapbase + 16 * t + m - kk];
#pragma unroll
                for (int ta = 0; ta < 2; ++ta)
#pragma unroll
                    for (int tb = 0; tb < 2; ++tb)
                        acc[ta][tb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ta], b[tb],
                                                                           acc[ta][tb], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int ik = 0; ik < kKC; ik += 32) {
                // lane: A[m = lane%16][k = 8(lane/16) + e], B[k = 8(lane/16) + e][n = lane%16]
                const int m = lane & 15, kb = ik + 8 * (lane >> 4);
                bf16x8 ahi[2], alo[2], bhi[2], blo[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const float xa = a_lds[(wr + 16 * t + m) * (kKC + kAPad) + kb + e];
                        const __bf16 h = static_cast<__bf16>(xa);
                        ahi[t][e] = h;
                        alo[t][e] = static_cast<__bf16>(xa - static_cast<float>(h));
                        const float xb = tap[tapbase + 16 * t + m - (kb + e)];
                        const __bf16 g = static_cast<__bf16>(xb);
                        bhi[t][e] = g;
                        blo[t][e] = static_cast<__bf16>(xb - static_cast<float>(g));
                    }
                }
#pragma unroll
                for (int ta = 0; ta < 2; ++ta)
#pragma unroll
                    for (int tb = 0; tb < 2; ++tb) {
                        acc[ta][tb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            alo[ta], bhi[tb], acc[ta][tb], 0, 0, 0);
                        acc[ta][tb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            ahi[ta], blo[tb], acc[ta][tb], 0, 0, 0);
                        acc[ta][tb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            ahi[ta], bhi[tb], acc[ta][tb], 0, 0, 0);
                    }
            }
        }
    }

    near_field_epilogue(P, wl, a_lds, acc, r0, j0, pi, kpar, kdim, nout, wr, wc, tid, lane);
}

// BF16x3 kernel with the operand splits precomputed in LDS.  A chunk is
// stored once per chunk as hi / lo BF16 rows (stride kAS, 16-byte rows), so a
// lane's 8 consecutive K values are one ds_read_b128.  A lane's 8 Toeplitz
// values are 8 consecutive taps in decreasing order; the far-field tap
// vector is stored reversed (hi / lo BF16) in 8 copies shifted by one entry,
// and the copy matching the window's alignment gives them as one aligned
// ds_read_b128.  Per K-step each wave issues 8 such reads for 12 MFMAs
// (2 x 2 tiles x hi*hi, hi*lo, lo*hi).
constexpr int kAS = kKC + 8;  // bf16 row stride of the A splits (144 B)

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Entries per tap copy (a multiple of 8 so every copy starts 16-byte aligned).
__host__ __device__ __forceinline__ int rev_len(int nq) { return ((nq + 7) & ~7) + 8; }

// NC tap copies: 8 (each fragment one ds_read_b128) or 4 (two ds_read_b64,
// half the LDS).  NT: column tiles of one parity per workgroup, all fed by
// the same staged A chunks (the row block's input is read once per NT tiles).
template <int NC, int NT>
__global__ __launch_bounds__(256) void ramp_bf16x3_kernel(const ToeplitzParams P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int qmin = -(P.ns / 2) - kKC - 4;
    const int nq = P.ns + kKC + kTN + 10;
    const int nrev = rev_len(nq);
    // LDS: weights [ns] | B copies hi/lo [8][nrev] bf16 (the near window reuses
    // them after the K loop) | A splits hi/lo [kTM][kAS] bf16
    float* wl = reinterpret_cast<float*>(smem);
    unsigned char* bbase = smem + ((static_cast<size_t>(P.ns) * 4 + 15) & ~static_cast<size_t>(15));
    __bf16* bhi = reinterpret_cast<__bf16*>(bbase);
    __bf16* blo = bhi + NC * nrev;
    const size_t bbytes = std::max(static_cast<size_t>(2 * NC) * nrev * 2,
                                   static_cast<size_t>(kTM) * kWin * 4);
    __bf16* ahi = reinterpret_cast<__bf16*>(bbase + ((bbytes + 15) & ~static_cast<size_t>(15)));
    __bf16* alo = ahi + kTM * kAS;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int bx, by;
    toeplitz_tile(P, bx, by);
    const int r0 = bx * kTM;
    // column groups of NT tiles: ceil(tiles_even / NT) of parity 0, then parity 1
    const int groups_even = (P.tiles_even + NT - 1) / NT;
    const int pi = by < groups_even ? 0 : 1;
    const int ntiles_pi = pi == 0 ? P.tiles_even : P.col_tiles - P.tiles_even;
    const int t0 = (pi == 0 ? by : by - groups_even) * NT;  // first tile of the group
    const int ntv = min(NT, ntiles_pi - t0);                  // its valid tiles
    const int nout = pi == 0 ? P.n_even : P.n_odd;
    const int kdim = pi == 0 ? P.ns / 2 : (P.ns + 1) / 2;
    const int kpar = 1 - pi;

    // the tap copies, built once per call by bf16_tap_copies_kernel in this
    // layout, arrive as 16-byte loads
    {
        const uint4* src = reinterpret_cast<const uint4*>(P.bcopies);
        uint4* dst = reinterpret_cast<uint4*>(bbase);
        for (int e = tid; e < NC * nrev / 4; e += 256) dst[e] = src[e];  // 2 NC nrev bf16
    }
    for (int e = tid; e < P.ns; e += 256) wl[e] = wt(P, e);

    constexpr int kPer = kTM * kKC / 256;
    float v[kPer];
    auto load_chunk = [&](int i0) {
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = tid + u * 256;
            const int rr = e / kKC, ii = e % kKC;
            const int r = r0 + rr, i = i0 + ii;
            v[u] = (r < P.nrow && i < kdim)
                       ? P.x[static_cast<size_t>(r) * P.ns + 2 * i + kpar] : 0.0f;
        }
    };
    load_chunk(0);

    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    f32x4 acc[NT][2][2];
#pragma unroll
    for (int tt = 0; tt < NT; ++tt)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[tt][a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int i0 = 0; i0 < kdim; i0 += kKC) {
        __syncthreads();  // B copies / weights written, previous chunk consumed
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = tid + u * 256;
            const int rr = e / kKC, ii = e % kKC, i = i0 + ii;
            const float a = v[u] * (i < kdim ? wl[2 * i + kpar] : 0.0f);
            const __bf16 h = static_cast<__bf16>(a);
            ahi[rr * kAS + ii] = h;
            alo[rr * kAS + ii] = static_cast<__bf16>(a - static_cast<float>(h));
        }
        __syncthreads();
        if (i0 + kKC < kdim) load_chunk(i0 + kKC);  // in flight during the MFMAs
#pragma unroll
        for (int ik = 0; ik < kKC; ik += 32) {
            // lane: A[m = lane%16][k = 8(lane/16) + e], B[k = 8(lane/16) + e][n = lane%16]
            const int m = lane & 15, kb = ik + 8 * (lane >> 4);
            bf16x8 fa_hi[2], fa_lo[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int ro = (wr + 16 * t + m) * kAS + kb;
                fa_hi[t] = *reinterpret_cast<const bf16x8*>(ahi + ro);
                fa_lo[t] = *reinterpret_cast<const bf16x8*>(alo + ro);
            }
#pragma unroll
            for (int tt = 0; tt < NT; ++tt) {
                if (tt >= ntv) break;  // uniform: past the parity's last tile
                // G[i][j] = ho[j - i + pi - 1]
                const int tapbase = (t0 + tt) * kTN + wc - i0 + pi - 1 - qmin;
                bf16x8 fb_hi[2], fb_lo[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    // taps Q, Q-1, ..., Q-7 = reversed entries S .. S+7, S = nq-1-Q
                    const int S = nq - 1 - (tapbase + 16 * t + m - kb);
                    const int c = S & (NC - 1), bo = c * nrev + (S - c);
                    if constexpr (NC == 8) {
                        fb_hi[t] = *reinterpret_cast<const bf16x8*>(bhi + bo);
                        fb_lo[t] = *reinterpret_cast<const bf16x8*>(blo + bo);
                    } else {
                        const bf16x4 h0 = *reinterpret_cast<const bf16x4*>(bhi + bo);
                        const bf16x4 h1 = *reinterpret_cast<const bf16x4*>(bhi + bo + 4);
                        const bf16x4 l0 = *reinterpret_cast<const bf16x4*>(blo + bo);
                        const bf16x4 l1 = *reinterpret_cast<const bf16x4*>(blo + bo + 4);
                        fb_hi[t] = __builtin_shufflevector(h0, h1, 0, 1, 2, 3, 4, 5, 6, 7);
                        fb_lo[t] = __builtin_shufflevector(l0, l1, 0, 1, 2, 3, 4, 5, 6, 7);
                    }
                }
#pragma unroll
                for (int ta = 0; ta < 2; ++ta)
#pragma unroll
                    for (int tb = 0; tb < 2; ++tb) {
                        f32x4& d = acc[tt][ta][tb];
                        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa_lo[ta], fb_hi[tb], d, 0, 0, 0);
                        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa_hi[ta], fb_lo[tb], d, 0, 0, 0);
                        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa_hi[ta], fb_hi[tb], d, 0, 0, 0);
                    }
            }
        }
    }
#pragma unroll
    for (int tt = 0; tt < NT; ++tt) {
        if (tt >= ntv) break;
        near_field_epilogue(P, wl, reinterpret_cast<float*>(bbase), acc[tt], r0, (t0 + tt) * kTN,
                            pi, kpar, kdim, nout, wr, wc, tid, lane);
    }
}

// Whole-row BF16x3 filter for rows of up to 1024 channels: one workgroup
// filters kRows rows.  The rows are loaded once, fan-weighted, into LDS in
// FP32 with kRowPad zeros on the left and at least 16 on the right:
//   xr[r][kRowPad + k] = w[k] x[r0 + r][k].
// Wave w computes output parity pi = w >> 1 and half w & 1 of that parity's
// 16-wide column tiles (TPW of them, the half rounded up to an even count):
// the far field on the matrix cores (row fragments split into BF16 hi / lo
// from the FP32 rows as they are read; tap fragments from the reversed tap
// vector, one word per entry holding its BF16 hi and lo, kept in a register
// ring across K steps), then the centre and near taps in FP64 from the same
// rows.  The results go back into the row buffer and leave as whole rows,
// so input and output move through HBM once each, with contiguous accesses.
constexpr int kRows = 16, kRowPad = 16;

// Row stride (floats): kRowPad + ns + >= 16 zeros (the near field reaches
// 2 kNear - 1 channels past an output), odd, so that the 64 lanes of an A-fragment
// read (16 rows x 4 column groups 16 floats apart) hit 64 different banks.
__host__ __device__ __forceinline__ int row_stride(int ns) { return (kRowPad + ns + 16) | 1; }

// LDS bytes of the whole-row kernel: the rows, then one word per entry of
// the reversed far-field tap vector.
__host__ __device__ __forceinline__ size_t rows_kernel_lds(int ns)
{
    return (static_cast<size_t>(kRows) * row_stride(ns) + rev_len(ns + kKC + kTN + 10)) *
           sizeof(float);
}

// Up to 12 tiles per wave the kernel fits 170 VGPRs, so three workgroups
// share a CU when their LDS allows (rows of up to 763 channels).
template <int TPW>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(TPW <= 12 ? 3 : 2, TPW <= 12 ? 3 : 2))) void
ramp_rows_bf16x3_kernel(const ToeplitzParams P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int ns = P.ns;
    const int qmin = -(ns / 2) - kKC - 4;
    const int nq = ns + kKC + kTN + 10;
    const int nrev = rev_len(nq);
    const int xs = row_stride(ns);
    // LDS: rows [kRows][xs] floats | taps [nrev] words (bf16 hi | bf16 lo << 16).
    // The taps follow the rows, so reads a little past the last row (by
    // lanes whose results are dropped) stay inside the allocation.
    float* xr = reinterpret_cast<float*>(smem);
    unsigned int* bc = reinterpret_cast<unsigned int*>(xr + kRows * xs);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int r0 = blockIdx.x * kRows;

    // the reversed far-field taps, straight from the FP64 taps, split into
    // BF16 hi and lo in one word
    for (int e = tid; e < nrev; e += 256) {
        const float f = far_tap_entry(P.taps64, ns, nq, qmin, e);
        const __bf16 h = static_cast<__bf16>(f);
        const __bf16 l = static_cast<__bf16>(f - static_cast<float>(h));
        bc[e] = static_cast<unsigned int>(__builtin_bit_cast(unsigned short, h)) |
                (static_cast<unsigned int>(__builtin_bit_cast(unsigned short, l)) << 16);
    }
    // Rows in, weighted: channel k = tid + 256 c (ns <= 1024), eight rows'
    // loads in flight at a time; zeros in the margins.
    float wk[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int k = tid + 256 * c;
        wk[c] = k >= ns ? 0.0f
                : P.wtype ? cos_weight(k, P.wtype, P.ds, P.dso, P.dsd, P.cents) : 1.0f;
    }
#pragma unroll
    for (int rb = 0; rb < kRows; rb += 8) {
        float v[8][4];
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int k = tid + 256 * c, row = r0 + rb + r;
                v[r][c] = (k < ns && row < P.nrow) ? P.x[static_cast<size_t>(row) * ns + k] : 0.0f;
            }
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int k = tid + 256 * c;
                if (k < ns) xr[(rb + r) * xs + kRowPad + k] = v[r][c] * wk[c];
            }
    }
    const int npad = xs - ns;  // kRowPad on the left, the rest on the right
    for (int e = tid; e < kRows * npad; e += 256) {
        const int r = e / npad, c = e - r * npad;
        xr[r * xs + (c < kRowPad ? c : ns + c)] = 0.0f;
    }
    __syncthreads();

    const int pi = wave >> 1, half = wave & 1, kpar = 1 - pi;
    const int nout = pi == 0 ? P.n_even : P.n_odd;
    const int kdim = pi == 0 ? ns / 2 : (ns + 1) / 2;
    const int nt = (nout + 15) / 16;
    const int tper = (nt + 1) / 2;
    const int tb0 = half * tper, ntw = min(tper, nt - tb0);  // this wave's tiles
    const int m = lane & 15, g = lane >> 4;

    f32x4 acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    // B fragment of tile t at K step i0: B[k][n] = ho[q], q = j - i + pi - 1,
    // j = 16 (tb0 + t) + n, i = i0 + 8g + e; taps Q, Q-1, ..., Q-7 are the
    // reversed entries S .. S+7, S = nq-1-Q.  Every wave runs TPW tiles (at
    // most one past the parity's last: its taps are still inside the
    // copies, and its results are dropped).
    auto load_b = [&](int t, int i0, bf16x8& fb_hi, bf16x8& fb_lo) {
        const int Q = 16 * (tb0 + t) + m - (i0 + 8 * g) + pi - 1 - qmin;
        const unsigned int* w = bc + (nq - 1 - Q);
        u32x4 hi, lo;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const unsigned int a = w[2 * e], b = w[2 * e + 1];
            hi[e] = (a & 0xffffu) | (b << 16);
            lo[e] = (a >> 16) | (b & 0xffff0000u);
        }
        fb_hi = __builtin_bit_cast(bf16x8, hi);
        fb_lo = __builtin_bit_cast(bf16x8, lo);
    };
    // G is Toeplitz in 16 x 32 blocks: tile t's fragment at step i0 + 32 is
    // tile t - 2's at step i0.  The fragments stay in a ring of TPW register
    // slots (tile t of step s in slot (t - 2s) mod TPW); the K loop is
    // unrolled over the ring's period TPW / 2 so every slot index is static,
    // and a step loads only tiles 0 and 1 (their MFMAs come last in the step,
    // behind the loads' latency).
    constexpr int R = TPW / 2;
    bf16x8 rh[TPW], rl[TPW];
#pragma unroll
    for (int t = 2; t < TPW; ++t) load_b(t, 0, rh[t], rl[t]);
    for (int base = 0; base < kdim; base += 32 * R) {
#pragma unroll
        for (int s = 0; s < R; ++s) {
            const int i0 = base + 32 * s;
            if (i0 < kdim) {  // uniform; once false, false for the rest
                // A[m][k]: input i = i0 + 8g + e of parity kpar in row m
                const float* arow = xr + m * xs + kRowPad + 2 * (i0 + 8 * g) + kpar;
                bf16x8 fa_hi, fa_lo;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    // past kdim the row holds zeros for 16 channels, then the next row
                    const float a = i0 + 8 * g + e < kdim ? arow[2 * e] : 0.0f;
                    const __bf16 h = static_cast<__bf16>(a);
                    fa_hi[e] = h;
                    fa_lo[e] = static_cast<__bf16>(a - static_cast<float>(h));
                }
                load_b(0, i0, rh[(TPW - 2 * s) % TPW], rl[(TPW - 2 * s) % TPW]);
                load_b(1, i0, rh[(TPW + 1 - 2 * s) % TPW], rl[(TPW + 1 - 2 * s) % TPW]);
#pragma unroll
                for (int t = TPW - 1; t >= 0; --t) {
                    const int sl = (TPW + t - 2 * s) % TPW;
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh[sl], fa_lo, acc[t], 0, 0, 0);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rl[sl], fa_hi, acc[t], 0, 0, 0);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh[sl], fa_hi, acc[t], 0, 0, 0);
                }
            }
        }
    }

    // out[n] = ds (h0 xw[n] + sum_q h[2q+1] xw[n-1-2q] + far), n = 2j + pi,
    // centre and |2q+1| <= 2 kNear - 1 in FP64.  The MFMAs take the taps as
    // their first operand, so D[4g + u][m] is row m, column
    // j = 16 (tb0 + t) + 4g + u: a lane's four outputs are consecutive
    // channels of one parity in one row, and their near fields share one
    // window of 2 kNear + 3 opposite-parity inputs (each read and widened
    // once for the four).
    const double h0 = P.taps64[ns];
    double hn[2 * kNear];
#pragma unroll
    for (int q = -kNear; q < kNear; ++q) {
        const int mm = 2 * q + 1;
        hn[q + kNear] = (mm >= -ns && mm <= ns) ? P.taps64[ns + mm] : 0.0;
    }
    constexpr int kWinU = 2 * kNear + 3;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        if (t >= ntw) continue;
        const int jb = 16 * (tb0 + t) + 4 * g;
        // row[2u] is the input at n_u = 2 (jb + u) + pi; the near input
        // n_u - 1 - 2q is row[2u - 1 - 2q] = win[u - q + kNear - 1]
        const float* row = xr + m * xs + kRowPad + 2 * jb + pi;
        double win[kWinU];
#pragma unroll
        for (int i = 0; i < kWinU; ++i) win[i] = static_cast<double>(row[2 * i + 1 - 2 * kNear]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            double near = h0 * static_cast<double>(row[2 * u]);
#pragma unroll
            for (int q = -kNear; q < kNear; ++q)
                near = fma(hn[q + kNear], win[u - q + kNear - 1], near);
            acc[t][u] = static_cast<float>(P.ds * (near + static_cast<double>(acc[t][u])));
        }
    }
    __syncthreads();  // every wave has read the rows: they take the results
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        if (t >= ntw) continue;
        const int jb = 16 * (tb0 + t) + 4 * g;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (jb + u < nout) xr[m * xs + kRowPad + 2 * (jb + u) + pi] = acc[t][u];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
        if (r0 + r >= P.nrow) break;  // uniform
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int k = tid + 256 * c;
            if (k < ns) P.out[static_cast<size_t>(r0 + r) * ns + k] = xr[r * xs + kRowPad + k];
        }
    }
}

// Persistent whole-row BF16x3 filter for rows of 1025..1536 channels (it
// needs more than 512): NW waves per workgroup, each workgroup walking row
// blocks of kRows rows.  The first NW/2 waves compute output parity 0 and the rest parity 1,
// each an equal share of the parity's 16-wide column tiles.
//
// Per block: the rows, fan-weighted, in LDS in FP32 (for the FP64 near field
// and as the staging of the output); the matrix cores take the far field in
// K chunks of 64 inputs, each chunk's opposite-parity inputs split once into
// BF16 hi / lo planes by the whole workgroup (double-buffered, one barrier
// per chunk), so a wave's A fragment is one 16-byte LDS read per plane.  The
// far-field taps are kept as four shifted copies of the reversed tap vector
// (hi and lo planes), so a B fragment of 8 consecutive taps is two aligned
// 8-byte reads per plane, and G's Toeplitz structure lets a wave load only
// two new fragments per K step (a ring of TPW fragments).  The next block's
// rows arrive in registers (U floats per lane: each of the kRows rows as
// U / kRows chunks of NT channels) while the current block computes, and the
// results leave through the row buffer as whole coalesced rows.  NW/4 waves per SIMD (256 / (NW/4) VGPRs each).
constexpr int kChunk = 64;              // inputs per A chunk (two K steps)
constexpr int kAC = kChunk + 8;         // bf16 row stride of an A plane (144 B)
constexpr int kAPlane = kRows * kAC;    // bf16 per plane
// Zero taps ahead of each reversed copy: a wave's last tiles may run up to
// NW/2 tiles past its parity's end (their results are dropped), which reads
// up to 16 NW/2 + 9 entries before the vector's start; those taps
// (|m| > ns) are zero.
constexpr int kTapPre = 144;
constexpr int kCopies = 4;

__host__ __device__ __forceinline__ int persist_copy_len(int ns)
{
    return kTapPre + rev_len(ns + kKC + kTN + 10);
}

// LDS bytes: weights | rows | tap copies (hi, lo) | A planes [2 buffers][2
// parities][hi, lo], each region 16-byte aligned.
__host__ __device__ __forceinline__ size_t persist_kernel_lds(int ns)
{
    const size_t w = (static_cast<size_t>(ns) * 4 + 15) & ~static_cast<size_t>(15);
    const size_t rows = (static_cast<size_t>(kRows) * row_stride(ns) * 4 + 15) & ~static_cast<size_t>(15);
    const size_t taps = static_cast<size_t>(2) * kCopies * persist_copy_len(ns) * 2;
    const size_t a = static_cast<size_t>(8) * kAPlane * 2;
    return w + rows + taps + a;
}

template <int NW, int TPW, int U>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW / 4, NW / 4))) void
ramp_persist_bf16x3_kernel(const ToeplitzParams P, int nblocks)
{
    constexpr int NT = 64 * NW, NP = NW / 2;  // lanes; waves per parity
    extern __shared__ __align__(16) unsigned char smem[];
    const int ns = P.ns;
    const int qmin = -(ns / 2) - kKC - 4;
    const int nq = ns + kKC + kTN + 10;
    const int xs = row_stride(ns);
    const int cl = persist_copy_len(ns);
    unsigned char* p = smem;
    float* wl = reinterpret_cast<float*>(p);
    p += (static_cast<size_t>(ns) * 4 + 15) & ~static_cast<size_t>(15);
    float* xr = reinterpret_cast<float*>(p);
    p += (static_cast<size_t>(kRows) * xs * 4 + 15) & ~static_cast<size_t>(15);
    __bf16* bh = reinterpret_cast<__bf16*>(p);
    __bf16* bl = bh + kCopies * cl;
    __bf16* ap = bl + kCopies * cl;  // [buf][parity kp][hi, lo][kRows][kAC]
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    // tap copies: copy c, position kTapPre + j holds reversed entry j + c
    for (int e = tid; e < kCopies * cl; e += NT) {
        const int c = e / cl, j = e - c * cl - kTapPre;
        const float f = far_tap_entry(P.taps64, ns, nq, qmin, j + c);
        const __bf16 h = static_cast<__bf16>(f);
        bh[e] = h;
        bl[e] = static_cast<__bf16>(f - static_cast<float>(h));
    }
    for (int k = tid; k < ns; k += NT)
        wl[k] = P.wtype ? cos_weight(k, P.wtype, P.ds, P.dso, P.dsd, P.cents) : 1.0f;
    // row margins: kRowPad zeros on the left, the rest of the stride on the
    // right; never written again (results land on channels 0..ns-1 only)
    const int npad = xs - ns;
    for (int e = tid; e < kRows * npad; e += NT) {
        const int r = e / npad, c = e - r * npad;
        xr[r * xs + (c < kRowPad ? c : ns + c)] = 0.0f;
    }

    // lane's elements of a block: row r, channel k = tid + NT c for the kRC
    // chunks c of a row (a lane past the row's end reads the next row's
    // element and does not use it).  One buffer range per block, its extent:
    // elements past the last row read as zero and their stores are dropped;
    // the row and chunk offsets are scalar.
    constexpr int kRC = U / kRows;
    static_assert(kRC * kRows == U && NT * kRC >= 1536, "a row is kRC chunks of NT channels");
    float v[U];
    auto load_block = [&](int blk, int t0) {
        const float* xb = P.x + static_cast<size_t>(blk) * kRows * ns;  // uniform
        const int nel = min(kRows, P.nrow - blk * kRows) * ns;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float*>(xb), 0, static_cast<unsigned int>(nel) * 4u, 0x00020000);
#pragma unroll
        for (int r = 0; r < kRows; ++r)
#pragma unroll
            for (int c = 0; c < kRC; ++c)
                v[r * kRC + c] = __builtin_bit_cast(
                    float, __builtin_amdgcn_raw_buffer_load_b32(rs, t0 * 4, (r * ns + NT * c) * 4, 0));
    };
    int blk = blockIdx.x;
    if (blk < nblocks) load_block(blk, tid);

    const int pi = wave / NP, part = wave - pi * NP, kpar = 1 - pi;
    const int nout = pi == 0 ? P.n_even : P.n_odd;
    const int kdim = pi == 0 ? ns / 2 : (ns + 1) / 2;
    const int nt = (nout + 15) / 16;
    const int tper = (nt + NP - 1) / NP;
    const int tb0 = part * tper, ntw = min(tper, nt - tb0);
    // chunks cover the longer parity's inputs, (ns + 1) / 2
    const int nchunk = ((ns + 1) / 2 + kChunk - 1) / kChunk;

    for (; blk < nblocks; blk += gridDim.x) {
        // An opaque zero: every per-lane address of the block (global
        // offsets, LDS positions, tap reads) is derived from it, so the
        // compiler recomputes them per block instead of hoisting them out of
        // the loop and holding them live through the matrix work.
        int lz = 0;
        asm volatile("" : "+s"(lz));
        const int t0 = tid + lz;
        const int m = (lane + lz) & 15, g = (lane + lz) >> 4;
        __syncthreads();  // the previous block's rows have left the buffer
#pragma unroll
        for (int c = 0; c < kRC; ++c) {
            const int k = t0 + NT * c;
            if (k < ns) {
                const float w = wl[k];
#pragma unroll
                for (int r = 0; r < kRows; ++r) xr[r * xs + kRowPad + k] = v[r * kRC + c] * w;
            }
        }
        __syncthreads();
        if (blk + static_cast<int>(gridDim.x) < nblocks) load_block(blk + gridDim.x, t0);

        // chunk c of both parities into A buffer c & 1: input i = 64 c + ii
        // of parity kp is channel 2 i + kp, zero past the row
        auto split_chunk = [&](int c) {
            __bf16* dst = ap + (c & 1) * 4 * kAPlane;
#pragma unroll
            for (int j = 0; j < 2 * kRows * kChunk / NT; ++j) {
                const int idx = t0 + NT * j;
                const int kp = idx / (kRows * kChunk);
                const int rem = idx - kp * (kRows * kChunk);
                const int r = rem / kChunk, ii = rem - r * kChunk;
                const int ch = 2 * (kChunk * c + ii) + kp;
                const float a = ch < ns ? xr[r * xs + kRowPad + ch] : 0.0f;
                const __bf16 h = static_cast<__bf16>(a);
                __bf16* o = dst + kp * 2 * kAPlane + r * kAC + ii;
                o[0] = h;
                o[kAPlane] = static_cast<__bf16>(a - static_cast<float>(h));
            }
        };

        f32x4 acc[TPW];
#pragma unroll
        for (int t = 0; t < TPW; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        // B fragment of tile t at K step i0: taps Q, Q-1, ..., Q-7 are
        // reversed entries S .. S+7, S = nq-1-Q, read from copy S & 3
        const __bf16* bhb = bh + lz;
        const __bf16* blb = bl + lz;
        auto load_b = [&](int t, int i0, bf16x8& fb_hi, bf16x8& fb_lo) {
            const int Q = 16 * (tb0 + t) + m - (i0 + 8 * g) + pi - 1 - qmin;
            const int S = nq - 1 - Q;
            const int c = S & (kCopies - 1);
            const int o = c * cl + kTapPre + (S - c);
            const bf16x4 h0 = *reinterpret_cast<const bf16x4*>(bhb + o);
            const bf16x4 h1 = *reinterpret_cast<const bf16x4*>(bhb + o + 4);
            const bf16x4 l0 = *reinterpret_cast<const bf16x4*>(blb + o);
            const bf16x4 l1 = *reinterpret_cast<const bf16x4*>(blb + o + 4);
            fb_hi = __builtin_shufflevector(h0, h1, 0, 1, 2, 3, 4, 5, 6, 7);
            fb_lo = __builtin_shufflevector(l0, l1, 0, 1, 2, 3, 4, 5, 6, 7);
        };
        // ring of tap fragments: tile t of step s in slot (t - 2 s) mod TPW;
        // the body is unrolled over LS steps (a whole number of chunks and of
        // ring periods) so every slot index is static
        constexpr int R = TPW / 2;
        constexpr int LS = (R % 2 == 0) ? R : 2 * R;
        bf16x8 rh[TPW], rl[TPW];
#pragma unroll
        for (int t = 2; t < TPW; ++t) load_b(t, 0, rh[t], rl[t]);
        split_chunk(0);
        __syncthreads();
        for (int base = 0; base < 2 * nchunk; base += LS) {
#pragma unroll
            for (int s = 0; s < LS; ++s) {
                const int step = base + s, c = step >> 1;
                if (c < nchunk) {  // uniform over the workgroup
                    if ((s & 1) == 0 && c + 1 < nchunk) split_chunk(c + 1);
                    const int i0 = 32 * step;
                    if (i0 < kdim) {  // uniform over the wave
                        const __bf16* arow = ap + (c & 1) * 4 * kAPlane + kpar * 2 * kAPlane +
                                             m * kAC + 32 * (s & 1) + 8 * g;
                        const bf16x8 fa_hi = *reinterpret_cast<const bf16x8*>(arow);
                        const bf16x8 fa_lo = *reinterpret_cast<const bf16x8*>(arow + kAPlane);
                        load_b(0, i0, rh[(TPW - 2 * (s % R)) % TPW], rl[(TPW - 2 * (s % R)) % TPW]);
                        load_b(1, i0, rh[(TPW + 1 - 2 * (s % R)) % TPW],
                               rl[(TPW + 1 - 2 * (s % R)) % TPW]);
#pragma unroll
                        for (int t = TPW - 1; t >= 0; --t) {
                            const int sl = (TPW + t - 2 * (s % R)) % TPW;
                            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh[sl], fa_lo, acc[t], 0, 0, 0);
                            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rl[sl], fa_hi, acc[t], 0, 0, 0);
                            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh[sl], fa_hi, acc[t], 0, 0, 0);
                        }
                    }
                    if (s & 1) __syncthreads();  // chunk c consumed, c + 1 split
                }
            }
        }

        // near field in FP64, as in ramp_rows_bf16x3_kernel
        const double* t64 = P.taps64 + lz;
        double hn[2 * kNear];
#pragma unroll
        for (int q = -kNear; q < kNear; ++q) {
            const int mm = 2 * q + 1;
            hn[q + kNear] = (mm >= -ns && mm <= ns) ? t64[ns + mm] : 0.0;
        }
        constexpr int kWinU = 2 * kNear + 3;
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
            if (t >= ntw) continue;
            const int jb = 16 * (tb0 + t) + 4 * g;
            const float* row = xr + m * xs + kRowPad + 2 * jb + pi;
            double win[kWinU];
#pragma unroll
            for (int i = 0; i < kWinU; ++i) win[i] = static_cast<double>(row[2 * i + 1 - 2 * kNear]);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                double near = t64[ns] * static_cast<double>(row[2 * u]);
#pragma unroll
                for (int q = -kNear; q < kNear; ++q)
                    near = fma(hn[q + kNear], win[u - q + kNear - 1], near);
                acc[t][u] = static_cast<float>(P.ds * (near + static_cast<double>(acc[t][u])));
            }
        }
        __syncthreads();  // every wave has read the rows: they take the results
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
            if (t >= ntw) continue;
            const int jb = 16 * (tb0 + t) + 4 * g;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (jb + u < nout) xr[m * xs + kRowPad + 2 * (jb + u) + pi] = acc[t][u];
        }
        __syncthreads();
        {
            // buffer stores: those past the block's last row are dropped.  A
            // second opaque zero keeps the fill's row-buffer addresses from
            // being reused here.
            int sz = 0;
            asm volatile("" : "+s"(sz));
            const int t1 = tid + sz;
            float* ob = P.out + static_cast<size_t>(blk) * kRows * ns;  // uniform
            const int nel = min(kRows, P.nrow - blk * kRows) * ns;
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                ob, 0, static_cast<unsigned int>(nel) * 4u, 0x00020000);
#pragma unroll
            for (int c = 0; c < kRC; ++c) {
                const int k = t1 + NT * c;
                if (k < ns) {
#pragma unroll
                    for (int r = 0; r < kRows; ++r)
                        __builtin_amdgcn_raw_buffer_store_b32(
                            __builtin_bit_cast(unsigned int, xr[r * xs + kRowPad + k]), rs, t1 * 4,
                            (r * ns + NT * c) * 4, 0);
                }
            }
        }
    }
}

// The BF16x3 kernel's tap operand: copy c (c < nc), entry j holds the
// reversed far-field tap nq-1-(j+c) split as hi (first nc nrev entries) and
// lo (next nc nrev entries), i.e. exactly its LDS image.
__global__ void bf16_tap_copies_kernel(const double* __restrict__ h64, __bf16* __restrict__ out,
                                       int ns, int nq, int nrev, int qmin, int nc)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nc * nrev) return;
    const int c = e / nrev, j = e - c * nrev;
    const float f = far_tap_entry(h64, ns, nq, qmin, j + c);
    const __bf16 h = static_cast<__bf16>(f);
    out[e] = h;
    out[nc * nrev + e] = static_cast<__bf16>(f - static_cast<float>(h));
}

// FP32 taps and (optionally) per-channel cosine weights from the FP64 taps.
__global__ void toeplitz_prep_kernel(const double* __restrict__ h, float* __restrict__ hf,
                                     float* __restrict__ w, int ns, int wtype, double ds,
                                     double dso, double dsd, double cents)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * ns + 1) hf[i] = static_cast<float>(h[i]);
    if (w && i < ns) w[i] = cos_weight(i, wtype, ds, dso, dsd, cents);
}

// Tap copies of the BF16x3 kernel: 4 (two ds_read_b64 per fragment, 50 KB of
// LDS at 1472 channels) measured 0.33 ms on 8 x 1440 rows of 1472 against
// 0.42 ms for 8 copies (one ds_read_b128, 76 KB: two workgroups per CU).
constexpr int kBf16Copies = 4;
// Column tiles per BF16x3 workgroup: 2 measured fastest on 736- and
// 1472-channel rows (3 and 6 lose occupancy to the extra accumulators).
constexpr int kBf16Tiles = 2;
static int bf16_copies() { return kBf16Copies; }

int ramp_toeplitz(const ctr_geom& g, const float* in, float* out, unsigned int nrow,
                  const double* filter, int weight_type, bool bf16x3, hipStream_t st)
{
    CTR_REQUIRE(g.ns >= 2, CTR_ERR_ARG, "Toeplitz ramp filter: ns must be >= 2");
    const int ns = static_cast<int>(g.ns);
    const int nq = ns + kKC + kTN + 10;
    const int nrev = ((nq + 7) & ~7) + 8;
    const int nc = bf16_copies();
    if (bf16x3 && ns > 1024 && ns <= 1536) {
        // Persistent eight-wave whole-row kernel: one workgroup per CU, row
        // blocks prefetched into registers.  Measured against the tiled
        // kernel on 46080 rows of 1472 channels: 0.53 ms against 0.80 ms;
        // on rows of up to 1024 channels the four-wave kernel below stays
        // faster (0.045 ms against 0.070 ms on 11520 rows of 736).
        ToeplitzParams P{};
        P.x = in;
        P.out = out;
        P.taps64 = filter;
        P.ns = ns;
        P.nrow = static_cast<int>(nrow);
        P.n_even = (ns + 1) / 2;
        P.n_odd = ns / 2;
        P.wtype = weight_type;
        P.ds = g.ds;
        P.dso = g.dso;
        P.dsd = g.dsd;
        P.cents = cent_s(g);
        const size_t lds = persist_kernel_lds(ns);
        // waves per workgroup, tiles per wave, prefetch floats per lane
        const int nw = 8;
        auto pk = ns <= 1280 ? ramp_persist_bf16x3_kernel<8, 10, 48>
                             : ramp_persist_bf16x3_kernel<8, 12, 48>;
        if (lds > 64 * 1024)
            CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(pk),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
        int dev = 0, cus = 0, per_cu = 0;
        CTR_HIP(hipGetDevice(&dev));
        CTR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        CTR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(pk), 64 * nw, lds));
        const int nblocks = static_cast<int>(div_up(nrow, kRows));
        const int grid = std::max(1, std::min(nblocks, cus * std::max(per_cu, 1)));
        hipLaunchKernelGGL(pk, dim3(grid), dim3(64 * nw), lds, st, P, nblocks);
        return check_launch("ramp_persist_bf16x3_kernel");
    }
    if (bf16x3) {
        // Whole rows in LDS while two workgroups fit per CU and a wave's tiles
        // fit its accumulators: the kernel builds its own tap vector and
        // weights (no scratch, no preparation launches).
        const size_t rows_lds = rows_kernel_lds(ns);
        const int n_even = (ns + 1) / 2;
        const int tper = (div_up(n_even, 16) + 1) / 2;
        if (rows_lds <= 80 * 1024 && tper <= 16 && ns <= 1024) {
            ToeplitzParams P{};
            P.x = in;
            P.out = out;
            P.taps64 = filter;
            P.ns = ns;
            P.nrow = static_cast<int>(nrow);
            P.n_even = n_even;
            P.n_odd = ns / 2;
            P.wtype = weight_type;
            P.ds = g.ds;
            P.dso = g.dso;
            P.dsd = g.dsd;
            P.cents = cent_s(g);
            auto rk = tper <= 2    ? ramp_rows_bf16x3_kernel<2>
                      : tper <= 4  ? ramp_rows_bf16x3_kernel<4>
                      : tper <= 6  ? ramp_rows_bf16x3_kernel<6>
                      : tper <= 8  ? ramp_rows_bf16x3_kernel<8>
                      : tper <= 10 ? ramp_rows_bf16x3_kernel<10>
                      : tper <= 12 ? ramp_rows_bf16x3_kernel<12>
                      : tper <= 14 ? ramp_rows_bf16x3_kernel<14>
                                   : ramp_rows_bf16x3_kernel<16>;
            if (rows_lds > 64 * 1024)
                CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(rk),
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            static_cast<int>(rows_lds)));
            hipLaunchKernelGGL(rk, dim3(div_up(nrow, kRows)), dim3(256), rows_lds, st, P);
            return check_launch("ramp_rows_bf16x3_kernel");
        }
    }
    // scratch: FP32 taps [2ns+1] | weights [ns] | (BF16x3) tap copies [16 nrev] bf16
    const size_t fbytes = ((3 * static_cast<size_t>(ns) + 1) * sizeof(float) + 15) & ~static_cast<size_t>(15);
    const size_t cbytes = bf16x3 ? static_cast<size_t>(2 * nc) * nrev * sizeof(__bf16) : 0;
    unsigned char* scratch = nullptr;
    CTR_HIP(hipMallocAsync(reinterpret_cast<void**>(&scratch), fbytes + cbytes, st));
    float* hf = reinterpret_cast<float*>(scratch);
    float* w = weight_type ? hf + 2 * ns + 1 : nullptr;
    __bf16* copies = bf16x3 ? reinterpret_cast<__bf16*>(scratch + fbytes) : nullptr;
    hipLaunchKernelGGL(toeplitz_prep_kernel, dim3(div_up(2 * ns + 1, 256)), dim3(256), 0, st,
                       filter, hf, w, ns, weight_type, g.ds, g.dso, g.dsd, cent_s(g));
    int rc = check_launch("toeplitz_prep_kernel");
    if (rc == CTR_OK && bf16x3) {
        hipLaunchKernelGGL(bf16_tap_copies_kernel, dim3(div_up(nc * nrev, 256)), dim3(256), 0, st,
                           filter, copies, ns, nq, nrev, -(ns / 2) - kKC - 4, nc);
        rc = check_launch("bf16_tap_copies_kernel");
    }
    if (rc == CTR_OK) {
        ToeplitzParams P{};
        P.x = in;
        P.out = out;
        P.taps64 = filter;
        P.taps = hf;
        P.wts = w;
        P.bcopies = copies;
        P.ns = ns;
        P.nrow = static_cast<int>(nrow);
        P.n_even = (ns + 1) / 2;
        P.n_odd = ns / 2;
        P.tiles_even = div_up(P.n_even, kTN);
        P.ds = g.ds;
        const size_t nq4 = (static_cast<size_t>(ns) + kKC + kTN + 10 + 3) & ~static_cast<size_t>(3);
        size_t lds = (ns + nq4 + std::max(static_cast<size_t>(kTM) * (kKC + kAPad),
                                          static_cast<size_t>(kTM) * kWin)) * sizeof(float);
        if (bf16x3) {
            const size_t bbytes = std::max(static_cast<size_t>(2 * nc) * nrev * 2,
                                           static_cast<size_t>(kTM) * kWin * 4);
            lds = ((static_cast<size_t>(ns) * 4 + 15) & ~static_cast<size_t>(15)) +
                  ((bbytes + 15) & ~static_cast<size_t>(15)) +
                  static_cast<size_t>(2) * kTM * kAS * 2;
        }
        CTR_REQUIRE(lds <= 160 * 1024, CTR_ERR_SHAPE, "Toeplitz ramp filter: ns too large");
        auto kern = bf16x3 ? ramp_bf16x3_kernel<kBf16Copies, kBf16Tiles> : ramp_mfma_kernel<false>;
        if (lds > 64 * 1024)
            CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
        P.col_tiles = P.tiles_even + div_up(P.n_odd, kTN);
        P.row_tiles = static_cast<int>(div_up(nrow, kTM));
        // BF16x3: one workgroup per group of kBf16Tiles column tiles of a parity
        P.grid_cols = bf16x3 ? div_up(P.tiles_even, kBf16Tiles) +
                                   div_up(P.col_tiles - P.tiles_even, kBf16Tiles)
                             : P.col_tiles;
        const dim3 grid(static_cast<unsigned int>(P.row_tiles) * P.grid_cols);
        hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, P);
        rc = check_launch("ramp_mfma_kernel");
    }
    CTR_HIP(hipFreeAsync(scratch, st));
    return rc;
}

}  // namespace ctr

extern "C" int ctr_ramp_toeplitz(const ctr_geom* g, int dtype, const void* in, void* out,
                                 unsigned int nrow_total, const double* filter, int weight_type,
                                 int bf16x3, void* stream)
{
    CTR_REQUIRE(g && in && out && filter, CTR_ERR_ARG, "ctr_ramp_toeplitz: null argument");
    CTR_REQUIRE(dtype == CTR_F32, CTR_ERR_UNSUPPORTED,
                "ctr_ramp_toeplitz: float32 only (use ctr_ramp for float64)");
    CTR_REQUIRE(weight_type >= 0 && weight_type <= 2, CTR_ERR_ARG,
                "ctr_ramp_toeplitz: weight_type must be 0, 1 or 2");
    if (nrow_total == 0) return CTR_OK;
    return ctr::ramp_toeplitz(*g, static_cast<const float*>(in), static_cast<float*>(out),
                              nrow_total, filter, weight_type, bf16x3 != 0,
                              static_cast<hipStream_t>(stream));
}
