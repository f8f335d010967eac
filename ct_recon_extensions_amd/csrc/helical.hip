// Helical rebinning for gfx950: HelicalInterpolation360<T> and
// HelicalInterpolation180<T> (src/helical_interpolation.cc:81-160 and
// 162-269, driven per slice by :336-365), bit-identical to the reference's
// CPU path in float32 and float64.
//
// Input: a multi-row helical scan in[b][na][nt][ns] (view, detector row,
// channel); output: one rotation of fan-beam views per slice,
// out[b][nz][view_per_rot][ns].  For output sample (slice z, view iv,
// channel is) with fan angle g = (is - cs) ds / dsd, the detector rows a slice
// of the given thickness reaches in a view whose source sits z_dif below the
// slice are
//     [floor(ct - (z_dif + th/2) dsd cos g / dso), ceil(ct - (z_dif - th/2) dsd cos g / dso))
// clamped to the detector; a window entirely off the detector is unused.
// The window's mean is taken in the output view (cur) and in a conjugate
// view (cor): the same angle one rotation away for 360LI; for 180LI the
// opposite ray, view cur - (pi - 2g)/orbit wrapped into the slice's
// rotation, read with linear interpolation in view and in the mirrored
// channel cs - s/ds.  The two are interpolated linearly to the slice,
//     out = (z_dif cor - z_dif2 cur) / (z_dif - z_dif2),
// or the one usable value is taken.
//
// One wavefront per output row, its lanes on 64 consecutive channels, so
// every row read of the output view is one contiguous line.  The arithmetic
// is the reference's FP64 sequence with contraction off; the terms that
// depend only on the channel (cos(g): the one transcendental, from the
// host's libm) or only on the row are evaluated once and reused, which
// leaves the values unchanged.  The FP64 divisions that remain per sample
// (window means, the z interpolation; for 180LI the conjugate window) bound
// the kernel ahead of its HBM reads.
#include "ctr_common.hpp"
#include "tile_common.hpp"

#include <cmath>

#pragma clang fp contract(off)

namespace ctr {
namespace {

struct HeliProblem {
    double ds, center_s, center_t, couch_begin, couch_mov, half_th, dso, dsd, orbit;
    double nt_d, vpr_d, ns1_d;  // nt, view_per_rot and ns - 1 as the reference compares them
    int ns, nt, na, vpr, nz, nvx;
    size_t in_stride;  // na * nt * ns
};

constexpr int kHeliTX = 64, kHeliTY = 4;

// Rows [it0, it1) of the window whose edges sit at (z_dif -+ th/2) / dso
// (up, bo), or false when it misses the detector
// (src/helical_interpolation.cc:104-118).
__device__ __forceinline__ bool heli_window(const HeliProblem& P, double up, double bo,
                                            double cg, int& it0, int& it1)
{
    double upper_t = P.center_t - up * cg * P.dsd;
    double bottom_t = P.center_t - bo * cg * P.dsd;
    if (bottom_t < 0 || upper_t >= P.nt_d) return false;
    upper_t = upper_t < 0 ? 0 : upper_t;
    bottom_t = bottom_t >= P.nt_d ? P.nt_d : bottom_t;
    it0 = static_cast<int>(floor(upper_t));
    it1 = static_cast<int>(ceil(bottom_t));
    return true;
}

template <typename T>
__device__ __forceinline__ double heli_mean(const HeliProblem& P, const T* __restrict__ in,
                                            int view, int is, int it0, int it1)
{
    const T* p = in + (static_cast<size_t>(view) * P.nt + it0) * P.ns + is;
    const int n = it1 - it0;
    double v = 0;
    // four rows' loads issued together (rows past the window re-read its
    // last row and are not added), then added in row order
    for (int k = 0; k < n; k += 4) {
        T a[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] = p[static_cast<size_t>(min(k + j, n - 1)) * P.ns];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (k + j < n) v += static_cast<double>(a[j]);
    }
    return v / n;
}

// One wavefront per output row (slice, view): the row's view numbers, z
// offsets and their window terms (z_dif +- th/2) / dso -- the divisions the
// reference repeats per channel -- are evaluated once, then the lanes walk
// the row's channels 64 at a time.  Per-channel terms come from the host
// tables: cos(gamma), and for 180LI (pi - 2 gamma) / orbit and the mirrored
// channel cs - s/ds.  Slice table entries: z, centre view, window begin,
// first output view (integers stored exactly as doubles).
template <typename T, int MODE>
__global__ __launch_bounds__(kHeliTX* kHeliTY) void helical_kernel(
    HeliProblem P, const double* __restrict__ cosg, const double* __restrict__ difv,
    const double* __restrict__ sscor, const double* __restrict__ slices,
    const T* __restrict__ in, T* __restrict__ out, unsigned int* status)
{
    // Rows in (batch, slice, view) order, every XCD given a contiguous run of
    // them: neighbouring output views of a slice read neighbouring views (and,
    // for 180LI, conjugate lines shifted by one view), so they share L2.
    const int l = xcd_tile(blockIdx.x, gridDim.x);
    const int bz = l / P.nvx;  // batch * nz + slice
    const int iv = (l - bz * P.nvx) * kHeliTY + threadIdx.y;
    if (iv >= P.vpr) return;
    const int b = bz / P.nz, iz = bz - b * P.nz;
    const double z = slices[4 * iz];
    const int center_view = static_cast<int>(slices[4 * iz + 1]);
    const int begin = static_cast<int>(slices[4 * iz + 2]);
    const int first_view = static_cast<int>(slices[4 * iz + 3]);
    const int end = begin + P.vpr - 1;
    const T* src = in + static_cast<size_t>(b) * P.in_stride;
    T* dst = out + (static_cast<size_t>(bz) * P.vpr + iv) * P.ns;

    int cur_view = first_view + iv;
    if (MODE == 180 && cur_view > end) cur_view -= P.vpr;
    const double z_dif = z - P.couch_begin - cur_view * P.couch_mov;
    const double cu_up = (z_dif + P.half_th) / P.dso;
    const double cu_bo = (z_dif - P.half_th) / P.dso;
    // 360LI: the conjugate is the same angle one rotation away
    const int cor_view360 = cur_view >= center_view ? cur_view - P.vpr : cur_view + P.vpr;
    const double z_dif360 = z - P.couch_begin - cor_view360 * P.couch_mov;
    const double co_up360 = (z_dif360 + P.half_th) / P.dso;
    const double co_bo360 = (z_dif360 - P.half_th) / P.dso;

    for (int is = threadIdx.x; is < P.ns; is += kHeliTX) {
        const double cg = cosg[is];
        int it0, it1;
        const bool use_cur = heli_window(P, cu_up, cu_bo, cg, it0, it1);
        const double cur_value = use_cur ? heli_mean(P, src, cur_view, is, it0, it1) : 0.0;

        double z_dif2, cor_value = 0;
        bool use_cor;
        if (MODE == 360) {
            z_dif2 = z_dif360;
            use_cor = heli_window(P, co_up360, co_bo360, cg, it0, it1);
            if (use_cor) cor_value = heli_mean(P, src, cor_view360, is, it0, it1);
        } else {
            double cor_view = cur_view - difv[is];
            cor_view = cor_view < begin ? cor_view + P.vpr_d : cor_view;
            cor_view = cor_view > end ? cor_view - P.vpr_d : cor_view;
            z_dif2 = z - P.couch_begin - cor_view * P.couch_mov;
            use_cor = heli_window(P, (z_dif2 + P.half_th) / P.dso, (z_dif2 - P.half_th) / P.dso,
                                  cg, it0, it1);
            if (use_cor) {
                const int vl = static_cast<int>(floor(cor_view));
                const int vr = static_cast<int>(ceil(cor_view));
                const double u = cor_view - vl;
                const double ss_cor = sscor[is];
                if (ss_cor < 0 || ss_cor > P.ns1_d) {
                    // conjugate channel off the detector: the current value alone
                    dst[is] = static_cast<T>(cur_value);
                    continue;
                }
                if (vl < 0 || vr >= P.na) {
                    dst[is] = static_cast<T>(NAN);
                    if (status) atomicOr(status, 2u);
                    continue;
                }
                const int sl = static_cast<int>(floor(ss_cor));
                const int sr = static_cast<int>(ceil(ss_cor));
                const double v = ss_cor - sl;
                double left_value = 0, right_value = 0;
                int count = 0;
                if (P.ns >= 2) {
                    // both taps of a row from one pair of neighbouring
                    // channels (sb, sb + 1), which holds sl and sr
                    const int sb = min(sl, P.ns - 2);
                    const bool l_hi = sl != sb, r_hi = sr != sb;
                    const T* pl = src + (static_cast<size_t>(vl) * P.nt + it0) * P.ns + sb;
                    const T* pr = src + (static_cast<size_t>(vr) * P.nt + it0) * P.ns + sb;
                    for (int it = it0; it < it1; ++it, pl += P.ns, pr += P.ns) {
                        T a[2], c[2];
                        __builtin_memcpy(a, pl, sizeof(a));
                        __builtin_memcpy(c, pr, sizeof(c));
                        left_value += (1 - v) * static_cast<double>(l_hi ? a[1] : a[0]) +
                                      v * static_cast<double>(r_hi ? a[1] : a[0]);
                        right_value += (1 - v) * static_cast<double>(l_hi ? c[1] : c[0]) +
                                       v * static_cast<double>(r_hi ? c[1] : c[0]);
                        ++count;
                    }
                } else {
                    const T* pl = src + (static_cast<size_t>(vl) * P.nt + it0) * P.ns;
                    const T* pr = src + (static_cast<size_t>(vr) * P.nt + it0) * P.ns;
                    for (int it = it0; it < it1; ++it, pl += P.ns, pr += P.ns) {
                        left_value += (1 - v) * static_cast<double>(pl[sl]) +
                                      v * static_cast<double>(pl[sr]);
                        right_value += (1 - v) * static_cast<double>(pr[sl]) +
                                       v * static_cast<double>(pr[sr]);
                        ++count;
                    }
                }
                cor_value = ((1 - u) * left_value + u * right_value) / count;
            }
        }

        if (use_cur && use_cor)
            dst[is] = static_cast<T>((z_dif * cor_value - z_dif2 * cur_value) / (z_dif - z_dif2));
        else if (use_cur)
            dst[is] = static_cast<T>(cur_value);
        else if (use_cor)
            dst[is] = static_cast<T>(cor_value);
        else {
            dst[is] = static_cast<T>(NAN);
            if (status) atomicOr(status, 1u);
        }
    }
}

template <typename T>
int launch_helical(const HeliProblem& P, int mode, int nbatch, const double* tables,
                   const void* in, void* out, unsigned int* status, hipStream_t st)
{
    const dim3 grid(static_cast<unsigned>(P.nvx) * P.nz * nbatch);
    const dim3 block(kHeliTX, kHeliTY);
    const double* cosg = tables;
    const double* difv = tables + P.ns;
    const double* sscor = tables + 2 * static_cast<size_t>(P.ns);
    const double* slices = tables + 3 * static_cast<size_t>(P.ns);
    if (mode == 360)
        hipLaunchKernelGGL((helical_kernel<T, 360>), grid, block, 0, st, P, cosg, difv, sscor,
                           slices, static_cast<const T*>(in), static_cast<T*>(out), status);
    else
        hipLaunchKernelGGL((helical_kernel<T, 180>), grid, block, 0, st, P, cosg, difv, sscor,
                           slices, static_cast<const T*>(in), static_cast<T*>(out), status);
    return check_launch("helical_kernel");
}

}  // namespace

int helical(const ctr_helical_geom& g, int mode, int dtype, const void* in, void* out,
            int nbatch, const double* tables, unsigned int* status, hipStream_t st)
{
    HeliProblem P;
    P.ds = g.ds;
    P.center_s = (g.ns - 1) / 2.0 + g.offset_s;
    P.center_t = (g.nt - 1) / 2.0 + g.offset_t;
    P.couch_begin = g.couch_begin;
    P.couch_mov = g.couch_mov;
    P.half_th = g.thickness / 2.0;
    P.dso = g.dso;
    P.dsd = g.dsd;
    P.orbit = g.orbit;
    P.nt_d = g.nt;
    P.vpr_d = g.view_per_rot;
    P.ns1_d = g.ns - 1;
    P.ns = static_cast<int>(g.ns);
    P.nt = static_cast<int>(g.nt);
    P.na = static_cast<int>(g.na);
    P.vpr = static_cast<int>(g.view_per_rot);
    P.nz = static_cast<int>(g.nz);
    P.nvx = div_up(P.vpr, kHeliTY);
    P.in_stride = static_cast<size_t>(g.na) * g.nt * g.ns;
    // work-items of the grid stay below 2^32
    const long long nblk = static_cast<long long>(P.nvx) * P.nz * nbatch;
    CTR_REQUIRE(nblk * kHeliTX < (1LL << 32), CTR_ERR_SHAPE,
                "ctr_helical: output too large for one launch");
    return dtype == CTR_F64 ? launch_helical<double>(P, mode, nbatch, tables, in, out, status, st)
                            : launch_helical<float>(P, mode, nbatch, tables, in, out, status, st);
}

}  // namespace ctr
