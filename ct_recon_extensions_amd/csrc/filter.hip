// Ramp filter and fan weighting for gfx950.
//
// ramp_kernel: one workgroup filters ROWS detector rows.  The rows are
// staged in LDS (optionally multiplied by the fan / flat cosine weight on the
// way in, FanWeighting<T> src/fan_weighting.cc:19-35, so the weighting costs
// no extra pass over HBM), zero-padded by ns on both sides so no lane needs a
// range test.  The linear convolution of src/filter.cc:67-91,
//     out[is] = ds * sum_k in[k] h[is - k],
// is evaluated as h0*x[is] + sum over odd m of h[m] x[is - m]: even taps
// other than the centre are zero (src/filter.cc:26-30, :45-49), which halves
// the work.  The tap index is uniform across the workgroup (scalar loads),
// each tap is applied to all ROWS rows, and every lane reads consecutive
// LDS words (no bank conflicts).  The centre tap and the odd taps |m| <= 15
// (~97% of the filter's weight; on smooth rows they nearly cancel) are
// accumulated in FP64 with FP64 taps; the far field in the row type (FP32
// for float rows), added in FP64 at the end.
// The transpose (RampFilterGrad) is the same operator since h is symmetric.
#include "ctr_common.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace ctr {

constexpr int kRampRows = 4;

struct RampParams {
    int ns, nrow, wtype;
    double ds, dso, dsd, cents;
};

__device__ inline double cos_weight(const RampParams& P, int is)
{
    const double s = P.ds * (static_cast<double>(is) - P.cents);
    if (P.wtype == 2) return P.dso * fabs(cos(atan2(s, P.dsd))) / P.dsd;
    return P.dso * fabs(cos(s / P.dsd)) / P.dsd;
}

// hf: taps in the accumulation type, hf[ns + m] = h[m]; zero-padded rows in LDS.
constexpr int kRampNear = 15;  // |m| <= kRampNear: FP64 near field

template <typename T, typename Acc = typename std::conditional<sizeof(T) == 8, double, float>::type>
__global__ __launch_bounds__(256) void ramp_kernel(const T* __restrict__ in, T* __restrict__ out,
                                                   const Acc* __restrict__ hf,
                                                   const double* __restrict__ h64, RampParams P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    Acc* xs = reinterpret_cast<Acc*>(smem);  // [ROWS][3*ns]
    const int ns = P.ns, pitch = 3 * ns;
    const int r0 = blockIdx.x * kRampRows;
    const int nr = min(kRampRows, P.nrow - r0);
    for (int e = threadIdx.x; e < kRampRows * pitch; e += 256) {
        const int r = e / pitch, c = e - r * pitch - ns;
        Acc v = 0;
        if (r < nr && c >= 0 && c < ns) {
            const T x = in[static_cast<size_t>(r0 + r) * ns + c];
            v = P.wtype ? static_cast<Acc>(cos_weight(P, c) * static_cast<double>(x))
                        : static_cast<Acc>(x);
        }
        xs[e] = v;
    }
    __syncthreads();
    const double h0 = h64[ns];
    // odd m in [-(ns-1), ns-1]; x index ns + is - m stays in [1, 3ns-1]
    const int mlo = -(ns - 1) | 1;  // smallest odd >= -(ns-1)
    const int nlo = max(mlo, -kRampNear), nhi = min(ns - 1, kRampNear);
    for (int is = threadIdx.x; is < ns; is += 256) {
        Acc acc[kRampRows];
        double near[kRampRows];
#pragma unroll
        for (int r = 0; r < kRampRows; ++r) {
            acc[r] = Acc(0);
            near[r] = h0 * static_cast<double>(xs[r * pitch + ns + is]);
        }
        for (int m = mlo; m < nlo; m += 2) {
            const Acc h = hf[ns + m];
            const int xi = ns + is - m;
#pragma unroll
            for (int r = 0; r < kRampRows; ++r) acc[r] = fma(h, xs[r * pitch + xi], acc[r]);
        }
        for (int m = nlo; m <= nhi; m += 2) {
            const double h = h64[ns + m];
            const int xi = ns + is - m;
#pragma unroll
            for (int r = 0; r < kRampRows; ++r)
                near[r] = fma(h, static_cast<double>(xs[r * pitch + xi]), near[r]);
        }
        for (int m = nhi + 2; m <= ns - 1; m += 2) {
            const Acc h = hf[ns + m];
            const int xi = ns + is - m;
#pragma unroll
            for (int r = 0; r < kRampRows; ++r) acc[r] = fma(h, xs[r * pitch + xi], acc[r]);
        }
#pragma unroll
        for (int r = 0; r < kRampRows; ++r)
            if (r < nr)
                out[static_cast<size_t>(r0 + r) * ns + is] =
                    static_cast<T>((near[r] + static_cast<double>(acc[r])) * P.ds);
    }
}

template <typename T>
__global__ void taps_kernel(const double* __restrict__ h, T* __restrict__ hf, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) hf[i] = static_cast<T>(h[i]);
}

template <typename T>
__global__ void fanw_kernel(const T* __restrict__ in, T* __restrict__ out, RampParams P)
{
    const size_t n = static_cast<size_t>(P.nrow) * P.ns;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int is = static_cast<int>(i % P.ns);
        out[i] = static_cast<T>(cos_weight(P, is) * static_cast<double>(in[i]));
    }
}

namespace {

template <typename T>
int ramp_typed(const ctr_geom& g, const T* in, T* out, unsigned int nrow, const double* filter,
               int wtype, hipStream_t st)
{
    using Acc = typename std::conditional<sizeof(T) == 8, double, float>::type;
    RampParams P{static_cast<int>(g.ns), static_cast<int>(nrow), wtype, g.ds, g.dso, g.dsd,
                 cent_s(g)};
    const size_t lds = static_cast<size_t>(kRampRows) * 3 * g.ns * sizeof(Acc);
    CTR_REQUIRE(lds <= 160 * 1024, CTR_ERR_SHAPE, "ramp filter: ns too large for LDS");
    const int ntaps = 2 * g.ns + 1;
    Acc* hf = nullptr;
    CTR_HIP(hipMallocAsync(reinterpret_cast<void**>(&hf), ntaps * sizeof(Acc), st));
    hipLaunchKernelGGL(taps_kernel<Acc>, dim3(div_up(ntaps, 256)), dim3(256), 0, st, filter, hf,
                       ntaps);
    int rc = check_launch("taps_kernel");
    if (rc == CTR_OK) {
        auto kern = ramp_kernel<T>;
        if (lds > 64 * 1024)
            CTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
        hipLaunchKernelGGL(kern, dim3(div_up(nrow, kRampRows)), dim3(256), lds, st, in, out,
                           static_cast<const Acc*>(hf), filter, P);
        rc = check_launch("ramp_kernel");
    }
    CTR_HIP(hipFreeAsync(hf, st));
    return rc;
}

}  // namespace

int ramp(const ctr_geom& g, int dtype, const void* in, void* out, unsigned int nrow,
         const double* filter, int weight_type, hipStream_t st)
{
    if (dtype == CTR_F32)
        return ramp_typed(g, static_cast<const float*>(in), static_cast<float*>(out), nrow, filter,
                          weight_type, st);
    return ramp_typed(g, static_cast<const double*>(in), static_cast<double*>(out), nrow, filter,
                      weight_type, st);
}

int fan_weighting(const ctr_geom& g, int dtype, const void* in, void* out, unsigned int nrow,
                  hipStream_t st)
{
    RampParams P{static_cast<int>(g.ns), static_cast<int>(nrow), g.type, g.ds, g.dso, g.dsd,
                 cent_s(g)};
    const size_t n = static_cast<size_t>(nrow) * g.ns;
    const dim3 grid(static_cast<unsigned int>(std::min<size_t>(8192, (n + 255) / 256)));
    if (dtype == CTR_F32)
        hipLaunchKernelGGL(fanw_kernel<float>, grid, dim3(256), 0, st,
                           static_cast<const float*>(in), static_cast<float*>(out), P);
    else
        hipLaunchKernelGGL(fanw_kernel<double>, grid, dim3(256), 0, st,
                           static_cast<const double*>(in), static_cast<double*>(out), P);
    return check_launch("fanw_kernel");
}

}  // namespace ctr
