// Single-view parallel BP, ParallelSingleViewBp2DPixDriven<T>
// (src/bp_par_2d_sv.cc:54-92): every (pixel, view) keeps its own value,
//     img[slice][iy][ix][ia] = (1 - u) proj[ia][is1] + u proj[ia][is2],
//     s = (xcos[ix][ia] + ysin[iy][ia]) / ds + cents, is1 = floor(s),
//     is2 = ceil(s), u = s - is1, zero unless 0 <= s <= ns - 1,
// with no sum over views and no scale; the tables are the parallel BP's.
//
// The output (nx ny na values per slice, 755 MB at 512^2 x 720 views) makes
// this an HBM-write-bound kernel.  A workgroup owns kSvPix consecutive
// pixels (flattened iy*nx + ix) of one slice and walks the views in chunks
// of kSvViews: it stages x cos + y sin of its pixels for the chunk in LDS
// (16-byte table reads, consecutive lanes on consecutive pieces of a
// pixel's contiguous run of views), computes with lanes over pixels (so the
// two taps of neighbouring pixels in one view row come from neighbouring
// channels), and writes each pixel's chunk of views back with 16-byte
// stores, consecutive lanes on consecutive pieces.  The arithmetic is the
// reference's, in FP64 without contraction, rounded to T once: the result
// matches it bit for bit.
#include "ctr_common.hpp"

namespace ctr {

namespace {

// kVpt views per computing thread: 4 keeps the kernel at 64 VGPRs (8 waves
// per SIMD); 8 needs 120 (4 waves per SIMD) and measured 7% slower.
constexpr int kSvPix = 64, kSvViews = 32, kSvPad = kSvViews + 1;
constexpr int kVpt = 4, kThreads = kSvPix * (kSvViews / kVpt), kStage = 1024 / kThreads;

struct SvbpProblem {
    int ns, na, npix;
    double ds, cents;
    double top;  // ns - 1
};

// 16-byte accesses of a pixel's run of views: VEC when the per-pixel strides
// (na doubles of table, na values of output) and the buffers are 16-byte
// aligned; otherwise element by element.
template <typename T>
__global__ __launch_bounds__(kThreads) void par_svbp_kernel(const T* __restrict__ proj,
                                                       T* __restrict__ img,
                                                       const double* __restrict__ xcos,
                                                       const double* __restrict__ ysin,
                                                       const SvbpProblem P, int nx, int vec)
{
#pragma clang fp contract(off)
    constexpr int kTv = 16 / sizeof(T);        // output values per 16 bytes
    __shared__ double pos[kSvPix][kSvPad];  // x cos + y sin, [pixel][view]
    __shared__ T res[kSvPix][kSvPad];       // results, [pixel][view]
    const int tid = threadIdx.x;
    const int pix0 = blockIdx.x * kSvPix;
    const int b = blockIdx.y;
    const T* pr = proj + static_cast<size_t>(b) * P.na * P.ns;
    T* out = img + (static_cast<size_t>(b) * P.npix + pix0) * P.na;
    // staging: element e = tid + kThreads k (k < kStage) of [64 pixels][16
    // pairs of views]: pixel e >> 4, views 2 (e & 15) and 2 (e & 15) + 1
    const double* xc[kStage];
    const double* ys[kStage];
    bool s_ok[kStage];
#pragma unroll
    for (int k = 0; k < kStage; ++k) {
        const int pix = pix0 + ((tid + kThreads * k) >> 4);
        s_ok[k] = pix < P.npix;
        const int iy = s_ok[k] ? pix / nx : 0, ix = s_ok[k] ? pix - iy * nx : 0;
        xc[k] = xcos + static_cast<size_t>(ix) * P.na;
        ys[k] = ysin + static_cast<size_t>(iy) * P.na;
    }
    // computing: thread -> (pixel tid & 63, views kVpt (tid >> 6) .. + kVpt - 1)
    const int cp = tid & (kSvPix - 1), cv = kVpt * (tid >> 6);
    const bool c_ok = pix0 + cp < P.npix;

    for (int va = 0; va < P.na; va += kSvViews) {
        const int nv = min(kSvViews, P.na - va);
        __syncthreads();  // the previous chunk's results are written out
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const int e = tid + kThreads * k, sp = e >> 4, v = 2 * (e & 15);
            if (!s_ok[k] || v >= nv) continue;
            if (vec && v + 1 < nv) {
                const double2 x2 = *reinterpret_cast<const double2*>(xc[k] + va + v);
                const double2 y2 = *reinterpret_cast<const double2*>(ys[k] + va + v);
                pos[sp][v] = x2.x + y2.x;
                pos[sp][v + 1] = x2.y + y2.y;
            } else {
                pos[sp][v] = xc[k][va + v] + ys[k][va + v];
                if (v + 1 < nv) pos[sp][v + 1] = xc[k][va + v + 1] + ys[k][va + v + 1];
            }
        }
        __syncthreads();
        if (c_ok) {
            // positions and taps of the thread's views first, then their
            // gathers back to back (an unused view reads channel 0 of a valid row)
            double u[kVpt];
            int i1[kVpt], i2[kVpt];
            bool ok[kVpt];
#pragma unroll
            for (int j = 0; j < kVpt; ++j) {
                const int v = min(cv + j, nv - 1);
                const double s = pos[cp][v] / P.ds + P.cents;
                ok[j] = cv + j < nv && s >= 0 && s <= P.top;
                const double f = floor(s);
                i1[j] = ok[j] ? static_cast<int>(f) : 0;
                i2[j] = ok[j] ? static_cast<int>(ceil(s)) : 0;
                u[j] = s - f;
            }
            T p0[kVpt], p1[kVpt];
#pragma unroll
            for (int j = 0; j < kVpt; ++j) {
                const T* row = pr + static_cast<size_t>(va + min(cv + j, nv - 1)) * P.ns;
                p0[j] = row[i1[j]];
                p1[j] = row[i2[j]];
            }
#pragma unroll
            for (int j = 0; j < kVpt; ++j) {
                const T r = ok[j] ? static_cast<T>((1 - u[j]) * static_cast<double>(p0[j]) +
                                                   u[j] * static_cast<double>(p1[j]))
                                  : T(0);
                if (cv + j < nv) res[cp][cv + j] = r;
            }
        }
        __syncthreads();
        // writing: element e of [64 pixels][32 / kTv pieces of 16 bytes]
        constexpr int kPieces = kSvViews / kTv;
        for (int e = tid; e < kSvPix * kPieces; e += kThreads) {
            const int wp = e / kPieces, v = kTv * (e - wp * kPieces);
            if (pix0 + wp >= P.npix || v >= nv) continue;
            T* o = out + static_cast<size_t>(wp) * P.na + va + v;
            if (vec && v + kTv <= nv) {
                if constexpr (sizeof(T) == 4) {
                    *reinterpret_cast<float4*>(o) =
                        make_float4(res[wp][v], res[wp][v + 1], res[wp][v + 2], res[wp][v + 3]);
                } else {
                    *reinterpret_cast<double2*>(o) = make_double2(res[wp][v], res[wp][v + 1]);
                }
            } else {
                for (int j = 0; j < kTv && v + j < nv; ++j) o[j] = res[wp][v + j];
            }
        }
    }
}

}  // namespace

int par_svbp(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
             const double* xcos, const double* ysin, hipStream_t st)
{
    CTR_REQUIRE(g.ns >= 1 && g.na >= 1 && g.nx >= 1 && g.ny >= 1, CTR_ERR_ARG,
                "single-view BP: sizes must be positive");
    CTR_REQUIRE(g.ds > 0, CTR_ERR_ARG, "single-view BP: ds must be positive");
    CTR_REQUIRE(nbatch <= 65535, CTR_ERR_ARG, "single-view BP: at most 65535 slices per call");
    const long long npix = static_cast<long long>(g.nx) * g.ny;
    CTR_REQUIRE(npix < (1LL << 31) && static_cast<long long>(g.na) * g.ns < (1LL << 31),
                CTR_ERR_SHAPE, "single-view BP: image or sinogram too large");
    SvbpProblem P{};
    P.ns = static_cast<int>(g.ns);
    P.na = static_cast<int>(g.na);
    P.npix = static_cast<int>(npix);
    P.ds = g.ds;
    P.cents = cent_s(g);
    P.top = static_cast<double>(g.ns) - 1;
    const dim3 grid(static_cast<unsigned int>((npix + kSvPix - 1) / kSvPix), nbatch);
    const size_t esize = dtype == CTR_F64 ? 8 : 4;
    const auto a16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    const int vec = (g.na * 8) % 16 == 0 && (g.na * esize) % 16 == 0 && a16(img) && a16(xcos) &&
                    a16(ysin);
    if (dtype == CTR_F64)
        hipLaunchKernelGGL(par_svbp_kernel<double>, grid, dim3(kThreads), 0, st,
                           static_cast<const double*>(proj), static_cast<double*>(img), xcos,
                           ysin, P, static_cast<int>(g.nx), vec);
    else
        hipLaunchKernelGGL(par_svbp_kernel<float>, grid, dim3(kThreads), 0, st,
                           static_cast<const float*>(proj), static_cast<float*>(img), xcos, ysin,
                           P, static_cast<int>(g.nx), vec);
    return check_launch("par_svbp_kernel");
}

}  // namespace ctr
