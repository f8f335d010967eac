// Single-view parallel BP, ParallelSingleViewBp2DPixDriven<T>
// (src/bp_par_2d_sv.cc:54-92): every (pixel, view) keeps its own value,
//     img[slice][iy][ix][ia] = (1 - u) proj[ia][is1] + u proj[ia][is2],
//     s = (xcos[ix][ia] + ysin[iy][ia]) / ds + cents, is1 = floor(s),
//     is2 = ceil(s), u = s - is1, zero unless 0 <= s <= ns - 1,
// with no sum over views and no scale; the tables are the parallel BP's.
//
// The output (nx ny na values per slice, 755 MB at 512^2 x 720 views) makes
// this an HBM-write-bound kernel.  A workgroup owns an 8 x 8 block of pixels
// of one slice and walks the views in chunks of kSvViews: it stages the
// chunk's x cos of its 8 columns and y sin of its 8 rows in LDS (16 runs of
// consecutive views instead of one pair per pixel), computes with lanes over
// pixels (the 64 pixels of a block project to about a dozen neighbouring
// channels in each view, so a wave's taps share a few cache lines), and writes
// each pixel's chunk of views back with non-temporal 16-byte stores,
// consecutive lanes on consecutive pieces.  Each chunk's division by the
// channel pitch is one multiply and two FMAs (div_rn).  The arithmetic is the
// reference's, in FP64 without contraction, rounded to T once: the result
// matches it bit for bit.
//
// Measured on 512^2 x 720 views x 729 channels (kernel, rocprof): 418 us with
// 64 consecutive flattened pixels per workgroup, 32-view chunks and the
// division; 402 us with the 8 x 8 block; 376 us with 64-view chunks (128:
// 395 us); 296 us with non-temporal stores.
#include "ctr_common.hpp"

namespace ctr {

namespace {

// kVpt views per computing thread: 4 keeps the kernel at 64 VGPRs (8 waves
// per SIMD); 8 needs 120 (4 waves per SIMD) and measured 7% slower.
constexpr int kSvB = 8, kSvPix = kSvB * kSvB, kSvViews = 64, kSvPad = kSvViews + 1;
constexpr int kVpt = 4, kThreads = 512, kPass = kVpt * (kThreads / kSvPix);
static_assert(kSvViews % kPass == 0 && (2 * kSvB * kSvViews) % kThreads == 0, "chunk shape");

typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef double f64x2v __attribute__((ext_vector_type(2)));

struct SvbpProblem {
    int ns, na, nx, ny;
    double ds, cents;
    double top;    // ns - 1
    double inv_ds; // 1 / ds, rounded to nearest
};

// a / b rounded to nearest, from y = RN(1/b): q0 = RN(a y) is within an
// ulp of a / b, the remainder a - b q0 is exact in one FMA, and RN(q0 + r y)
// is the correctly rounded quotient (Markstein's theorem; IEEE-754 binary64,
// no underflow).  Tiny or zero |a| (where the remainder could underflow)
// takes the division itself.
__device__ __forceinline__ double div_rn(double a, double b, double y)
{
    if (!(fabs(a) >= 0x1p-960)) return a / b;
    const double q0 = a * y;
    const double r = fma(-q0, b, a);
    return fma(r, y, q0);
}

// 16-byte stores of a pixel's run of views: VEC when the per-pixel stride (na
// values of output) and the buffer are 16-byte aligned; otherwise element by
// element.
template <typename T>
__global__ __launch_bounds__(kThreads) void par_svbp_kernel(const T* __restrict__ proj,
                                                       T* __restrict__ img,
                                                       const double* __restrict__ xcos,
                                                       const double* __restrict__ ysin,
                                                       const SvbpProblem P, int vec)
{
#pragma clang fp contract(off)
    constexpr int kTv = 16 / sizeof(T);          // output values per 16 bytes
    __shared__ double tx[kSvB][kSvPad];       // x cos of the block's columns, [ix][view]
    __shared__ double ty[kSvB][kSvPad];       // y sin of its rows, [iy][view]
    __shared__ T res[kSvPix][kSvPad];         // results, [pixel][view]
    const int tid = threadIdx.x;
    const int ix0 = blockIdx.x * kSvB, iy0 = blockIdx.y * kSvB;
    const int b = blockIdx.z;
    const T* pr = proj + static_cast<size_t>(b) * P.na * P.ns;
    T* out = img + static_cast<size_t>(b) * P.nx * P.ny * P.na;
    // staging: element e = tid + kThreads k of the 16 runs of kSvViews
    // views (run r = e / kSvViews: x cos of column ix0 + r, or y sin of row
    // iy0 + r - 8)
    constexpr int kStage = 2 * kSvB * kSvViews / kThreads;
    const double* srow[kStage];
    double* sdst[kStage];
    int sv[kStage];
    bool s_ok[kStage];
#pragma unroll
    for (int k = 0; k < kStage; ++k) {
        const int e = tid + kThreads * k, sr = e / kSvViews;
        sv[k] = e - sr * kSvViews;
        const bool sx = sr < kSvB;
        const int sc = sx ? ix0 + sr : iy0 + sr - kSvB;
        s_ok[k] = sc < (sx ? P.nx : P.ny);
        srow[k] = (sx ? xcos : ysin) + static_cast<size_t>(s_ok[k] ? sc : 0) * P.na;
        sdst[k] = (sx ? &tx[0][0] : &ty[0][0]) + (sx ? sr : sr - kSvB) * kSvPad + sv[k];
    }
    // computing: thread -> (pixel tid & 63 = column i + 8 row j, views
    // kPass h + kVpt (tid >> 6) .. + kVpt - 1 in pass h)
    const int cp = tid & (kSvPix - 1);
    const int ci = cp & (kSvB - 1), cj = cp / kSvB;
    const bool c_ok = ix0 + ci < P.nx && iy0 + cj < P.ny;

    for (int va = 0; va < P.na; va += kSvViews) {
        const int nv = min(kSvViews, P.na - va);
        __syncthreads();  // the previous chunk's results are written out
#pragma unroll
        for (int k = 0; k < kStage; ++k)
            if (s_ok[k] && sv[k] < nv) *sdst[k] = srow[k][va + sv[k]];
        __syncthreads();
#pragma unroll 1
        for (int cv = kVpt * (tid >> 6); cv < nv && c_ok; cv += kPass) {
            // positions and taps of the thread's views first, then their
            // gathers back to back (an unused view reads channel 0 of a valid row)
            double u[kVpt];
            int i1[kVpt], i2[kVpt];
            bool ok[kVpt];
#pragma unroll
            for (int j = 0; j < kVpt; ++j) {
                const int v = min(cv + j, nv - 1);
                const double s = div_rn(tx[ci][v] + ty[cj][v], P.ds, P.inv_ds) + P.cents;
                ok[j] = cv + j < nv && s >= 0 && s <= P.top;
                const double f = floor(s);
                i1[j] = ok[j] ? static_cast<int>(f) : 0;
                i2[j] = ok[j] ? static_cast<int>(ceil(s)) : 0;
                u[j] = s - f;
            }
            T p0[kVpt], p1[kVpt];
#pragma unroll
            for (int j = 0; j < kVpt; ++j) {
                // the wave's 64 pixels share their views: one row per view
                const int vr = __builtin_amdgcn_readfirstlane(va + min(cv + j, nv - 1));
                const T* row = pr + static_cast<size_t>(vr) * P.ns;
                p0[j] = row[static_cast<unsigned int>(i1[j])];
                p1[j] = row[static_cast<unsigned int>(i2[j])];
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
        // writing: element e of [64 pixels][kSvViews / kTv pieces of 16 bytes]
        constexpr int kPieces = kSvViews / kTv;
        for (int e = tid; e < kSvPix * kPieces; e += kThreads) {
            const int wp = e / kPieces, v = kTv * (e - wp * kPieces);
            const int ix = ix0 + (wp & (kSvB - 1)), iy = iy0 + wp / kSvB;
            if (ix >= P.nx || iy >= P.ny || v >= nv) continue;
            T* o = out + (static_cast<size_t>(iy) * P.nx + ix) * P.na + va + v;
            // non-temporal: the output streams past L2 and is not read again
            if (vec && v + kTv <= nv) {
                if constexpr (sizeof(T) == 4) {
                    __builtin_nontemporal_store(
                        f32x4v{res[wp][v], res[wp][v + 1], res[wp][v + 2], res[wp][v + 3]},
                        reinterpret_cast<f32x4v*>(o));
                } else {
                    __builtin_nontemporal_store(f64x2v{res[wp][v], res[wp][v + 1]},
                                                reinterpret_cast<f64x2v*>(o));
                }
            } else {
                for (int j = 0; j < kTv && v + j < nv; ++j)
                    __builtin_nontemporal_store(res[wp][v + j], o + j);
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
    P.nx = static_cast<int>(g.nx);
    P.ny = static_cast<int>(g.ny);
    P.ds = g.ds;
    P.cents = cent_s(g);
    P.top = static_cast<double>(g.ns) - 1;
    P.inv_ds = 1.0 / g.ds;
    const dim3 grid(static_cast<unsigned int>(div_up(g.nx, kSvB)),
                    static_cast<unsigned int>(div_up(g.ny, kSvB)), nbatch);
    const size_t esize = dtype == CTR_F64 ? 8 : 4;
    const int vec = (g.na * esize) % 16 == 0 && (reinterpret_cast<uintptr_t>(img) & 15) == 0;
    if (dtype == CTR_F64)
        hipLaunchKernelGGL(par_svbp_kernel<double>, grid, dim3(kThreads), 0, st,
                           static_cast<const double*>(proj), static_cast<double*>(img), xcos,
                           ysin, P, vec);
    else
        hipLaunchKernelGGL(par_svbp_kernel<float>, grid, dim3(kThreads), 0, st,
                           static_cast<const float*>(proj), static_cast<float*>(img), xcos, ysin,
                           P, vec);
    return check_launch("par_svbp_kernel");
}

}  // namespace ctr
