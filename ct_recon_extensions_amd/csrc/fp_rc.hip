// Ray-casting parallel FP, ParallelProjection2DRayCasting<T>
// (src/fp_par_2d.cc:91-144): every ray starts at its entry point on the FOV
// circle (pixel-index units) and takes nsteps steps of (step sin, step cos);
// a sample inside [0, nx-1] x [0, ny-1] adds the bilinear interpolation of
// the image, summed in FP64 in ray order, and the ray is sum * step * dx.
//
// A wave holds 8 adjacent rays x 8 step phases: lane j of a ray samples the
// ray's steps j, j + 8, j + 16, ..., so one gather instruction covers a
// patch of 8 rays x 8 steps (8 x 1 pixels), and a sample's two taps in an
// image row are one 8-byte load.  With one ray per lane, a
// wave's 64 samples of one step lie on a segment across the rays that at
// steep angles crosses one image row per lane (53 cache lines per gather
// measured, which saturated the vector L1); the patch needs a few.
//
// The arithmetic is the reference's, in FP64 without contraction and in the
// same order, rounded to T once, so the result matches it bit for bit:
// - every lane advances its ray one step at a time, as the reference
//   accumulates x += step sin;
// - the 8 lanes of a ray sum the ray's samples in step order, read back
//   through LDS, each lane computing the same sum;
// - a step past the ray's end or outside the image contributes +0.  The sum
//   starts at +0 and never becomes -0 under round-to-nearest, so adding +0
//   leaves it unchanged.
#include "ctr_common.hpp"

namespace ctr {

namespace {

constexpr int kRcPhases = 8, kRcRays = 256 / kRcPhases;  // per 256-thread block

struct RcProblem {
    int ns, na, nx, ny, nrays;
    double step, dx, xmax, ymax;  // xmax = nx - 1, ymax = ny - 1
};

template <typename T>
__global__ __launch_bounds__(256) void fp_rc_kernel(const T* __restrict__ img,
                                                    T* __restrict__ proj,
                                                    const double* __restrict__ sincostbl,
                                                    const double* __restrict__ begins,
                                                    const int* __restrict__ nsteps,
                                                    const RcProblem P)
{
#pragma clang fp contract(off)
    // each lane's sample value, read back by the 8 lanes of its ray
    __shared__ __align__(16) double vals[256];
    const int j = threadIdx.x % kRcPhases;
    const int r = blockIdx.x * kRcRays + threadIdx.x / kRcPhases;
    const bool valid = r < P.nrays;
    const int b = blockIdx.y;
    const T* im = img + static_cast<size_t>(b) * P.nx * P.ny;
    const int n = valid ? nsteps[r] : 0;
    if (n > 0) {
        const int a = r / P.ns;
        const double stx = P.step * sincostbl[2 * a], sty = P.step * sincostbl[2 * a + 1];
        double x = begins[2 * static_cast<size_t>(r)], y = begins[2 * static_cast<size_t>(r) + 1];
        for (int t = 0; t < j; ++t) {
            x += stx;
            y += sty;
        }
        double sum = 0;
        // the ray's 8 values through LDS (one write, four 16-byte reads; a
        // wave's LDS accesses stay in order), added in step order
        auto add_ray = [&](double val) {
            vals[threadIdx.x] = val;
            __builtin_amdgcn_wave_barrier();
            const double2* g2 =
                reinterpret_cast<const double2*>(vals + (threadIdx.x & ~(kRcPhases - 1)));
            double2 w[kRcPhases / 2];
#pragma unroll
            for (int t = 0; t < kRcPhases / 2; ++t) w[t] = g2[t];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int t = 0; t < kRcPhases / 2; ++t) {
                sum += w[t].x;
                sum += w[t].y;
            }
        };
        auto advance = [&]() {
#pragma unroll
            for (int t = 0; t < kRcPhases; ++t) {
                x += stx;
                y += sty;
            }
        };
        if (P.nx >= 2) {
            // Software-pipelined over two register sets: the next sample's
            // taps are loaded before this sample's value and the ray's sum.
            // A sample's two taps in a row come from one load of the column
            // pair (xb, xb + 1), which holds ix1 and ix2.
            struct Tap {
                bool in, h1, h2;
                double u, v;
                T q1[2], q2[2];
            };
            auto fetch = [&](int k, Tap& s) {
                s.in = k + j < n && x >= 0 && x <= P.xmax && y >= 0 && y <= P.ymax;
                if (s.in) {
                    const double fx = floor(x), fy = floor(y);
                    const int ix1 = static_cast<int>(fx), ix2 = static_cast<int>(ceil(x));
                    const int iy1 = static_cast<int>(fy), iy2 = static_cast<int>(ceil(y));
                    s.u = x - fx;
                    s.v = y - fy;
                    const int xb = min(ix1, P.nx - 2);
                    s.h1 = ix1 != xb;
                    s.h2 = ix2 != xb;
                    __builtin_memcpy(s.q1, im + xb + iy1 * P.nx, sizeof(s.q1));
                    __builtin_memcpy(s.q2, im + xb + iy2 * P.nx, sizeof(s.q2));
                }
            };
            auto value = [&](const Tap& s) {
                double val = 0;
                if (s.in) {
                    const double u = s.u, v = s.v;
                    const double p11 = static_cast<double>(s.h1 ? s.q1[1] : s.q1[0]);
                    const double p21 = static_cast<double>(s.h2 ? s.q1[1] : s.q1[0]);
                    const double p12 = static_cast<double>(s.h1 ? s.q2[1] : s.q2[0]);
                    const double p22 = static_cast<double>(s.h2 ? s.q2[1] : s.q2[0]);
                    val = (1 - v) * ((1 - u) * p11 + u * p21) + v * ((1 - u) * p12 + u * p22);
                }
                return val;
            };
            Tap ta, tb;
            fetch(0, ta);
            for (int k0 = 0;; k0 += 2 * kRcPhases) {
                advance();
                fetch(k0 + kRcPhases, tb);
                add_ray(value(ta));
                if (k0 + kRcPhases >= n) break;
                advance();
                fetch(k0 + 2 * kRcPhases, ta);
                add_ray(value(tb));
                if (k0 + 2 * kRcPhases >= n) break;
            }
        } else {
            for (int k0 = 0; k0 < n; k0 += kRcPhases) {
                double val = 0;
                if (k0 + j < n && x >= 0 && x <= P.xmax && y >= 0 && y <= P.ymax) {
                    const double fx = floor(x), fy = floor(y);
                    const int ix1 = static_cast<int>(fx), ix2 = static_cast<int>(ceil(x));
                    const int iy1 = static_cast<int>(fy), iy2 = static_cast<int>(ceil(y));
                    const double u = x - fx, v = y - fy;
                    const double p11 = static_cast<double>(im[ix1 + iy1 * P.nx]);
                    const double p21 = static_cast<double>(im[ix2 + iy1 * P.nx]);
                    const double p12 = static_cast<double>(im[ix1 + iy2 * P.nx]);
                    const double p22 = static_cast<double>(im[ix2 + iy2 * P.nx]);
                    val = (1 - v) * ((1 - u) * p11 + u * p21) + v * ((1 - u) * p12 + u * p22);
                }
                add_ray(val);
                advance();
            }
        }
        if (j == 0) proj[static_cast<size_t>(b) * P.nrays + r] = static_cast<T>(sum * P.step * P.dx);
    } else if (valid && j == 0) {
        proj[static_cast<size_t>(b) * P.nrays + r] = T(0);
    }
}

}  // namespace

int fp_rc(const ctr_geom& g, int dtype, const void* img, void* proj, int nbatch, double step,
          const double* sincostbl, const double* begins, const int* nsteps, hipStream_t st)
{
    CTR_REQUIRE(g.ns >= 1 && g.na >= 1 && g.nx >= 1 && g.ny >= 1, CTR_ERR_ARG,
                "ray-casting FP: sizes must be positive");
    CTR_REQUIRE(step > 0, CTR_ERR_ARG, "ray-casting FP: step must be positive");
    CTR_REQUIRE(nbatch <= 65535, CTR_ERR_ARG, "ray-casting FP: at most 65535 slices per call");
    const long long nrays = static_cast<long long>(g.na) * g.ns;
    CTR_REQUIRE(nrays < (1LL << 31) && static_cast<long long>(g.nx) * g.ny < (1LL << 31),
                CTR_ERR_SHAPE, "ray-casting FP: image or sinogram too large");
    RcProblem P{};
    P.ns = static_cast<int>(g.ns);
    P.na = static_cast<int>(g.na);
    P.nx = static_cast<int>(g.nx);
    P.ny = static_cast<int>(g.ny);
    P.nrays = static_cast<int>(nrays);
    P.step = step;
    P.dx = g.dx;
    P.xmax = static_cast<double>(g.nx) - 1;
    P.ymax = static_cast<double>(g.ny) - 1;
    const dim3 grid(static_cast<unsigned int>((nrays + kRcRays - 1) / kRcRays), nbatch);
    if (dtype == CTR_F64)
        hipLaunchKernelGGL(fp_rc_kernel<double>, grid, dim3(256), 0, st,
                           static_cast<const double*>(img), static_cast<double*>(proj), sincostbl,
                           begins, nsteps, P);
    else
        hipLaunchKernelGGL(fp_rc_kernel<float>, grid, dim3(256), 0, st,
                           static_cast<const float*>(img), static_cast<float*>(proj), sincostbl,
                           begins, nsteps, P);
    return check_launch("fp_rc_kernel");
}

}  // namespace ctr
