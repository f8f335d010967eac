// Geometry policy of the ray-driven FP's transpose on the BP engine
// (src/fp_par_2d.cc:332-391): per pixel-view the channels within |o1| of the
// pixel's projected centre mid, weight 1 - |c - mid|/|o1|, times the view
// weight.  The reference keeps one side of mid only for the first and last
// pixel along a view's marched axis; there the weight jumps from 1 to 0 as
// mid crosses a channel, so FP32 rounding of mid would flip whole taps.
// Those border pixels are recomputed afterwards with the reference's FP64
// arithmetic (fpt_kernel over the border); this policy serves the interior.
// Host-callable so the tap rule can be checked on the CPU.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

#include "tile_common.hpp"

namespace ctr {

struct alignas(16) FptAnchor {
    float a, b;    // channels per pixel step in x / in y
    float fc;      // mid at the anchor pixel minus is0
    float inv_o1;  // 1/|o1|
    float wv;      // the view's weight |dy/cos| or |dx/sin|
    int is0;
    int inside;    // unused by the custom taps (kept for the engine)
    int nch;       // channels from is0 the tile's taps can touch (one of margin)
    int pad[3];
};

struct FptProblem {
    int ns, na, nx, ny;
};

struct FptTables {
    const double* weights;  // [na]
    const double* pos;      // [na][stride][2]  (begin + i*offset2, offset1)
    const int* usex;        // [na]
    int stride;
};

struct FptGeo {
    using Problem = FptProblem;
    using Tables = FptTables;
    using Anchor = FptAnchor;
    static constexpr bool kCustomTaps = true;

    __host__ __device__ static Anchor anchor(const Problem& P, const Tables& T, const TileFrame& F,
                                    int view, int ilo, int ihi, int jlo, int jhi)
    {
        const bool xb = T.usex[view] != 0;
        const double* row = T.pos + 2 * static_cast<size_t>(view) * T.stride;
        double midc, a, b, o1;
        FptAnchor A;
        if (xb) {  // mid = pos[a][iy].begin + pos[a][iy].o1 * ix
            o1 = row[2 * F.jc + 1];
            midc = row[2 * F.jc] + o1 * F.ic;
            const int j2 = F.jc + 1 < P.ny ? F.jc + 1 : F.jc - 1;
            a = o1;
            b = P.ny > 1 ? (row[2 * j2] - row[2 * F.jc]) / (j2 - F.jc) : 0.0;
        } else {  // mid = pos[a][ix].begin + pos[a][ix].o1 * iy
            o1 = row[2 * F.ic + 1];
            midc = row[2 * F.ic] + o1 * F.jc;
            const int i2 = F.ic + 1 < P.nx ? F.ic + 1 : F.ic - 1;
            a = P.nx > 1 ? (row[2 * i2] - row[2 * F.ic]) / (i2 - F.ic) : 0.0;
            b = o1;
        }
        const double lo = fmin(a * ilo, a * ihi) + fmin(b * jlo, b * jhi);
        const double hi = fmax(a * ilo, a * ihi) + fmax(b * jlo, b * jhi);
        const int is0 = static_cast<int>(floor(midc + lo)) - 1;
        A.a = static_cast<float>(a);
        A.b = static_cast<float>(b);
        A.fc = static_cast<float>(midc - is0);
        A.inv_o1 = static_cast<float>(1.0 / fabs(o1));
        A.wv = static_cast<float>(T.weights[view]);
        A.is0 = is0;
        A.inside = 1;
        A.nch = static_cast<int>(floor(midc + hi)) + 3 - is0;
        A.pad[0] = A.pad[1] = A.pad[2] = 0;
        return A;
    }

    __host__ __device__ __forceinline__ static void taps(const Problem&, const Anchor& A, float fi,
                                                float fj, int& idx, float& w0, float& w1)
    {
        const float f = fmaf(A.a, fi, fmaf(A.b, fj, A.fc));
        const float fl = floorf(f);
        const float fr = f - fl;
        idx = static_cast<int>(fl);
        const float io = A.inv_o1;
        w0 = fmaxf(0.0f, fmaf(-fr, io, 1.0f)) * A.wv;        // channel floor(mid)
        w1 = fmaxf(0.0f, fmaf(fr - 1.0f, io, 1.0f)) * A.wv;  // channel floor(mid) + 1
    }
};


}  // namespace ctr
