// Per-(pixel tile, view) fan-beam geometry, the policy the BP engine uses
// for bp_fan_2d and bp_fan_2d_angle; the forward kernel and its transpose
// evaluate bit-identical channel positions and weights through it.
//
// For a pixel at (x, y) and source angle beta (source at
// (-dso sin b, dso cos b)) the reference computes, in FP64
// (src/bp_fan_2d.cc:66-70):
//     d = dso + x sin b - y cos b,  r = x cos b + y sin b,
//     s = atan2(r, d) * dsd/ds + cents,  w = dsd^2 / (d^2 + r^2).
// Here the tile anchor (xc, yc) is evaluated once in FP64 and every pixel of
// the tile is expressed relative to it.  With (dd, dr) the pixel's offsets
// of (d, r) from the anchor's (dc, rc) and Lc^2 = dc^2 + rc^2,
//     tan(gamma - gamma_c) = u = (dr dc - dd rc) / (Lc^2 + dd dc + dr rc),
// which is a ratio of two functions linear in the pixel offsets (i, j), and
//     d^2 + r^2 = Lc^2 (1 + e)^2 (1 + u^2),  e = (dd dc + dr rc)/Lc^2.
// So s = s_c + (dsd/ds) atan(u) and w = (dsd^2/Lc^2) / ((1+e)^2 (1+u^2)),
// evaluated per pixel in FP32 on small numbers: the channel position keeps
// ~1e-6 channel accuracy where a direct FP32 evaluation of s ~ 700 would
// lose ~5e-5.  The FP64 anchor carries the large part (s_c) exactly.
#pragma once

#include <hip/hip_runtime.h>

#include "ctr_common.hpp"
#include "tile_common.hpp"

namespace ctr {

struct alignas(16) FanAnchor {
    float ai, aj;  // numerator slopes per pixel step in x / y, pre-scaled
    float bi, bj;  // denominator slopes per pixel step
    float b0;      // denominator constant
    float fc;      // anchor channel position minus is0
    int is0;       // first channel of the staged detector span
    int inside;    // 1: every pixel of the tile lands strictly inside [0, ns-1]
    int nch;       // channels from is0 the tile's taps can touch (one of margin)
    int pad[3];
};

// Geometry of one fan problem as the kernels see it.
struct FanProblem {
    double dso, dsd, ds, cents, dx, dy;
    int ns, na, nx, ny;
    // polynomial coefficients in v = k*u (k = dsd/ds):
    //   atan:      ds(v) = v * (1 + v2*(p1 + v2*p2)),  p1 = -1/(3k^2), p2 = 1/(5k^4)
    //   1/(1+u^2) = 1 + v2*(g1 + v2*g2),                 g1 = -1/k^2,   g2 = 1/k^4
    float p1, p2, g1, g2;
    float k;  // dsd/ds in FP32 (exact mode)
};

struct FanTables {
    const double* xpos;     // [nx]   (src/bp_fan_2d.cc:37-40)
    const double* ypos;     // [ny]
    const double* sincos;   // [2na]  (sin, cos) of the source angle
};

// Anchor of the tile whose local pixel offsets span [ilo, ihi] x [jlo, jhi]
// around (xc, yc).  FP64 for the anchor itself, FP32 for the span bounds.
__device__ inline FanAnchor fan_anchor(const FanProblem& P, double xc, double yc, double sb,
                                       double cb, int ilo, int ihi, int jlo, int jhi)
{
    const double d = P.dso + xc * sb - yc * cb;
    const double r = xc * cb + yc * sb;
    const double L2 = d * d + r * r;
    const double inv = 1.0 / L2;
    const double k = P.dsd / P.ds;
    const double sc = atan2(r, d) * k + P.cents;
    // dd = i*dx*sb + j*dy*cb,  dr = i*dx*cb - j*dy*sb  (y decreases with iy)
    const double nI = P.dx * (cb * d - sb * r) * inv;
    const double nJ = -P.dy * (sb * d + cb * r) * inv;
    const double eI = P.dx * (sb * d + cb * r) * inv;
    const double eJ = P.dy * (cb * d - sb * r) * inv;
    // The channel offset is monotone along any line (a ratio of linear
    // functions through atan), so the tile's extreme channels are at corners.
    const float fnI = static_cast<float>(nI), fnJ = static_cast<float>(nJ);
    const float feI = static_cast<float>(eI), feJ = static_cast<float>(eJ);
    const float fk = static_cast<float>(k);
    float lo = 0.0f, hi = 0.0f;
    const int ci[2] = {ilo, ihi}, cj[2] = {jlo, jhi};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const float u = (fnI * ci[a] + fnJ * cj[b]) / (1.0f + feI * ci[a] + feJ * cj[b]);
            const float dsv = fk * atanf(u);
            lo = fminf(lo, dsv);
            hi = fmaxf(hi, dsv);
        }
    const double smin = sc + lo, smax = sc + hi;
    const int is0 = static_cast<int>(floor(smin)) - 1;
    const double q = P.dsd * sqrt(inv);  // sqrt(dsd^2 / Lc^2)
    const double iq = 1.0 / q;
    FanAnchor A;
    A.ai = static_cast<float>(k * nI * iq);
    A.aj = static_cast<float>(k * nJ * iq);
    A.bi = static_cast<float>(eI * iq);
    A.bj = static_cast<float>(eJ * iq);
    A.b0 = static_cast<float>(iq);
    A.fc = static_cast<float>(sc - is0);
    A.is0 = is0;
    A.inside = (smin >= 0.25 && smax <= P.ns - 1.25) ? 1 : 0;
    A.nch = static_cast<int>(floor(smax)) + 3 - is0;
    A.pad[0] = A.pad[1] = A.pad[2] = 0;
    return A;
}

// Channel position (relative to is0) and weight of local pixel (fi, fj).
// Default mode: short series (the host checks |u| stays small enough);
// EXACT: atan and 1/(1+u^2) evaluated directly.
template <bool EXACT>
__device__ __forceinline__ void fan_pixel(const FanProblem& P, const FanAnchor& A, float fi,
                                          float fj, float& f, float& w)
{
    const float num = fmaf(A.ai, fi, A.aj * fj);
    const float den = fmaf(A.bi, fi, fmaf(A.bj, fj, A.b0));
    const float rd = __builtin_amdgcn_rcpf(den);
    const float v = num * rd;  // k * u, channels
    if constexpr (EXACT) {
        const float u = v / P.k;
        f = A.fc + P.k * atanf(u);
        w = rd * rd / fmaf(u, u, 1.0f);
    } else {
        const float v2 = v * v;
        const float t = v2 * fmaf(v2, P.p2, P.p1);
        f = A.fc + fmaf(v, t, v);
        w = rd * rd * fmaf(v2, fmaf(v2, P.g2, P.g1), 1.0f);
    }
}

// Host side (fan_bp.hip): geometry checks for a TW x TH tiling, the
// channels one tile's span needs (lmax) and whether the series is accurate
// enough (exact = false) -- and the smallest source-to-pixel distance.
int fan_problem(const ctr_geom& g, int TW, int TH, FanProblem& P, int& lmax, bool& exact);
double fan_min_distance(const ctr_geom& g, int TW, int TH);

// Policy for the BP engine.
template <bool EXACT>
struct FanGeo {
    using Problem = FanProblem;
    using Tables = FanTables;
    using Anchor = FanAnchor;
    __device__ static Anchor anchor(const Problem& P, const Tables& T, const TileFrame& F,
                                    int view, int ilo, int ihi, int jlo, int jhi)
    {
        return fan_anchor(P, T.xpos[F.ic], T.ypos[F.jc], T.sincos[2 * view],
                          T.sincos[2 * view + 1], ilo, ihi, jlo, jhi);
    }
    __device__ __forceinline__ static void pixel(const Problem& P, const Anchor& A, float fi,
                                                 float fj, float& f, float& w)
    {
        fan_pixel<EXACT>(P, A, fi, fj, f, w);
    }
    // the series' addends p1 and g1 meet the uniform p2 and g2 in one FMA each
    __device__ __forceinline__ static void pin(Problem& P)
    {
        asm volatile("" : "+v"(P.p1), "+v"(P.g1));
    }
    // Level line s = s_star in pixel-index coordinates: the ray at fan angle
    // g* = (s_star - cents) ds / dsd, r - tan(g*) d = 0, i.e.
    // L(ix, iy) = pi ix + qj iy + rr >= 0 exactly where s >= s_star for every
    // pixel in front of the source (d > 0; the host admits no other).  g* is
    // clamped to +-1.47 rad: every admitted pixel has |g| <= asin(0.95), so a
    // clamped line still bounds the same set of image pixels.
    __device__ static void boundary(const Problem& P, const Tables& T, int view, double s_star,
                                    double& pi, double& qj, double& rr)
    {
        const double sb = T.sincos[2 * view], cb = T.sincos[2 * view + 1];
        const double g = fmin(fmax((s_star - P.cents) * P.ds / P.dsd, -1.47), 1.47);
        const double t = tan(g);
        const double pp = cb - t * sb, qq = sb + t * cb, r0 = -t * P.dso;
        const double x0 = T.xpos[0], y0 = T.ypos[0];
        const double xs = P.nx > 1 ? (T.xpos[P.nx - 1] - x0) / (P.nx - 1) : P.dx;
        const double ys = P.ny > 1 ? (T.ypos[P.ny - 1] - y0) / (P.ny - 1) : -P.dy;
        pi = pp * xs;
        qj = qq * ys;
        rr = pp * x0 + qq * y0 + r0;
    }
};

}  // namespace ctr
