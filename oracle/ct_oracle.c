/*
 * ct_oracle.c -- CPU restatement of the reference's 2-D CT hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product path (ct_recon_extensions_amd) never
 * links, loads or falls back to it.
 *
 * Every function restates one routine of the reference's single-threaded CPU
 * implementation (the src/ directory of ltldlzltl/CT_RECON_EXTENSIONS), keeping its
 * arithmetic type, its evaluation order and its edge rules, so that the
 * results agree with the compiled reference to the last bit where the
 * reference is well defined.  Where the reference is broken (the transposes
 * of bp_par_2d and bp_fan_2d_angle, the unsigned-underflow edge of the
 * ray-driven FP transpose) the oracle states the mathematically intended
 * operator instead, and says so at the function.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   - against golden vectors written by the compiled reference
 *     (oracle/_ref, tests/golden/make_golden.py): prep tables, fan BP,
 *     fan BP transpose, fan-angle BP, par BP, ray-driven FP and its
 *     transpose, ramp filter and its transpose, fan weighting;
 *   - against closed forms / brute force: ramp taps (Ram-Lak), the filter
 *     (numpy.convolve in FP64), constant-sinogram BP identities, the disk
 *     chord length for FP, fan weights;
 *   - transposes without a usable reference counterpart (par BP transpose,
 *     fan-angle BP transpose) are pinned by the dot-product adjoint test
 *     against their pinned forward operators, on sizes where a planted
 *     index or sign error moves the inner product far beyond rounding.
 *
 * Build: oracle/Makefile -> oracle/libct_oracle.so (gcc -O2, C99, no FMA
 * contraction, so the FP64 expressions round exactly like the reference's
 * g++ -O3 -std=c++11 build on x86-64).  Sines and cosines of one angle go
 * through glibc's sincos: g++ -O3 merges the reference's sin(angle) and
 * cos(angle) into that call, and its cosine can differ from cos() in the
 * last bit (view 464 of 720 over 180 deg).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stddef.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_PI_4
#define M_PI_4 0.78539816339744830962
#endif

/* One geometry record for every operator (the union of the reference's
 * Param structs, include/bp_fan_2d.h:20-59, include/bp_par_2d.h:20-55,
 * include/fp_par_2d.h:20-55, include/filter.h:21-36,
 * include/fan_weighting.h:21-39). */
typedef struct {
    unsigned int ns, na;
    double ds, orbit, offset_s, orbit_start;
    unsigned int nx, ny;
    double dx, dy, offset_x, offset_y;
    double dso, dsd, fov;
    int type;            /* filter: 0 par, 1 fan, 2 flat; weighting: 1 fan, 2 flat */
    unsigned int nrow;   /* filter / weighting rows */
} orc_geom;

/* ------------------------------------------------------------------ */
/* Fan-beam tables: src/bp_fan_2d.cc:18-44 (GradPrep :96-122 identical). */
int orc_fan_prep(const orc_geom* g, double* xpos, double* ypos, double* sincostbl)
{
    const double centx = ((double)(g->nx - 1)) / 2 + g->offset_x;
    const double centy = ((double)(g->ny - 1)) / 2 + g->offset_y;
    double angle = g->orbit_start;
    double posx = -centx * g->dx, posy = centy * g->dy;
    unsigned int i;
    for (i = 0; i < g->na; ++i) {
        sincos(angle, &sincostbl[2 * i], &sincostbl[2 * i + 1]);
        angle += g->orbit;
    }
    for (i = 0; i < g->nx; ++i) { xpos[i] = posx; posx += g->dx; }
    for (i = 0; i < g->ny; ++i) { ypos[i] = posy; posy -= g->dy; }
    return 1;
}

/* Fan-beam tables from caller angles: src/bp_fan_2d_angle.cc:18-45. */
int orc_fan_angle_prep(const orc_geom* g, const double* angles, double* xpos,
                       double* ypos, double* sincostbl)
{
    const double centx = ((double)(g->nx - 1)) / 2 + g->offset_x;
    const double centy = ((double)(g->ny - 1)) / 2 + g->offset_y;
    double posx = -centx * g->dx, posy = centy * g->dy;
    unsigned int i;
    for (i = 0; i < g->na; ++i) {
        sincos(angles[i], &sincostbl[2 * i], &sincostbl[2 * i + 1]);
    }
    for (i = 0; i < g->nx; ++i) { xpos[i] = posx; posx += g->dx; }
    for (i = 0; i < g->ny; ++i) { ypos[i] = posy; posy -= g->dy; }
    return 1;
}

/* Scale applied by bp_fan_2d: |orbit| / round(|na*orbit|/pi)
 * (src/bp_fan_2d.cc:58,83).  bp_fan_2d_angle applies none (:83). */
static double fan_scale(const orc_geom* g)
{
    const double factor = round(fabs(g->na * g->orbit) / M_PI);
    return fabs(g->orbit) / factor;
}

/* Pixel-driven fan BP, src/bp_fan_2d.cc:46-91 (scaled=1) and
 * src/bp_fan_2d_angle.cc:47-91 (scaled=0).  FP64 accumulation. */
#define DEF_FAN_BP(T, SUF)                                                        \
int orc_fan_bp_##SUF(const orc_geom* g, const T* proj, T* img, const double* xpos, \
                     const double* ypos, const double* sc, int scaled)            \
{                                                                                 \
    const double cents = ((double)(g->ns - 1)) / 2 + g->offset_s;                 \
    const double factor = round(fabs(g->na * g->orbit) / M_PI);                   \
    unsigned int ix, iy, ia, is1, is2;                                            \
    for (iy = 0; iy < g->ny; ++iy) {                                              \
        for (ix = 0; ix < g->nx; ++ix) {                                          \
            double sum = 0.0;                                                     \
            const double x = xpos[ix], y = ypos[iy];                              \
            for (ia = 0; ia < g->na; ++ia) {                                      \
                const double sb = sc[2 * ia], cb = sc[2 * ia + 1];                \
                const double d = g->dso + x * sb - y * cb;                        \
                const double r = x * cb + y * sb;                                 \
                const double gam = atan2(r, d);                                   \
                const double s = gam * g->dsd / g->ds + cents;                    \
                const double w = g->dsd * g->dsd / (d * d + r * r);               \
                if (s >= 0 && s <= g->ns - 1) {                                   \
                    double u;                                                     \
                    is1 = (unsigned int)floor(s);                                 \
                    is2 = (unsigned int)ceil(s);                                  \
                    u = s - is1;                                                  \
                    sum += ((1 - u) * proj[is1 + ia * g->ns]                      \
                            + u * proj[is2 + ia * g->ns]) * w;                    \
                }                                                                 \
            }                                                                     \
            img[ix + iy * g->nx] = scaled ? (T)(sum * fabs(g->orbit) / factor)    \
                                          : (T)sum;                               \
        }                                                                         \
    }                                                                             \
    return 1;                                                                     \
}
DEF_FAN_BP(float, f)
DEF_FAN_BP(double, d)

/* Transpose of the fan BP.  scaled=1 restates src/bp_fan_2d.cc:124-170
 * (accumulation in T, scale applied at the end in T).  scaled=0 is the
 * transpose of the unscaled bp_fan_2d_angle; the reference's own version
 * (src/bp_fan_2d_angle.cc:124-165) divides by the weights and is not a
 * transpose, so this branch is pinned by the adjoint test instead. */
#define DEF_FAN_BPT(T, SUF)                                                       \
int orc_fan_bpt_##SUF(const orc_geom* g, const T* img, T* grad, const double* xpos, \
                      const double* ypos, const double* sc, int scaled)           \
{                                                                                 \
    const double cents = ((double)(g->ns - 1)) / 2 + g->offset_s;                 \
    unsigned int ix, iy, ia, is1, is2;                                            \
    size_t i, n = (size_t)g->ns * g->na;                                          \
    for (i = 0; i < n; ++i) grad[i] = 0.0;                                        \
    for (iy = 0; iy < g->ny; ++iy) {                                              \
        for (ix = 0; ix < g->nx; ++ix) {                                          \
            const T v = img[ix + iy * g->nx];                                     \
            const double x = xpos[ix], y = ypos[iy];                              \
            for (ia = 0; ia < g->na; ++ia) {                                      \
                const double sb = sc[2 * ia], cb = sc[2 * ia + 1];                \
                const double d = g->dso + x * sb - y * cb;                        \
                const double r = x * cb + y * sb;                                 \
                const double w = g->dsd * g->dsd / (d * d + r * r);               \
                const double s = atan2(r, d) * g->dsd / g->ds + cents;            \
                if (s >= 0 && s <= g->ns - 1) {                                   \
                    double u;                                                     \
                    is1 = (unsigned int)floor(s);                                 \
                    is2 = (unsigned int)ceil(s);                                  \
                    u = s - is1;                                                  \
                    grad[is1 + ia * g->ns] += v * (T)(w * (1 - u));               \
                    grad[is2 + ia * g->ns] += v * (T)(w * u);                     \
                }                                                                 \
            }                                                                     \
        }                                                                         \
    }                                                                             \
    if (scaled) {                                                                 \
        const T wt = (T)fan_scale(g);                                             \
        for (i = 0; i < n; ++i) grad[i] *= wt;                                    \
    }                                                                             \
    return 1;                                                                     \
}
DEF_FAN_BPT(float, f)
DEF_FAN_BPT(double, d)

/* ------------------------------------------------------------------ */
/* Parallel-beam tables: src/bp_par_2d.cc:18-49.
 * xcos[ia + ix*na] = x*cos(angle), ysin[ia + iy*na] = y*sin(angle). */
int orc_par_prep(const orc_geom* g, double* xcos, double* ysin)
{
    const double centx = ((double)(g->nx - 1)) / 2 + g->offset_x;
    const double centy = ((double)(g->ny - 1)) / 2 + g->offset_y;
    double angle = g->orbit_start;
    unsigned int ia, ix, iy;
    for (ia = 0; ia < g->na; ++ia) {
        double sa, ca;
        sincos(angle, &sa, &ca);
        double posx = -centx * g->dx, posy = centy * g->dy;
        for (ix = 0; ix < g->nx; ++ix) { xcos[ia + ix * g->na] = posx * ca; posx += g->dx; }
        for (iy = 0; iy < g->ny; ++iy) { ysin[ia + iy * g->na] = posy * sa; posy -= g->dy; }
        angle += g->orbit;
    }
    return 1;
}

/* Pixel-driven parallel BP, src/bp_par_2d.cc:51-90 (signed orbit scale). */
#define DEF_PAR_BP(T, SUF)                                                        \
int orc_par_bp_##SUF(const orc_geom* g, const T* proj, T* img, const double* xcos, \
                     const double* ysin)                                          \
{                                                                                 \
    const double cents = ((double)(g->ns - 1)) / 2 + g->offset_s;                 \
    unsigned int ix, iy, ia, is1, is2;                                            \
    for (iy = 0; iy < g->ny; ++iy) {                                              \
        for (ix = 0; ix < g->nx; ++ix) {                                          \
            double sum = 0.0;                                                     \
            for (ia = 0; ia < g->na; ++ia) {                                      \
                const double s = (xcos[ia + ix * g->na] + ysin[ia + iy * g->na])  \
                                 / g->ds + cents;                                 \
                if (s >= 0 && s <= g->ns - 1) {                                   \
                    double u;                                                     \
                    is1 = (unsigned int)floor(s);                                 \
                    is2 = (unsigned int)ceil(s);                                  \
                    u = s - is1;                                                  \
                    sum += (1 - u) * proj[is1 + ia * g->ns]                       \
                           + u * proj[is2 + ia * g->ns];                          \
                }                                                                 \
            }                                                                     \
            img[ix + iy * g->nx] = (T)(sum * g->orbit);                           \
        }                                                                         \
    }                                                                             \
    return 1;                                                                     \
}
DEF_PAR_BP(float, f)
DEF_PAR_BP(double, d)

/* Ray-casting parallel FP tables, ParallelProjection2DRayCastingPrepare
 * (src/fp_par_2d.cc:18-89): per view (sin, cos) of the accumulated angle;
 * per ray the entry point (x1, y1) in pixel-index units on the FOV circle of
 * radius fov and the number of steps of length `step` (index units along the
 * marched axis: x when sin > 1/1.414, else y); rays missing the circle get 0
 * steps and a zero entry point. */
int orc_fp_rc_prep(const orc_geom* g, double step, double* sincostbl, double* begins,
                   int* nsteps)
{
    const double xc = (double)(g->nx - 1) / 2.0 + g->offset_x;
    const double yc = (double)(g->ny - 1) / 2.0 + g->offset_y;
    const double fov2 = g->fov * g->fov;
    const double s0 = -((double)(g->ns - 1) / 2.0 + g->offset_s) * g->ds;
    double angle = g->orbit_start;
    unsigned int ia, is;
    for (ia = 0; ia < g->na; ++ia) {
        double sa, ca, s = s0;
        sincos(angle, &sa, &ca);
        sincostbl[2 * ia] = sa;
        sincostbl[2 * ia + 1] = ca;
        {
            const double stx = step * sa, sty = step * ca;
            const int usex = sa > (1.0 / 1.414);
            for (is = 0; is < g->ns; ++is) {
                const size_t r = (size_t)ia * g->ns + is;
                const double h2 = fov2 - s * s;
                if (h2 > 0) {
                    const double h = sqrt(h2);
                    const double x1 = (s * ca - h * sa) / g->dx + xc;
                    const double y1 = -(s * sa + h * ca) / g->dy + yc;
                    const double x2 = (s * ca + h * sa) / g->dx + xc;
                    const double y2 = -(s * sa - h * ca) / g->dy + yc;
                    nsteps[r] = usex ? (int)((x2 - x1) / stx) : (int)((y2 - y1) / sty);
                    begins[2 * r] = x1;
                    begins[2 * r + 1] = y1;
                } else {
                    nsteps[r] = 0;
                    begins[2 * r] = 0;
                    begins[2 * r + 1] = 0;
                }
                s += g->ds;
            }
        }
        angle += g->orbit;
    }
    return 1;
}

/* Ray-casting parallel FP, ParallelProjection2DRayCasting<T>
 * (src/fp_par_2d.cc:91-144): from the entry point, nsteps steps of
 * (step sin, step cos) in index units; samples inside [0, nx-1] x [0, ny-1]
 * add the bilinear interpolation of the image (FP64 sum), and the ray is
 * sum * step * dx. */
#define DEF_FP_RC(T, SUF)                                                         \
int orc_fp_rc_##SUF(const orc_geom* g, double step, const T* img, T* proj,       \
                    const double* sincostbl, const double* begins, const int* nsteps) \
{                                                                                 \
    unsigned int ia, is;                                                          \
    for (ia = 0; ia < g->na; ++ia) {                                              \
        const double sa = sincostbl[2 * ia], ca = sincostbl[2 * ia + 1];          \
        const double stx = step * sa, sty = step * ca;                            \
        for (is = 0; is < g->ns; ++is) {                                          \
            const size_t r = (size_t)ia * g->ns + is;                             \
            const int n = nsteps[r];                                              \
            if (n > 0) {                                                          \
                double x = begins[2 * r], y = begins[2 * r + 1], sum = 0;         \
                int k;                                                            \
                for (k = 0; k < n; ++k) {                                         \
                    if (x >= 0 && x <= g->nx - 1 && y >= 0 && y <= g->ny - 1) {   \
                        const int ix1 = (int)floor(x), ix2 = (int)ceil(x);        \
                        const int iy1 = (int)floor(y), iy2 = (int)ceil(y);        \
                        const double u = x - ix1, v = y - iy1;                    \
                        sum += (1 - v) * ((1 - u) * img[ix1 + iy1 * g->nx]        \
                                          + u * img[ix2 + iy1 * g->nx])           \
                               + v * ((1 - u) * img[ix1 + iy2 * g->nx]            \
                                      + u * img[ix2 + iy2 * g->nx]);              \
                    }                                                             \
                    x += stx;                                                     \
                    y += sty;                                                     \
                }                                                                 \
                proj[r] = (T)(sum * step * g->dx);                                \
            } else {                                                              \
                proj[r] = 0;                                                      \
            }                                                                     \
        }                                                                         \
    }                                                                             \
    return 1;                                                                     \
}
DEF_FP_RC(float, f)
DEF_FP_RC(double, d)

/* Single-view parallel BP, src/bp_par_2d_sv.cc:54-92: every (pixel, view)
 * keeps its own value, img[(iy*nx + ix)*na + ia] = the view's linear
 * interpolation at s (zero off the detector), with no sum over views and no
 * scale.  Its tables are the parallel BP's: src/bp_par_2d_sv.cc:18-49
 * repeats src/bp_par_2d.cc:18-49 (orc_par_prep). */
#define DEF_PAR_SVBP(T, SUF)                                                      \
int orc_par_svbp_##SUF(const orc_geom* g, const T* proj, T* img,                  \
                       const double* xcos, const double* ysin)                    \
{                                                                                 \
    const double cents = ((double)(g->ns - 1)) / 2 + g->offset_s;                 \
    unsigned int ix, iy, ia, is1, is2;                                            \
    for (iy = 0; iy < g->ny; ++iy) {                                              \
        for (ix = 0; ix < g->nx; ++ix) {                                          \
            T* out = img + ((size_t)iy * g->nx + ix) * g->na;                     \
            for (ia = 0; ia < g->na; ++ia) {                                      \
                const double s = (xcos[ia + ix * g->na] + ysin[ia + iy * g->na])  \
                                 / g->ds + cents;                                 \
                if (s >= 0 && s <= g->ns - 1) {                                   \
                    double u;                                                     \
                    is1 = (unsigned int)floor(s);                                 \
                    is2 = (unsigned int)ceil(s);                                  \
                    u = s - is1;                                                  \
                    out[ia] = (T)((1 - u) * proj[is1 + ia * g->ns]                \
                                  + u * proj[is2 + ia * g->ns]);                  \
                } else {                                                          \
                    out[ia] = 0;                                                  \
                }                                                                 \
            }                                                                     \
        }                                                                         \
    }                                                                             \
    return 1;                                                                     \
}
DEF_PAR_SVBP(float, f)
DEF_PAR_SVBP(double, d)

/* Transpose of the parallel BP (scatter form, FP64 accumulation).  The
 * reference's transpose (src/bp_par_2d.cc:95-199) reads uninitialised
 * pointers, so this one is the mathematically intended operator, pinned by
 * the adjoint test against orc_par_bp. */
#define DEF_PAR_BPT(T, SUF)                                                       \
int orc_par_bpt_##SUF(const orc_geom* g, const T* img, T* grad, const double* xcos, \
                      const double* ysin, double* work)                           \
{                                                                                 \
    const double cents = ((double)(g->ns - 1)) / 2 + g->offset_s;                 \
    unsigned int ix, iy, ia, is1, is2;                                            \
    size_t i, n = (size_t)g->ns * g->na;                                          \
    for (i = 0; i < n; ++i) work[i] = 0.0;                                        \
    for (iy = 0; iy < g->ny; ++iy) {                                              \
        for (ix = 0; ix < g->nx; ++ix) {                                          \
            const double v = img[ix + iy * g->nx];                                \
            for (ia = 0; ia < g->na; ++ia) {                                      \
                const double s = (xcos[ia + ix * g->na] + ysin[ia + iy * g->na])  \
                                 / g->ds + cents;                                 \
                if (s >= 0 && s <= g->ns - 1) {                                   \
                    double u;                                                     \
                    is1 = (unsigned int)floor(s);                                 \
                    is2 = (unsigned int)ceil(s);                                  \
                    u = s - is1;                                                  \
                    work[is1 + ia * g->ns] += (1 - u) * v;                        \
                    work[is2 + ia * g->ns] += u * v;                              \
                }                                                                 \
            }                                                                     \
        }                                                                         \
    }                                                                             \
    for (i = 0; i < n; ++i) grad[i] = (T)(work[i] * g->orbit);                    \
    return 1;                                                                     \
}
DEF_PAR_BPT(float, f)
DEF_PAR_BPT(double, d)

/* ------------------------------------------------------------------ */
/* Ray-driven parallel FP tables, src/fp_par_2d.cc:149-208.
 * sincostbl[2a]=sin, [2a+1]=cos of the wrapped, accumulated angle;
 * beginoffset[2(a*ns+s)] = begin + s*offset2, [..+1] = offset;
 * usex[a] = 0 when angle in [pi/4,3pi/4) U [5pi/4,7pi/4). */
int orc_fp_prep(const orc_geom* g, double* sincostbl, double* beginoffset, int* usex)
{
    const double centx = (double)(g->nx - 1) / 2.0 + g->offset_x;
    const double centy = (double)(g->ny - 1) / 2.0 + g->offset_y;
    const double cents = (double)(g->ns - 1) / 2.0 + g->offset_s;
    double angle = g->orbit_start;
    unsigned int ia, is;
    double* bp = beginoffset;
    for (ia = 0; ia < g->na; ++ia) {
        double sa, ca, offset, offset2, begin;
        while (angle < 0) angle += 2 * M_PI;
        while (angle >= 2 * M_PI) angle -= 2 * M_PI;
        sincos(angle, &sa, &ca);
        sincostbl[2 * ia] = sa;
        sincostbl[2 * ia + 1] = ca;
        if ((angle >= M_PI_4 && angle < 3 * M_PI_4) ||
            (angle >= 5 * M_PI_4 && angle < 7 * M_PI_4))
            usex[ia] = 0;
        else
            usex[ia] = 1;
        if (usex[ia]) {
            offset = g->dy * sa / (ca * g->dx);
            offset2 = g->ds / (ca * g->dx);
            begin = centx - centy * offset - cents * offset2;
        } else {
            offset = g->dx * ca / (sa * g->dy);
            offset2 = -g->ds / (sa * g->dy);
            begin = centy - centx * offset - cents * offset2;
        }
        for (is = 0; is < g->ns; ++is) {
            bp[0] = begin;
            bp[1] = offset;
            begin += offset2;
            bp += 2;
        }
        angle += g->orbit;
    }
    return 1;
}

/* Ray-driven parallel FP, src/fp_par_2d.cc:210-280.  The running sum is
 * of type T (the reference accumulates in T, :219); positions are FP64
 * and advanced by repeated addition (:247, :271). */
#define DEF_FP(T, SUF)                                                            \
int orc_fp_##SUF(const orc_geom* g, const T* img, T* proj, const double* sc,      \
                 const double* beginoffset, const int* usex)                      \
{                                                                                 \
    unsigned int ia, is, k, i1, i2;                                               \
    const double* bp = beginoffset;                                               \
    T* out = proj;                                                                \
    for (ia = 0; ia < g->na; ++ia) {                                              \
        if (usex[ia]) {                                                           \
            const double dw = fabs(g->dy / sc[2 * ia + 1]);                       \
            for (is = 0; is < g->ns; ++is) {                                      \
                double x = bp[0];                                                 \
                const double off = bp[1];                                         \
                T sum = 0.0;                                                      \
                for (k = 0; k < g->ny; ++k) {                                     \
                    if (x >= 0 && x <= g->nx - 1) {                               \
                        double u;                                                 \
                        i1 = (unsigned int)floor(x);                              \
                        i2 = (unsigned int)ceil(x);                               \
                        u = x - i1;                                               \
                        sum += (1 - u) * img[i1 + k * g->nx] + u * img[i2 + k * g->nx]; \
                    }                                                             \
                    x += off;                                                     \
                }                                                                 \
                *out++ = (T)(sum * dw);                                           \
                bp += 2;                                                          \
            }                                                                     \
        } else {                                                                  \
            const double dw = fabs(g->dx / sc[2 * ia]);                           \
            for (is = 0; is < g->ns; ++is) {                                      \
                double y = bp[0];                                                 \
                const double off = bp[1];                                         \
                T sum = 0.0;                                                      \
                for (k = 0; k < g->nx; ++k) {                                     \
                    if (y >= 0 && y <= g->ny - 1) {                               \
                        double u;                                                 \
                        i1 = (unsigned int)floor(y);                              \
                        i2 = (unsigned int)ceil(y);                               \
                        u = y - i1;                                               \
                        sum += (1 - u) * img[k + i1 * g->nx] + u * img[k + i2 * g->nx]; \
                    }                                                             \
                    y += off;                                                     \
                }                                                                 \
                *out++ = (T)(sum * dw);                                           \
                bp += 2;                                                          \
            }                                                                     \
        }                                                                         \
    }                                                                             \
    return 1;                                                                     \
}
DEF_FP(float, f)
DEF_FP(double, d)

/* Tables of the FP transpose, src/fp_par_2d.cc:285-330.  No angle wrap
 * here (the reference's GradPrep does not wrap, :301-307). pos is laid out
 * [na][nx or ny][2]; the reference indexes it with a stride of 2*nx per view,
 * so it requires nx == ny (as its TF op enforces). */
int orc_fp_grad_prep(const orc_geom* g, double* weights, double* pos, int* usex)
{
    const double centx = (double)(g->nx - 1) / 2.0 + g->offset_x;
    const double centy = (double)(g->ny - 1) / 2.0 + g->offset_y;
    const double cents = (double)(g->ns - 1) / 2.0 + g->offset_s;
    double angle = g->orbit_start;
    unsigned int ia, k;
    double* pp = pos;
    for (ia = 0; ia < g->na; ++ia) {
        double sa, ca;
        sincos(angle, &sa, &ca);
        const int bx = !((angle >= M_PI_4 && angle < 3 * M_PI_4) ||
                         (angle >= 5 * M_PI_4 && angle < 7 * M_PI_4));
        double o1, o2, begin;
        usex[ia] = bx;
        if (bx) {
            weights[ia] = fabs(g->dy / ca);
            o1 = g->dx * ca / g->ds;
            o2 = -g->dy * sa / g->ds;
            begin = cents - centx * o1 - centy * o2;
            for (k = 0; k < g->ny; ++k) { pp[0] = begin; pp[1] = o1; begin += o2; pp += 2; }
        } else {
            weights[ia] = fabs(g->dx / sa);
            o1 = -g->dy * sa / g->ds;
            o2 = g->dx * ca / g->ds;
            begin = cents - centy * o1 - centx * o2;
            for (k = 0; k < g->nx; ++k) { pp[0] = begin; pp[1] = o1; begin += o2; pp += 2; }
        }
        angle += g->orbit;
    }
    return 1;
}

/* FP transpose in gather form, src/fp_par_2d.cc:332-391.  Differences
 * from the reference, both where the reference is undefined: the channel
 * loop is clamped to [0, ns-1] (the reference converts a negative ceil() to
 * unsigned and reads past the row when right > ns-1, :375-378).  Inside the
 * detector the result is the reference's. */
#define DEF_FP_GRAD(T, SUF)                                                       \
int orc_fp_grad_##SUF(const orc_geom* g, const T* proj, T* img,                   \
                      const double* weights, const double* pos, const int* usex)  \
{                                                                                 \
    unsigned int ix, iy, ia;                                                      \
    for (iy = 0; iy < g->ny; ++iy) {                                              \
        for (ix = 0; ix < g->nx; ++ix) {                                          \
            double sum = 0.0;                                                     \
            for (ia = 0; ia < g->na; ++ia) {                                      \
                const T* pr = proj + (size_t)ia * g->ns;                          \
                double begin, off, mid, left, right, tmp = 0.0, c;                \
                if (usex[ia]) {                                                   \
                    const double* pp = pos + 2 * ((size_t)ia * g->nx + iy);       \
                    begin = pp[0]; off = pp[1];                                   \
                    mid = begin + off * ix;                                       \
                    left = ix == 0 ? mid : mid - off;                             \
                    right = ix == g->nx - 1 ? mid : mid + off;                    \
                } else {                                                          \
                    const double* pp = pos + 2 * ((size_t)ia * g->nx + ix);       \
                    begin = pp[0]; off = pp[1];                                   \
                    mid = begin + off * iy;                                       \
                    left = iy == 0 ? mid : mid - off;                             \
                    right = iy == g->ny - 1 ? mid : mid + off;                    \
                }                                                                 \
                if (left > right) { double t = left; left = right; right = t; }   \
                c = ceil(left);                                                   \
                if (c < 0) c = 0;                                                 \
                for (; c <= right && c <= (double)(g->ns - 1); c += 1.0) {        \
                    const unsigned int is = (unsigned int)c;                      \
                    tmp += (1 - fabs((mid - (double)is) / off)) * pr[is];         \
                }                                                                 \
                sum += tmp * weights[ia];                                         \
            }                                                                     \
            img[ix + iy * g->nx] = (T)sum;                                        \
        }                                                                         \
    }                                                                             \
    return 1;                                                                     \
}
DEF_FP_GRAD(float, f)
DEF_FP_GRAD(double, d)

/* ------------------------------------------------------------------ */
/* Distance-driven parallel FP tables, src/fp_par_2d.cc:396-457: per view the
 * angle wrapped into [0, 2pi), (sin, cos), the branch (x-branch outside
 * [pi/4, 3pi/4) and [5pi/4, 7pi/4)), and per channel (begin, offset,
 * half-width): the channel's footprint centre on the first row (x-branch) or
 * column (y-branch) in pixel units, accumulated channel by channel, its
 * shift per row or column, and half its width. */
int orc_fp_dd_prep(const orc_geom* g, double* sincostbl, double* bo, int* usex)
{
    const double centx = (double)(g->nx - 1) / 2.0 + g->offset_x;
    const double centy = (double)(g->ny - 1) / 2.0 + g->offset_y;
    const double cents = (double)(g->ns - 1) / 2.0 + g->offset_s;
    double angle = g->orbit_start;
    unsigned int ia, is;
    double* bp = bo;
    for (ia = 0; ia < g->na; ++ia) {
        double sa, ca, off, off2, begin;
        while (angle < 0) angle += 2 * M_PI;
        while (angle >= 2 * M_PI) angle -= 2 * M_PI;
        sincos(angle, &sa, &ca);
        sincostbl[2 * ia] = sa;
        sincostbl[2 * ia + 1] = ca;
        usex[ia] = !((angle >= M_PI_4 && angle < 3 * M_PI_4) ||
                     (angle >= 5 * M_PI_4 && angle < 7 * M_PI_4));
        if (usex[ia]) {
            off = g->dy * sa / (ca * g->dx);
            off2 = g->ds / (ca * g->dx);
            begin = centx - centy * off - cents * off2;
        } else {
            off = g->dx * ca / (sa * g->dy);
            off2 = -g->ds / (sa * g->dy);
            begin = centy - centx * off - cents * off2;
        }
        for (is = 0; is < g->ns; ++is) {
            bp[0] = begin;
            bp[1] = off;
            bp[2] = off2 / 2;
            begin += off2;
            bp += 3;
        }
        angle += g->orbit;
    }
    return 1;
}

/* Distance-driven parallel FP, src/fp_par_2d.cc:459-543.  Per channel and
 * per row (x-branch; per column for the y-branch) the footprint
 * [left, right] = centre -+ half-width, shifted by `offset` from row to row,
 * averages the cell values (img[v] + img[v+1]) / 2 of the cells v it
 * overlaps, weighted by the overlap length; the row averages are summed and
 * scaled by |dy/cos| (|dx/sin|).  Arithmetic as the reference's: each term
 * in double, the sums tsum, lsum and sum in T.  The reference's cell index
 * is unsigned: a footprint starting left of the first cell turns
 * floor(left) < 0 into a huge index, the loop does not run and the row adds
 * nothing (x86-64 g++ build); that rule is kept.  Cells v >= n-1 are
 * skipped (the pair v, v+1 must exist). */
#define DEF_FP_DD(T, SUF)                                                         \
int orc_fp_dd_##SUF(const orc_geom* g, const T* img, T* proj, const double* sc,   \
                    const double* bo, const int* usex)                            \
{                                                                                 \
    unsigned int ia, is, u, v;                                                    \
    for (ia = 0; ia < g->na; ++ia) {                                              \
        const int bx = usex[ia];                                                  \
        const double dw = bx ? fabs(g->dy / sc[2 * ia + 1]) : fabs(g->dx / sc[2 * ia]); \
        const unsigned int nu = bx ? g->ny : g->nx, nv = bx ? g->nx : g->ny;     \
        const size_t su = bx ? g->nx : 1, sv = bx ? 1 : g->nx;                   \
        for (is = 0; is < g->ns; ++is) {                                          \
            const double* b = bo + 3 * ((size_t)ia * g->ns + is);                 \
            double left = b[0] - b[2], right = b[0] + b[2];                       \
            const double off = b[1];                                              \
            T sum = 0;                                                            \
            if (left > right) { double t = left; left = right; right = t; }       \
            for (u = 0; u < nu; ++u) {                                            \
                T tsum = 0, lsum = 0;                                             \
                const double fl = floor(left);                                    \
                if (fl >= 0) {                                                    \
                    for (v = (unsigned int)fl; v < ceil(right); ++v) {            \
                        if (v >= nv - 1) continue;                                \
                        const double len = fmin((double)(v + 1), right) -         \
                                           fmax((double)v, left);                 \
                        tsum += len * (img[u * su + v * sv] +                     \
                                       img[u * su + (v + 1) * sv]) / 2;           \
                        lsum += len;                                              \
                    }                                                             \
                }                                                                 \
                if (lsum > 0) sum += tsum / lsum;                                 \
                left += off;                                                      \
                right += off;                                                     \
            }                                                                     \
            proj[(size_t)ia * g->ns + is] = (T)(sum * dw);                        \
        }                                                                         \
    }                                                                             \
    return 1;                                                                     \
}
DEF_FP_DD(float, f)
DEF_FP_DD(double, d)

/* Gradient of the distance-driven FP as the reference defines it,
 * src/fp_par_2d.cc:601-655 (the TF op's gradient for method "disdriven").
 * Its tables (src/fp_par_2d.cc:548-599) are those of orc_fp_grad_prep:
 * weights |dy/cos| or |dx/sin| and per row or column (begin, o1).  Per pixel
 * and view: mid = begin + o1 * i along the branch's axis, the span
 * [mid - o1, mid + o1] (one-sided at the first and last pixel), and every
 * channel is whose cell [is - 1/2, is + 1/2] it overlaps adds
 * overlap * proj[is] / 2; sums in double.  Like the reference's unsigned
 * channel index, a span with ceil(left - 1/2) < 0 adds nothing (kept);
 * channels past ns - 1, which the reference reads beyond the row, are not
 * read here. */
#define DEF_FP_DD_GRAD(T, SUF)                                                    \
int orc_fp_dd_grad_##SUF(const orc_geom* g, const T* proj, T* img,                \
                         const double* weights, const double* pos, const int* usex) \
{                                                                                 \
    unsigned int ix, iy, ia, is;                                                  \
    for (iy = 0; iy < g->ny; ++iy) {                                              \
        for (ix = 0; ix < g->nx; ++ix) {                                          \
            double sum = 0.0;                                                     \
            for (ia = 0; ia < g->na; ++ia) {                                      \
                const T* pr = proj + (size_t)ia * g->ns;                          \
                double begin, off, mid, left, right, tmp = 0.0, c0;               \
                if (usex[ia]) {                                                   \
                    const double* pp = pos + 2 * ((size_t)ia * g->nx + iy);       \
                    begin = pp[0]; off = pp[1];                                   \
                    mid = begin + off * ix;                                       \
                    left = ix == 0 ? mid : mid - off;                             \
                    right = ix == g->nx - 1 ? mid : mid + off;                    \
                } else {                                                          \
                    const double* pp = pos + 2 * ((size_t)ia * g->nx + ix);       \
                    begin = pp[0]; off = pp[1];                                   \
                    mid = begin + off * iy;                                       \
                    left = iy == 0 ? mid : mid - off;                             \
                    right = iy == g->ny - 1 ? mid : mid + off;                    \
                }                                                                 \
                if (left > right) { double t = left; left = right; right = t; }   \
                c0 = ceil(left - 0.5);                                            \
                if (c0 >= 0) {                                                    \
                    for (is = (unsigned int)c0; is <= floor(right + 0.5) && is < g->ns; ++is) { \
                        const double len = fmin((double)is + 0.5, right) -        \
                                           fmax((double)is - 0.5, left);          \
                        tmp += len * pr[is] / 2;                                  \
                    }                                                             \
                }                                                                 \
                sum += tmp * weights[ia];                                         \
            }                                                                     \
            img[ix + iy * g->nx] = (T)sum;                                        \
        }                                                                         \
    }                                                                             \
    return 1;                                                                     \
}
DEF_FP_DD_GRAD(float, f)
DEF_FP_DD_GRAD(double, d)

/* ------------------------------------------------------------------ */
/* Helical rebinning of a 3-D multi-row sinogram proj_in[na][nt][ns] into
 * nz 2-D fan sinograms proj_out[nz][view_per_rot][ns]: 360LI
 * (src/helical_interpolation.cc:81-160, the helper HelicalInterpolation360
 * uses) and 180LI (:162-269, HelicalInterpolation180).  The geometry record
 * for these two: ns, nt, na, ds, dt (unused by the reference's arithmetic),
 * offset_s, offset_t, orbit_start (unused), orbit, couch_begin, couch_mov,
 * view_per_rot, nz, dz, z0, thickness, dso, dsd.  Slice z advances by
 * accumulation (z += dz, :304-306).  Per output (slice, view, channel) the
 * detector rows a slice of that thickness reaches are averaged in the view
 * (cur) and in the conjugate view (cor: one rotation away for 360LI; the
 * opposite ray, view cur - (pi - 2 gamma)/orbit with linear interpolation
 * in view and channel, for 180LI), and the two are interpolated linearly in
 * z.  A window entirely off the detector is not used.  Returns 1, or where
 * the reference throws: 0 for a slice whose +-1 rotation (360LI) or half
 * rotation (180LI) of views is not in the scan (nothing further written),
 * -1 for a sample with neither window on the detector (NaN there).  -2: a
 * 180LI conjugate view that interpolates from a view outside the scan (the
 * reference reads outside its input there; NaN here). */
typedef struct {
    unsigned int ns, nt, na;
    double ds, dt, offset_s, offset_t, orbit_start, orbit, couch_begin, couch_mov;
    unsigned int view_per_rot, nz;
    double dz, z0, thickness, dso, dsd;
} orc_heli;

/* the rows [floor(upper), ceil(bottom)) of the window z_dif -+ thickness/2 */
static int heli_window(const orc_heli* h, double center_t, double z_dif, double cg,
                       int* it0, int* it1)
{
    double upper_t = center_t - (z_dif + h->thickness / 2.0) / h->dso * cg * h->dsd;
    double bottom_t = center_t - (z_dif - h->thickness / 2.0) / h->dso * cg * h->dsd;
    if (bottom_t < 0 || upper_t >= h->nt) return 0;
    upper_t = upper_t < 0 ? 0 : upper_t;
    bottom_t = bottom_t >= h->nt ? h->nt : bottom_t;
    *it0 = (int)floor(upper_t);
    *it1 = (int)ceil(bottom_t);
    return 1;
}

/* 1 when every slice's view window lies in the scan: [cv - vpr, cv + vpr - 1]
 * for 360LI, [cv - vpr/2, cv - vpr/2 + vpr - 1] for 180LI.  The reference
 * checks slice by slice and throws at the first short one, after writing
 * the slices before it; the restatement writes nothing then. */
static int heli_views_cover(const orc_heli* h, int mode)
{
    const int vpr = (int)h->view_per_rot;
    double z = h->z0;
    for (unsigned int iz = 0; iz < h->nz; ++iz, z += h->dz) {
        const int cv = (int)round((z - h->couch_begin) / h->couch_mov);
        const int begin = mode == 360 ? cv - vpr : cv - vpr / 2;
        const int end = mode == 360 ? cv + vpr - 1 : begin + vpr - 1;
        if (begin < 0 || end >= (int)h->na) return 0;
    }
    return 1;
}

#define DEF_HELI360(T, SUF)                                                       \
int orc_heli360_##SUF(const orc_heli* h, const T* in, T* out)                     \
{                                                                                 \
    const int vpr = (int)h->view_per_rot;                                         \
    const double center_t = (h->nt - 1) / 2.0 + h->offset_t;                     \
    const double center_s = (h->ns - 1) / 2.0 + h->offset_s;                     \
    double z = h->z0;                                                             \
    int ok = 1;                                                                   \
    if (!heli_views_cover(h, 360)) return 0;                                      \
    for (unsigned int iz = 0; iz < h->nz; ++iz, z += h->dz) {                     \
        T* po = out + (size_t)iz * vpr * h->ns;                                   \
        const int center_view = (int)round((z - h->couch_begin) / h->couch_mov);  \
        const int begin = center_view - vpr;                                      \
        const int first_view = (begin / vpr + 1) * vpr;                          \
        for (int iv = 0; iv < vpr; ++iv) {                                        \
            const int cur_view = first_view + iv;                                 \
            const double z_dif = z - h->couch_begin - cur_view * h->couch_mov;    \
            const int cor_view = cur_view >= center_view ? cur_view - vpr : cur_view + vpr; \
            const double z_dif2 = z - h->couch_begin - cor_view * h->couch_mov;   \
            for (int is = 0; is < (int)h->ns; ++is) {                             \
                const double s = (is - center_s) * h->ds;                         \
                const double cg = cos(s / h->dsd);                                \
                double cur = 0, cor = 0;                                          \
                int a0, a1, use_cur, use_cor, it, cnt;                            \
                use_cur = heli_window(h, center_t, z_dif, cg, &a0, &a1);          \
                if (use_cur) {                                                    \
                    for (cnt = 0, it = a0; it < a1; ++it, ++cnt)                  \
                        cur += in[is + ((size_t)it + (size_t)cur_view * h->nt) * h->ns]; \
                    cur /= cnt;                                                   \
                }                                                                 \
                use_cor = heli_window(h, center_t, z_dif2, cg, &a0, &a1);         \
                if (use_cor) {                                                    \
                    for (cnt = 0, it = a0; it < a1; ++it, ++cnt)                  \
                        cor += in[is + ((size_t)it + (size_t)cor_view * h->nt) * h->ns]; \
                    cor /= cnt;                                                   \
                }                                                                 \
                T* o = po + is + (size_t)iv * h->ns;                              \
                if (use_cur && use_cor)                                           \
                    *o = (T)((z_dif * cor - z_dif2 * cur) / (z_dif - z_dif2));    \
                else if (use_cur)                                                 \
                    *o = (T)cur;                                                  \
                else if (use_cor)                                                 \
                    *o = (T)cor;                                                  \
                else {                                                            \
                    *o = (T)NAN;                                                  \
                    if (ok > 0) ok = -1;                                          \
                }                                                                 \
            }                                                                     \
        }                                                                         \
    }                                                                             \
    return ok;                                                                    \
}
DEF_HELI360(float, f)
DEF_HELI360(double, d)

#define DEF_HELI180(T, SUF)                                                       \
int orc_heli180_##SUF(const orc_heli* h, const T* in, T* out)                     \
{                                                                                 \
    const int vpr = (int)h->view_per_rot;                                         \
    const double center_t = (h->nt - 1) / 2.0 + h->offset_t;                     \
    const double center_s = (h->ns - 1) / 2.0 + h->offset_s;                     \
    double z = h->z0;                                                             \
    int ok = 1;                                                                   \
    if (!heli_views_cover(h, 180)) return 0;                                      \
    for (unsigned int iz = 0; iz < h->nz; ++iz, z += h->dz) {                     \
        T* po = out + (size_t)iz * vpr * h->ns;                                   \
        const int center_view = (int)round((z - h->couch_begin) / h->couch_mov);  \
        const int begin = center_view - vpr / 2, end = begin + vpr - 1;           \
        const int first_view = (int)(ceil(begin / (double)vpr) * h->view_per_rot); \
        for (int iv = 0; iv < vpr; ++iv) {                                        \
            int cur_view = first_view + iv;                                       \
            if (cur_view > end) cur_view -= vpr;                                  \
            const double z_dif = z - h->couch_begin - cur_view * h->couch_mov;    \
            for (int is = 0; is < (int)h->ns; ++is) {                             \
                const double s = (is - center_s) * h->ds;                         \
                const double gamma = s / h->dsd;                                  \
                const double cg = cos(gamma);                                     \
                double cur = 0, cor = 0, z_dif2, cv;                              \
                int a0, a1, use_cur, use_cor, it, cnt;                            \
                T* o = po + is + (size_t)iv * h->ns;                              \
                use_cur = heli_window(h, center_t, z_dif, cg, &a0, &a1);          \
                if (use_cur) {                                                    \
                    for (cnt = 0, it = a0; it < a1; ++it, ++cnt)                  \
                        cur += in[is + ((size_t)it + (size_t)cur_view * h->nt) * h->ns]; \
                    cur /= cnt;                                                   \
                }                                                                 \
                cv = cur_view - (M_PI - 2 * gamma) / h->orbit;                    \
                cv = cv < begin ? cv + h->view_per_rot : cv;                      \
                cv = cv > end ? cv - h->view_per_rot : cv;                        \
                z_dif2 = z - h->couch_begin - cv * h->couch_mov;                  \
                use_cor = heli_window(h, center_t, z_dif2, cg, &a0, &a1);         \
                if (use_cor) {                                                    \
                    const int vl = (int)floor(cv), vr = (int)ceil(cv);            \
                    const double u = cv - vl;                                     \
                    const double ss = center_s - s / h->ds;                       \
                    int sl, sr;                                                   \
                    double v, lv = 0, rv = 0;                                     \
                    if (ss < 0 || ss > h->ns - 1) {                               \
                        *o = (T)cur;  /* the conjugate channel is off the detector */ \
                        continue;                                                 \
                    }                                                             \
                    if (vl < 0 || vr >= (int)h->na) {                             \
                        *o = (T)NAN;                                              \
                        ok = -2;                                                  \
                        continue;                                                 \
                    }                                                             \
                    sl = (int)floor(ss);                                          \
                    sr = (int)ceil(ss);                                           \
                    v = ss - sl;                                                  \
                    for (cnt = 0, it = a0; it < a1; ++it, ++cnt) {                \
                        const T* rl = in + ((size_t)it + (size_t)vl * h->nt) * h->ns; \
                        const T* rr = in + ((size_t)it + (size_t)vr * h->nt) * h->ns; \
                        lv += (1 - v) * rl[sl] + v * rl[sr];                      \
                        rv += (1 - v) * rr[sl] + v * rr[sr];                      \
                    }                                                             \
                    cor = ((1 - u) * lv + u * rv) / cnt;                          \
                }                                                                 \
                if (use_cur && use_cor)                                           \
                    *o = (T)((z_dif * cor - z_dif2 * cur) / (z_dif - z_dif2));    \
                else if (use_cur)                                                 \
                    *o = (T)cur;                                                  \
                else if (use_cor)                                                 \
                    *o = (T)cor;                                                  \
                else {                                                            \
                    *o = (T)NAN;                                                  \
                    if (ok > 0) ok = -1;                                          \
                }                                                                 \
            }                                                                     \
        }                                                                         \
    }                                                                             \
    return ok;                                                                    \
}
DEF_HELI180(float, f)
DEF_HELI180(double, d)

/* ------------------------------------------------------------------ */
/* Ramp filter taps filter[ns + n], n in [-ns, ns], src/filter.cc:18-65.
 * Returns 0 for an unknown type (:62). */
int orc_ramp_prep(const orc_geom* g, double* filter)
{
    double* pf = filter + g->ns;
    int n, ns = (int)g->ns;
    if (g->type == 0 || g->type == 2) {
        const double ds2_inv = 1.0 / (g->ds * g->ds);
        const double pi2_inv = 1.0 / (M_PI * M_PI);
        for (n = -ns; n <= ns; ++n) {
            if (n == 0) pf[n] = 0.25 * ds2_inv;
            else if (n % 2 == 0) pf[n] = 0;
            else pf[n] = -pi2_inv * ds2_inv / (double)(n * n);
        }
    } else if (g->type == 1) {
        const double k = 1.0 / (M_PI * M_PI * g->dsd * g->dsd);
        for (n = -ns; n <= ns; ++n) {
            if (n == 0) pf[n] = 0.25 / (g->ds * g->ds);
            else if (n % 2 == 0) pf[n] = 0;
            else {
                const double sa = sin((double)n * g->ds / g->dsd);
                pf[n] = -k / (sa * sa);
            }
        }
    } else {
        return 0;
    }
    return 1;
}

/* Linear convolution per row, src/filter.cc:67-91 (FP64 sum, all taps). */
#define DEF_RAMP(T, SUF)                                                          \
int orc_ramp_##SUF(const orc_geom* g, const T* in, const double* filter, T* out)  \
{                                                                                 \
    const double* pf = filter + g->ns;                                            \
    const int ns = (int)g->ns;                                                    \
    unsigned int row;                                                             \
    int is, n;                                                                    \
    for (row = 0; row < g->nrow; ++row) {                                         \
        const T* ip = in + (size_t)row * g->ns;                                   \
        T* op = out + (size_t)row * g->ns;                                        \
        for (is = 0; is < ns; ++is) {                                             \
            double sum = 0;                                                       \
            for (n = -ns; n <= ns; ++n) {                                         \
                const int k = is - n;                                             \
                if (k >= 0 && k < ns) sum += ip[k] * pf[n];                       \
            }                                                                     \
            op[is] = (T)(sum * g->ds);                                            \
        }                                                                         \
    }                                                                             \
    return 1;                                                                     \
}
DEF_RAMP(float, f)
DEF_RAMP(double, d)

/* Correlation (the transpose), src/filter.cc:96-124; zero taps skipped. */
#define DEF_RAMP_GRAD(T, SUF)                                                     \
int orc_ramp_grad_##SUF(const orc_geom* g, const T* in, const double* filter, T* out) \
{                                                                                 \
    const double* pf = filter + g->ns;                                            \
    const int ns = (int)g->ns;                                                    \
    unsigned int row;                                                             \
    int is, n;                                                                    \
    for (row = 0; row < g->nrow; ++row) {                                         \
        const T* ip = in + (size_t)row * g->ns;                                   \
        T* op = out + (size_t)row * g->ns;                                        \
        for (is = 0; is < ns; ++is) {                                             \
            double sum = 0;                                                       \
            for (n = -ns; n <= ns; ++n) {                                         \
                const int k = is + n;                                             \
                if (k >= 0 && k < ns && (pf[n] > 0 || pf[n] < 0))                 \
                    sum += ip[k] * pf[n];                                         \
            }                                                                     \
            op[is] = (T)(sum * g->ds);                                            \
        }                                                                         \
    }                                                                             \
    return 1;                                                                     \
}
DEF_RAMP_GRAD(float, f)
DEF_RAMP_GRAD(double, d)

/* ------------------------------------------------------------------ */
/* Fan / flat cosine weighting, src/fan_weighting.cc:19-59.  The transpose
 * is the same diagonal (:39-56), so one routine serves both (grad=1 only
 * changes the operand order of the product, as the reference does). */
#define DEF_FANW(T, SUF)                                                          \
int orc_fanw_##SUF(const orc_geom* g, const T* in, T* out, int grad)              \
{                                                                                 \
    const double cents = (double)(g->ns - 1) / 2 + g->offset_s;                   \
    int is, ia;                                                                   \
    for (is = 0; is < (int)g->ns; ++is) {                                         \
        const double s = g->ds * ((double)is - cents);                            \
        double w;                                                                 \
        if (g->type == 2) w = g->dso * fabs(cos(atan2(s, g->dsd))) / g->dsd;      \
        else if (g->type == 1) w = g->dso * fabs(cos(s / g->dsd)) / g->dsd;       \
        else return 0;                                                            \
        for (ia = 0; ia < (int)g->nrow; ++ia) {                                   \
            const size_t i = is + (size_t)ia * g->ns;                             \
            out[i] = grad ? (T)(in[i] * w) : (T)(w * in[i]);                      \
        }                                                                         \
    }                                                                             \
    return 1;                                                                     \
}
DEF_FANW(float, f)
DEF_FANW(double, d)
