// ref_shim.cpp -- extern "C" entry points over the reference's own CPU classes.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the unchanged reference sources src/{bp_fan_2d,bp_fan_2d_angle,bp_par_2d,
// bp_par_2d_sv,fp_par_2d,filter,fan_weighting,
// helical_interpolation}.cc, read where they lie under the reference
// checkout, into oracle/_ref/libctref.so.  Nothing of the reference is copied
// into this repository; the built library is git-ignored.  It is used to
// write golden vectors (tests/golden/make_golden.py), to cross-check the C
// restatement in oracle/ct_oracle.c, and as bench.py's CPU baseline.
//
// Each function builds the reference's Param struct in its own constructor
// order and calls calculate_on_cpu.  The geometry record is the one declared
// in oracle/ct_oracle.c (orc_geom).
#include "include/bp_fan_2d.h"
#include "include/bp_fan_2d_angle.h"
#include "include/bp_par_2d.h"
#include "include/bp_par_2d_sv.h"
#include "include/fp_par_2d.h"
#include "include/filter.h"
#include "include/fan_weighting.h"
#include "include/helical_interpolation.h"

#include <stdexcept>

using namespace ct_recon;

extern "C" {

struct orc_geom {
    unsigned int ns, na;
    double ds, orbit, offset_s, orbit_start;
    unsigned int nx, ny;
    double dx, dy, offset_x, offset_y;
    double dso, dsd, fov;
    int type;
    unsigned int nrow;
};

static FanBackprojection2DParam fan_param(const orc_geom* g)
{
    return FanBackprojection2DParam(g->ns, g->na, g->ds, g->orbit, g->offset_s, g->orbit_start,
                                    g->nx, g->ny, g->dx, g->dy, g->offset_x, g->offset_y,
                                    g->dso, g->dsd, g->fov);
}

static FanBackprojection2DAngleParam fan_angle_param(const orc_geom* g)
{
    return FanBackprojection2DAngleParam(g->ns, g->na, g->ds, g->offset_s, g->nx, g->ny, g->dx,
                                         g->dy, g->offset_x, g->offset_y, g->dso, g->dsd, g->fov);
}

static ParallelBackprojection2DParam par_param(const orc_geom* g)
{
    return ParallelBackprojection2DParam(g->ns, g->na, g->ds, g->orbit, g->offset_s,
                                         g->orbit_start, g->nx, g->ny, g->dx, g->dy,
                                         g->offset_x, g->offset_y, g->fov);
}

static ParallelProjection2DParam fp_param(const orc_geom* g)
{
    return ParallelProjection2DParam(g->ns, g->na, g->ds, g->orbit, g->offset_s, g->orbit_start,
                                     g->nx, g->ny, g->dx, g->dy, g->offset_x, g->offset_y, g->fov);
}

int ref_fan_prep(const orc_geom* g, double* xpos, double* ypos, double* sc)
{
    return FanBackprojection2DPixDrivenPrep(fan_param(g)).calculate_on_cpu(xpos, ypos, sc);
}

int ref_fan_grad_prep(const orc_geom* g, double* xpos, double* ypos, double* sc)
{
    return FanBackprojection2DPixDrivenGradPrep(fan_param(g)).calculate_on_cpu(xpos, ypos, sc);
}

int ref_fan_bp_f(const orc_geom* g, const float* p, float* img, const double* x,
                 const double* y, const double* sc)
{
    return FanBackprojection2DPixDriven<float>(fan_param(g)).calculate_on_cpu(p, img, x, y, sc);
}

int ref_fan_bp_d(const orc_geom* g, const double* p, double* img, const double* x,
                 const double* y, const double* sc)
{
    return FanBackprojection2DPixDriven<double>(fan_param(g)).calculate_on_cpu(p, img, x, y, sc);
}

int ref_fan_bpt_f(const orc_geom* g, const float* img, float* grad, const double* x,
                  const double* y, const double* sc)
{
    return FanBackprojection2DPixDrivenGrad<float>(fan_param(g))
        .calculate_on_cpu(img, grad, x, y, sc);
}

int ref_fan_bpt_d(const orc_geom* g, const double* img, double* grad, const double* x,
                  const double* y, const double* sc)
{
    return FanBackprojection2DPixDrivenGrad<double>(fan_param(g))
        .calculate_on_cpu(img, grad, x, y, sc);
}

int ref_fan_angle_prep(const orc_geom* g, const double* angles, double* xpos, double* ypos,
                       double* sc)
{
    return FanBackprojection2DAnglePixDrivenPrep(fan_angle_param(g))
        .calculate_on_cpu(angles, xpos, ypos, sc);
}

int ref_fan_angle_bp_f(const orc_geom* g, const float* p, float* img, const double* x,
                       const double* y, const double* sc)
{
    return FanBackprojection2DAnglePixDriven<float>(fan_angle_param(g))
        .calculate_on_cpu(p, img, x, y, sc);
}

int ref_fan_angle_bp_d(const orc_geom* g, const double* p, double* img, const double* x,
                       const double* y, const double* sc)
{
    return FanBackprojection2DAnglePixDriven<double>(fan_angle_param(g))
        .calculate_on_cpu(p, img, x, y, sc);
}

int ref_par_prep(const orc_geom* g, double* xcos, double* ysin)
{
    int unused = 0;
    return ParallelBackprojection2DPixDrivenPrep(par_param(g)).calculate_on_cpu(xcos, ysin, &unused);
}

int ref_par_bp_f(const orc_geom* g, const float* p, float* img, const double* xcos,
                 const double* ysin)
{
    return ParallelBackprojection2DPixDriven<float>(par_param(g))
        .calculate_on_cpu(p, img, xcos, ysin, nullptr);
}

int ref_par_bp_d(const orc_geom* g, const double* p, double* img, const double* xcos,
                 const double* ysin)
{
    return ParallelBackprojection2DPixDriven<double>(par_param(g))
        .calculate_on_cpu(p, img, xcos, ysin, nullptr);
}

static ParallelSingleViewBp2DParam svbp_param(const orc_geom* g)
{
    return ParallelSingleViewBp2DParam(g->ns, g->na, g->ds, g->orbit, g->offset_s,
                                       g->orbit_start, g->nx, g->ny, g->dx, g->dy,
                                       g->offset_x, g->offset_y, g->fov);
}

int ref_par_svbp_prep(const orc_geom* g, double* xcos, double* ysin)
{
    int unused = 0;
    return ParallelSingleViewBp2DPixDrivenPrep(svbp_param(g)).calculate_on_cpu(xcos, ysin,
                                                                               &unused);
}

int ref_par_svbp_f(const orc_geom* g, const float* p, float* img, const double* xcos,
                   const double* ysin)
{
    return ParallelSingleViewBp2DPixDriven<float>(svbp_param(g))
        .calculate_on_cpu(p, img, xcos, ysin, nullptr);
}

int ref_par_svbp_d(const orc_geom* g, const double* p, double* img, const double* xcos,
                   const double* ysin)
{
    return ParallelSingleViewBp2DPixDriven<double>(svbp_param(g))
        .calculate_on_cpu(p, img, xcos, ysin, nullptr);
}

int ref_fp_rc_prep(const orc_geom* g, double step, double* sc, double* begins, int* nsteps)
{
    return ParallelProjection2DRayCastingPrepare(fp_param(g), step)
        .calculate_on_cpu(sc, begins, nsteps);
}

int ref_fp_rc_f(const orc_geom* g, double step, const float* img, float* proj, const double* sc,
                const double* begins, const int* nsteps)
{
    return ParallelProjection2DRayCasting<float>(fp_param(g), step)
        .calculate_on_cpu(img, proj, sc, begins, nsteps);
}

int ref_fp_rc_d(const orc_geom* g, double step, const double* img, double* proj,
                const double* sc, const double* begins, const int* nsteps)
{
    return ParallelProjection2DRayCasting<double>(fp_param(g), step)
        .calculate_on_cpu(img, proj, sc, begins, nsteps);
}

int ref_fp_prep(const orc_geom* g, double* sc, double* bo, int* usex)
{
    return ParallelProjection2DRayDrivenPrepare(fp_param(g)).calculate_on_cpu(sc, bo, usex);
}

int ref_fp_f(const orc_geom* g, const float* img, float* proj, const double* sc,
             const double* bo, const int* usex)
{
    return ParallelProjection2DRayDriven<float>(fp_param(g)).calculate_on_cpu(img, proj, sc, bo, usex);
}

int ref_fp_d(const orc_geom* g, const double* img, double* proj, const double* sc,
             const double* bo, const int* usex)
{
    return ParallelProjection2DRayDriven<double>(fp_param(g))
        .calculate_on_cpu(img, proj, sc, bo, usex);
}

int ref_fp_grad_prep(const orc_geom* g, double* weights, double* pos, int* usex)
{
    return ParallelProjection2DRayDrivenGradPrep(fp_param(g)).calculate_on_cpu(weights, pos, usex);
}

int ref_fp_grad_f(const orc_geom* g, const float* proj, float* img, const double* w,
                  const double* pos, const int* usex)
{
    return ParallelProjection2DRayDrivenGrad<float>(fp_param(g))
        .calculate_on_cpu(proj, img, w, pos, usex);
}

int ref_fp_grad_d(const orc_geom* g, const double* proj, double* img, const double* w,
                  const double* pos, const int* usex)
{
    return ParallelProjection2DRayDrivenGrad<double>(fp_param(g))
        .calculate_on_cpu(proj, img, w, pos, usex);
}

int ref_fp_dd_prep(const orc_geom* g, double* sc, double* bo, int* usex)
{
    return ParallelProjection2DDisDrivenPrep(fp_param(g)).calculate_on_cpu(sc, bo, usex);
}

int ref_fp_dd_f(const orc_geom* g, const float* img, float* proj, const double* sc,
                const double* bo, const int* usex)
{
    return ParallelProjection2DDisDriven<float>(fp_param(g)).calculate_on_cpu(img, proj, sc, bo, usex);
}

int ref_fp_dd_d(const orc_geom* g, const double* img, double* proj, const double* sc,
                const double* bo, const int* usex)
{
    return ParallelProjection2DDisDriven<double>(fp_param(g))
        .calculate_on_cpu(img, proj, sc, bo, usex);
}

int ref_fp_dd_grad_prep(const orc_geom* g, double* weights, double* pos, int* usex)
{
    return ParallelProjection2DDisDrivenGradPrep(fp_param(g)).calculate_on_cpu(weights, pos, usex);
}

int ref_fp_dd_grad_f(const orc_geom* g, const float* proj, float* img, const double* w,
                     const double* pos, const int* usex)
{
    return ParallelProjection2DDisDrivenGrad<float>(fp_param(g))
        .calculate_on_cpu(proj, img, w, pos, usex);
}

int ref_fp_dd_grad_d(const orc_geom* g, const double* proj, double* img, const double* w,
                     const double* pos, const int* usex)
{
    return ParallelProjection2DDisDrivenGrad<double>(fp_param(g))
        .calculate_on_cpu(proj, img, w, pos, usex);
}

static FilterParam filter_param(const orc_geom* g)
{
    return FilterParam(g->ns, g->nrow, g->ds, g->dsd, g->type, 0);
}

int ref_ramp_prep(const orc_geom* g, double* filter)
{
    return RampFilterPrep(filter_param(g)).calculate_on_cpu(filter);
}

int ref_ramp_f(const orc_geom* g, const float* in, const double* filter, float* out)
{
    return RampFilter<float>(filter_param(g)).calculate_on_cpu(in, filter, out);
}

int ref_ramp_d(const orc_geom* g, const double* in, const double* filter, double* out)
{
    return RampFilter<double>(filter_param(g)).calculate_on_cpu(in, filter, out);
}

int ref_ramp_grad_f(const orc_geom* g, const float* in, const double* filter, float* out)
{
    return RampFilterGrad<float>(filter_param(g)).calculate_on_cpu(in, filter, out);
}

int ref_ramp_grad_d(const orc_geom* g, const double* in, const double* filter, double* out)
{
    return RampFilterGrad<double>(filter_param(g)).calculate_on_cpu(in, filter, out);
}

static FanWeightingParam fanw_param(const orc_geom* g)
{
    return FanWeightingParam(g->ns, g->nrow, g->ds, g->offset_s, g->dso, g->dsd, g->type);
}

int ref_fanw_f(const orc_geom* g, const float* in, float* out, int grad)
{
    return grad ? FanWeightingGrad<float>(fanw_param(g)).calculate_on_cpu(in, out)
                : FanWeighting<float>(fanw_param(g)).calculate_on_cpu(in, out);
}

int ref_fanw_d(const orc_geom* g, const double* in, double* out, int grad)
{
    return grad ? FanWeightingGrad<double>(fanw_param(g)).calculate_on_cpu(in, out)
                : FanWeighting<double>(fanw_param(g)).calculate_on_cpu(in, out);
}

// Helical rebinning: the record is orc_heli of oracle/ct_oracle.c.  The
// reference throws std::runtime_error for a slice without enough views or a
// sample with no detector row in either window; both return 0 here.
struct orc_heli {
    unsigned int ns, nt, na;
    double ds, dt, offset_s, offset_t, orbit_start, orbit, couch_begin, couch_mov;
    unsigned int view_per_rot, nz;
    double dz, z0, thickness, dso, dsd;
};

static HelicalInterpolationParam heli_param(const orc_heli* h)
{
    return HelicalInterpolationParam(h->ns, h->nt, h->na, h->ds, h->dt, h->offset_s, h->offset_t,
                                     h->orbit_start, h->orbit, h->couch_begin, h->couch_mov,
                                     h->view_per_rot, h->nz, h->dz, h->z0, h->thickness, h->dso,
                                     h->dsd);
}

}  // extern "C"

template <class Op, typename T>
static int heli_run(const orc_heli* h, const T* in, T* out)
{
    try {
        return Op(heli_param(h)).calculate_on_cpu(in, out) ? 1 : 0;
    } catch (const std::runtime_error&) {
        return 0;
    }
}

extern "C" {

int ref_heli360_f(const orc_heli* h, const float* in, float* out)
{
    return heli_run<HelicalInterpolation360<float>>(h, in, out);
}

int ref_heli360_d(const orc_heli* h, const double* in, double* out)
{
    return heli_run<HelicalInterpolation360<double>>(h, in, out);
}

int ref_heli180_f(const orc_heli* h, const float* in, float* out)
{
    return heli_run<HelicalInterpolation180<float>>(h, in, out);
}

int ref_heli180_d(const orc_heli* h, const double* in, double* out)
{
    return heli_run<HelicalInterpolation180<double>>(h, in, out);
}

}  // extern "C"
