/*
 * bp_par_2d_sv.h -- single-view parallel-beam pixel-driven backprojection:
 * the class API of the reference's include/bp_par_2d_sv.h:17-143 with
 * hipStream_t GPU entry points.  The operator writes every (pixel, view)
 * value, img[(iy*nx + ix)*na + ia], with no sum and no scale; the Prep fills
 * the parallel BP's tables (xcos[ix*na + ia], ysin[iy*na + ia]).
 * Conventions as in bp_fan_2d.h: calculate_on_cpu of the operator takes
 * host pointers (computed on the GPU), calculate_on_gpu takes device
 * pointers and synchronises the stream before returning.
 */
#ifndef CT_RECON_MI355X_BP_PAR_2D_SV_H_
#define CT_RECON_MI355X_BP_PAR_2D_SV_H_

#include <hip/hip_runtime_api.h>

namespace ct_recon {

struct ParallelSingleViewBp2DParam {
    unsigned int ns, na;
    double ds, orbit, offset_s, orbit_start;
    unsigned int nx, ny;
    double dx, dy, offset_x, offset_y;
    double fov;
    ParallelSingleViewBp2DParam() {}
    ParallelSingleViewBp2DParam(unsigned int ns, unsigned int na, double ds, double orbit,
                                double offset_s, double orbit_start, unsigned int nx,
                                unsigned int ny, double dx, double dy, double offset_x,
                                double offset_y, double fov)
        : ns(ns), na(na), ds(ds), orbit(orbit), offset_s(offset_s), orbit_start(orbit_start),
          nx(nx), ny(ny), dx(dx), dy(dy), offset_x(offset_x), offset_y(offset_y), fov(fov)
    {
    }
};

class ParallelSingleViewBp2DPrepare {
public:
    explicit ParallelSingleViewBp2DPrepare(const ParallelSingleViewBp2DParam& param)
        : param_(param)
    {
    }
    virtual ~ParallelSingleViewBp2DPrepare() {}
    virtual bool calculate_on_cpu(double* buffer1, double* buffer2, int* buffer3) const = 0;
    virtual bool calculate_on_gpu(double* buffer1, double* buffer2, int* buffer3,
                                  hipStream_t) const = 0;

protected:
    ParallelSingleViewBp2DParam param_;
};

template <typename T>
class ParallelSingleViewBp2D {
public:
    explicit ParallelSingleViewBp2D(const ParallelSingleViewBp2DParam& param) : param_(param) {}
    virtual ~ParallelSingleViewBp2D() {}
    virtual bool calculate_on_cpu(const T* proj, T* img, const double*, const double*,
                                  const int*) const = 0;
    virtual bool calculate_on_gpu(const T* proj, T* img, const double*, const double*,
                                  const int*, hipStream_t) const = 0;

protected:
    ParallelSingleViewBp2DParam param_;
};

// src/bp_par_2d_sv.cc:18-49: xcos, ysin; the int buffer is unused.
class ParallelSingleViewBp2DPixDrivenPrep : public ParallelSingleViewBp2DPrepare {
public:
    explicit ParallelSingleViewBp2DPixDrivenPrep(const ParallelSingleViewBp2DParam& param)
        : ParallelSingleViewBp2DPrepare(param)
    {
    }
    bool calculate_on_cpu(double* xcos, double* ysin, int* buffer) const override;
    bool calculate_on_gpu(double* xcos, double* ysin, int* buffer, hipStream_t) const override;
};

// src/bp_par_2d_sv.cc:54-92 (ctr_par_sv_bp, one slice).
template <typename T>
class ParallelSingleViewBp2DPixDriven : public ParallelSingleViewBp2D<T> {
public:
    explicit ParallelSingleViewBp2DPixDriven(const ParallelSingleViewBp2DParam& param)
        : ParallelSingleViewBp2D<T>(param)
    {
    }
    bool calculate_on_cpu(const T* proj, T* img, const double* xcos, const double* ysin,
                          const int* buffer) const override;
    bool calculate_on_gpu(const T* proj, T* img, const double* xcos, const double* ysin,
                          const int* buffer, hipStream_t) const override;
};

}  // namespace ct_recon

#endif  // CT_RECON_MI355X_BP_PAR_2D_SV_H_
