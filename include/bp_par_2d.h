/*
 * bp_par_2d.h -- parallel-beam pixel-driven backprojection and its exact
 * transpose: the class API of the reference's include/bp_par_2d.h:20-222
 * (pixel-driven classes; the distance-driven stubs of :227-298 always
 * returned false and are not declared) with hipStream_t GPU entry points.
 * Conventions as in bp_fan_2d.h.
 */
#ifndef CT_RECON_MI355X_BP_PAR_2D_H_
#define CT_RECON_MI355X_BP_PAR_2D_H_

#include <hip/hip_runtime_api.h>

namespace ct_recon {

struct ParallelBackprojection2DParam {
    unsigned int ns, na;
    double ds, orbit, offset_s, orbit_start;
    unsigned int nx, ny;
    double dx, dy, offset_x, offset_y;
    double fov;

    ParallelBackprojection2DParam() {}
    ParallelBackprojection2DParam(unsigned int ns, unsigned int na, double ds, double orbit,
                                  double offset_s, double orbit_start, unsigned int nx,
                                  unsigned int ny, double dx, double dy, double offset_x,
                                  double offset_y, double fov)
        : ns(ns), na(na), ds(ds), orbit(orbit), offset_s(offset_s), orbit_start(orbit_start),
          nx(nx), ny(ny), dx(dx), dy(dy), offset_x(offset_x), offset_y(offset_y), fov(fov)
    {
    }
};

class ParallelBackprojection2DPrepare {
public:
    explicit ParallelBackprojection2DPrepare(const ParallelBackprojection2DParam& param)
        : param_(param)
    {
    }
    virtual ~ParallelBackprojection2DPrepare() {}
    virtual bool calculate_on_cpu(double* buffer1, double* buffer2, int* buffer3) const = 0;
    virtual bool calculate_on_gpu(double* buffer1, double* buffer2, int* buffer3,
                                  hipStream_t) const = 0;

protected:
    ParallelBackprojection2DParam param_;
};

template <typename T>
class ParallelBackprojection2D {
public:
    explicit ParallelBackprojection2D(const ParallelBackprojection2DParam& param) : param_(param) {}
    virtual ~ParallelBackprojection2D() {}
    virtual bool calculate_on_cpu(const T* proj, T* img, const double*, const double*,
                                  const int*) const = 0;
    virtual bool calculate_on_gpu(const T* proj, T* img, const double*, const double*,
                                  const int*, hipStream_t) const = 0;

protected:
    ParallelBackprojection2DParam param_;
};

class ParallelBackprojection2DGradPrep {
public:
    explicit ParallelBackprojection2DGradPrep(const ParallelBackprojection2DParam& param)
        : param_(param)
    {
    }
    virtual ~ParallelBackprojection2DGradPrep() {}
    virtual bool calculate_on_cpu(double* buffer1, double* buffer2, bool* buffer3) const = 0;
    virtual bool calculate_on_gpu(double* buffer1, double* buffer2, bool* buffer3,
                                  hipStream_t) const = 0;

protected:
    ParallelBackprojection2DParam param_;
};

template <typename T>
class ParallelBackprojection2DGrad {
public:
    explicit ParallelBackprojection2DGrad(const ParallelBackprojection2DParam& param)
        : param_(param)
    {
    }
    virtual ~ParallelBackprojection2DGrad() {}
    virtual bool calculate_on_cpu(const T* img, T* grad, const double*, const double*,
                                  const bool*) const = 0;
    virtual bool calculate_on_gpu(const T* img, T* grad, const double*, const double*,
                                  const bool*, hipStream_t) const = 0;

protected:
    ParallelBackprojection2DParam param_;
};

// Tables: xcos[ix*na + ia] = x cos(theta), ysin[iy*na + ia] = y sin(theta);
// buffer3 unused (src/bp_par_2d.cc:18-49).
class ParallelBackprojection2DPixDrivenPrep : public ParallelBackprojection2DPrepare {
public:
    explicit ParallelBackprojection2DPixDrivenPrep(const ParallelBackprojection2DParam& param)
        : ParallelBackprojection2DPrepare(param)
    {
    }
    bool calculate_on_cpu(double* xcos, double* ysin, int* buffer) const override;
    bool calculate_on_gpu(double* xcos, double* ysin, int* buffer, hipStream_t) const override;
};

template <typename T>
class ParallelBackprojection2DPixDriven : public ParallelBackprojection2D<T> {
public:
    explicit ParallelBackprojection2DPixDriven(const ParallelBackprojection2DParam& param)
        : ParallelBackprojection2D<T>(param)
    {
    }
    bool calculate_on_cpu(const T* proj, T* img, const double* xcos, const double* ysin,
                          const int*) const override;
    bool calculate_on_gpu(const T* proj, T* img, const double* xcos, const double* ysin,
                          const int*, hipStream_t) const override;
};

// The reference's transpose tables: begins[na*ns], offsets[na], usex[na]
// (src/bp_par_2d.cc:95-145), filled for compatibility; the transpose below
// does not need them.
class ParallelBackprojection2DPixDrivenGradPrep : public ParallelBackprojection2DGradPrep {
public:
    explicit ParallelBackprojection2DPixDrivenGradPrep(const ParallelBackprojection2DParam& param)
        : ParallelBackprojection2DGradPrep(param)
    {
    }
    bool calculate_on_cpu(double* begins, double* offsets, bool* usex) const override;
    bool calculate_on_gpu(double* begins, double* offsets, bool* usex,
                          hipStream_t) const override;
};

// Exact transpose of ParallelBackprojection2DPixDriven (the reference's,
// src/bp_par_2d.cc:147-199, reads uninitialised pointers).
template <typename T>
class ParallelBackprojection2DPixDrivenGrad : public ParallelBackprojection2DGrad<T> {
public:
    explicit ParallelBackprojection2DPixDrivenGrad(const ParallelBackprojection2DParam& param)
        : ParallelBackprojection2DGrad<T>(param)
    {
    }
    bool calculate_on_cpu(const T* img, T* grad, const double*, const double*,
                          const bool*) const override;
    bool calculate_on_gpu(const T* img, T* grad, const double*, const double*, const bool*,
                          hipStream_t) const override;
};

}  // namespace ct_recon

#endif
