/*
 * fp_par_2d.h -- parallel-beam forward projection: the class API of the
 * reference's include/fp_par_2d.h:20-352 (ray casting, ray-driven with its
 * gather-form transpose, distance-driven with the reference's gradient) with
 * hipStream_t GPU entry points.  Conventions as in bp_fan_2d.h.
 */
#ifndef CT_RECON_MI355X_FP_PAR_2D_H_
#define CT_RECON_MI355X_FP_PAR_2D_H_

#include <hip/hip_runtime_api.h>

namespace ct_recon {

struct ParallelProjection2DParam {
    unsigned int ns, na;
    double ds, orbit, offset_s, orbit_start;
    unsigned int nx, ny;
    double dx, dy, offset_x, offset_y;
    double fov;

    ParallelProjection2DParam() {}
    ParallelProjection2DParam(unsigned int ns, unsigned int na, double ds, double orbit,
                              double offset_s, double orbit_start, unsigned int nx,
                              unsigned int ny, double dx, double dy, double offset_x,
                              double offset_y, double fov)
        : ns(ns), na(na), ds(ds), orbit(orbit), offset_s(offset_s), orbit_start(orbit_start),
          nx(nx), ny(ny), dx(dx), dy(dy), offset_x(offset_x), offset_y(offset_y), fov(fov)
    {
    }
};

class ParallelProjection2DPrepare {
public:
    explicit ParallelProjection2DPrepare(const ParallelProjection2DParam& param) : param_(param) {}
    virtual ~ParallelProjection2DPrepare() {}
    virtual bool calculate_on_cpu(double* buffer1, double* buffer2, int* buffer3) const = 0;
    virtual bool calculate_on_gpu(double* buffer1, double* buffer2, int* buffer3,
                                  hipStream_t) const = 0;

protected:
    ParallelProjection2DParam param_;
};

template <typename T>
class ParallelProjection2D {
public:
    explicit ParallelProjection2D(const ParallelProjection2DParam& param) : param_(param) {}
    virtual ~ParallelProjection2D() {}
    virtual bool calculate_on_cpu(const T* img, T* proj, const double*, const double*,
                                  const int*) const = 0;
    virtual bool calculate_on_gpu(const T* img, T* proj, const double*, const double*,
                                  const int*, hipStream_t) const = 0;

protected:
    ParallelProjection2DParam param_;
};

class ParallelProjection2DGradPrepare {
public:
    explicit ParallelProjection2DGradPrepare(const ParallelProjection2DParam& param)
        : param_(param)
    {
    }
    virtual ~ParallelProjection2DGradPrepare() {}
    virtual bool calculate_on_cpu(double* buffer1, double* buffer2, int* buffer3) const = 0;
    virtual bool calculate_on_gpu(double* buffer1, double* buffer2, int* buffer3,
                                  hipStream_t) const = 0;

protected:
    ParallelProjection2DParam param_;
};

template <typename T>
class ParallelProjection2DGrad {
public:
    explicit ParallelProjection2DGrad(const ParallelProjection2DParam& param) : param_(param) {}
    virtual ~ParallelProjection2DGrad() {}
    virtual bool calculate_on_cpu(const T* proj, T* img, const double*, const double*,
                                  const int*) const = 0;
    virtual bool calculate_on_gpu(const T* proj, T* img, const double*, const double*,
                                  const int*, hipStream_t) const = 0;

protected:
    ParallelProjection2DParam param_;
};

// Ray casting (src/fp_par_2d.cc:18-144).  Tables: sincostbl[2na],
// begins[na][ns][2], nsteps[na][ns]; step_size in pixels (the reference's
// default 0.125).
class ParallelProjection2DRayCastingPrepare : public ParallelProjection2DPrepare {
public:
    explicit ParallelProjection2DRayCastingPrepare(const ParallelProjection2DParam& param,
                                                   const double step_size = 0.125)
        : ParallelProjection2DPrepare(param), step_size_(step_size)
    {
    }
    bool calculate_on_cpu(double* sincostbl, double* begins, int* nsteps) const override;
    bool calculate_on_gpu(double* sincostbl, double* begins, int* nsteps,
                          hipStream_t) const override;

protected:
    double step_size_;
};

// img [ny][nx] -> proj [na][ns], bit-identical to src/fp_par_2d.cc:91-144.
template <typename T>
class ParallelProjection2DRayCasting : public ParallelProjection2D<T> {
public:
    explicit ParallelProjection2DRayCasting(const ParallelProjection2DParam& param,
                                            const double step_size = 0.125)
        : ParallelProjection2D<T>(param), step_size_(step_size)
    {
    }
    bool calculate_on_cpu(const T* img, T* proj, const double* sincostbl, const double* begins,
                          const int* nsteps) const override;
    bool calculate_on_gpu(const T* img, T* proj, const double* sincostbl, const double* begins,
                          const int* nsteps, hipStream_t) const override;

protected:
    double step_size_;
};

// Tables: sincostbl[2na], beginoffset[na][ns][2], usex[na]
// (src/fp_par_2d.cc:149-208).
class ParallelProjection2DRayDrivenPrepare : public ParallelProjection2DPrepare {
public:
    explicit ParallelProjection2DRayDrivenPrepare(const ParallelProjection2DParam& param)
        : ParallelProjection2DPrepare(param)
    {
    }
    bool calculate_on_cpu(double* sincostbl, double* beginoffset, int* usex) const override;
    bool calculate_on_gpu(double* sincostbl, double* beginoffset, int* usex,
                          hipStream_t) const override;
};

// img [ny][nx] -> proj [na][ns] (src/fp_par_2d.cc:210-280).
template <typename T>
class ParallelProjection2DRayDriven : public ParallelProjection2D<T> {
public:
    explicit ParallelProjection2DRayDriven(const ParallelProjection2DParam& param)
        : ParallelProjection2D<T>(param)
    {
    }
    bool calculate_on_cpu(const T* img, T* proj, const double* sincostbl,
                          const double* beginoffset, const int* usex) const override;
    bool calculate_on_gpu(const T* img, T* proj, const double* sincostbl,
                          const double* beginoffset, const int* usex, hipStream_t) const override;
};

// Tables: weights[na], pos[na][nx][2], usex[na] (src/fp_par_2d.cc:285-330);
// the layout requires nx == ny, as in the reference.
class ParallelProjection2DRayDrivenGradPrep : public ParallelProjection2DGradPrepare {
public:
    explicit ParallelProjection2DRayDrivenGradPrep(const ParallelProjection2DParam& param)
        : ParallelProjection2DGradPrepare(param)
    {
    }
    bool calculate_on_cpu(double* weights, double* pos, int* usex) const override;
    bool calculate_on_gpu(double* weights, double* pos, int* usex,
                          hipStream_t stream) const override;
};

// proj [na][ns] -> img [ny][nx], gather form (src/fp_par_2d.cc:332-391).
template <typename T>
class ParallelProjection2DRayDrivenGrad : public ParallelProjection2DGrad<T> {
public:
    explicit ParallelProjection2DRayDrivenGrad(const ParallelProjection2DParam& param)
        : ParallelProjection2DGrad<T>(param)
    {
    }
    bool calculate_on_cpu(const T* proj, T* img, const double* weights, const double* pos,
                          const int* usex) const override;
    bool calculate_on_gpu(const T* proj, T* img, const double* weights, const double* pos,
                          const int* usex, hipStream_t stream) const override;
};

// Distance-driven FP (src/fp_par_2d.cc:396-543).  Tables: sincostbl[2na],
// beginoffset[na][ns][3] = (footprint centre, shift per row or column, half
// width), usex[na].
class ParallelProjection2DDisDrivenPrep : public ParallelProjection2DPrepare {
public:
    explicit ParallelProjection2DDisDrivenPrep(const ParallelProjection2DParam& param)
        : ParallelProjection2DPrepare(param)
    {
    }
    bool calculate_on_cpu(double* sincostbl, double* beginoffset, int* usex) const override;
    bool calculate_on_gpu(double* sincostbl, double* beginoffset, int* usex,
                          hipStream_t) const override;
};

// img [ny][nx] -> proj [na][ns], bit-identical to the reference's CPU path.
template <typename T>
class ParallelProjection2DDisDriven : public ParallelProjection2D<T> {
public:
    explicit ParallelProjection2DDisDriven(const ParallelProjection2DParam& param)
        : ParallelProjection2D<T>(param)
    {
    }
    bool calculate_on_cpu(const T* img, T* proj, const double* sincostbl,
                          const double* beginoffset, const int* usex) const override;
    bool calculate_on_gpu(const T* img, T* proj, const double* sincostbl,
                          const double* beginoffset, const int* usex, hipStream_t) const override;
};

// Tables of the distance-driven gradient (src/fp_par_2d.cc:548-599): those of
// the ray-driven transpose, weights[na], pos[na][nx][2], usex[na]; nx == ny.
class ParallelProjection2DDisDrivenGradPrep : public ParallelProjection2DGradPrepare {
public:
    explicit ParallelProjection2DDisDrivenGradPrep(const ParallelProjection2DParam& param)
        : ParallelProjection2DGradPrepare(param)
    {
    }
    bool calculate_on_cpu(double* weights, double* pos, int* usex) const override;
    bool calculate_on_gpu(double* weights, double* pos, int* usex,
                          hipStream_t stream) const override;
};

// proj [na][ns] -> img [ny][nx]: the reference's distance-driven gradient
// (src/fp_par_2d.cc:601-655), an approximate transpose of the FP above.
template <typename T>
class ParallelProjection2DDisDrivenGrad : public ParallelProjection2DGrad<T> {
public:
    explicit ParallelProjection2DDisDrivenGrad(const ParallelProjection2DParam& param)
        : ParallelProjection2DGrad<T>(param)
    {
    }
    bool calculate_on_cpu(const T* proj, T* img, const double* weights, const double* pos,
                          const int* usex) const override;
    bool calculate_on_gpu(const T* proj, T* img, const double* weights, const double* pos,
                          const int* usex, hipStream_t stream) const override;
};

}  // namespace ct_recon

#endif
