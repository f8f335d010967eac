/*
 * bp_fan_2d.h -- equiangular fan-beam pixel-driven backprojection and its
 * transpose: the C++ class API of the reference's include/bp_fan_2d.h:20-227
 * (same namespace, class names, constructor order and table layouts), with
 * the GPU entry points taking a hipStream_t and running the MI355X kernels.
 *
 * Prep classes build the FP64 tables (xpos[nx], ypos[ny], sincostbl[2na])
 * into caller-owned buffers: calculate_on_cpu on the host, calculate_on_gpu
 * into device buffers.  Operator classes take device pointers in
 * calculate_on_gpu, enqueue on the stream and synchronise it before
 * returning (the reference's contract: results ready on return).  The
 * operators' calculate_on_cpu takes host pointers, as the reference's does,
 * and computes on the GPU: it stages the buffers through device memory on a
 * stream of its own and runs the same kernels (same results as
 * calculate_on_gpu).  Failures return false with ctr_last_error() set.
 */
#ifndef CT_RECON_MI355X_BP_FAN_2D_H_
#define CT_RECON_MI355X_BP_FAN_2D_H_

#include <hip/hip_runtime_api.h>

namespace ct_recon {

struct FanBackprojection2DParam {
    unsigned int ns;     // number of channels
    unsigned int na;     // number of views
    double ds;           // channel pitch: arc length on the detector at radius dsd
    double orbit;        // source-angle step between views (rad)
    double offset_s;     // detector centre minus (ns-1)/2, in channels
    double orbit_start;  // source angle of view 0 (rad)
    unsigned int nx, ny;  // image size
    double dx, dy;        // pixel pitch
    double offset_x, offset_y;  // iso-centre minus (n-1)/2, in pixels
    double dso;          // source to rotation centre
    double dsd;          // source to detector
    double fov;          // unused by the pixel-driven operators

    FanBackprojection2DParam() {}
    FanBackprojection2DParam(unsigned int ns, unsigned int na, double ds, double orbit,
                             double offset_s, double orbit_start, unsigned int nx, unsigned int ny,
                             double dx, double dy, double offset_x, double offset_y, double dso,
                             double dsd, double fov)
        : ns(ns), na(na), ds(ds), orbit(orbit), offset_s(offset_s), orbit_start(orbit_start),
          nx(nx), ny(ny), dx(dx), dy(dy), offset_x(offset_x), offset_y(offset_y), dso(dso),
          dsd(dsd), fov(fov)
    {
    }
};

class FanBackprojection2DPrepare {
public:
    explicit FanBackprojection2DPrepare(const FanBackprojection2DParam& param) : param_(param) {}
    virtual ~FanBackprojection2DPrepare() {}
    virtual bool calculate_on_cpu(double* buffer1, double* buffer2, double* buffer3) const = 0;
    virtual bool calculate_on_gpu(double* buffer1, double* buffer2, double* buffer3,
                                  hipStream_t) const = 0;

protected:
    FanBackprojection2DParam param_;
};

template <typename T>
class FanBackprojection2D {
public:
    explicit FanBackprojection2D(const FanBackprojection2DParam& param) : param_(param) {}
    virtual ~FanBackprojection2D() {}
    virtual bool calculate_on_cpu(const T* proj, T* img, const double*, const double*,
                                  const double*) const = 0;
    virtual bool calculate_on_gpu(const T* proj, T* img, const double*, const double*,
                                  const double*, hipStream_t) const = 0;

protected:
    FanBackprojection2DParam param_;
};

class FanBackprojection2DGradPrep {
public:
    explicit FanBackprojection2DGradPrep(const FanBackprojection2DParam& param) : param_(param) {}
    virtual ~FanBackprojection2DGradPrep() {}
    virtual bool calculate_on_cpu(double* buffer1, double* buffer2, double* buffer3) const = 0;
    virtual bool calculate_on_gpu(double* buffer1, double* buffer2, double* buffer3,
                                  hipStream_t) const = 0;

protected:
    FanBackprojection2DParam param_;
};

template <typename T>
class FanBackprojection2DGrad {
public:
    explicit FanBackprojection2DGrad(const FanBackprojection2DParam& param) : param_(param) {}
    virtual ~FanBackprojection2DGrad() {}
    virtual bool calculate_on_cpu(const T* img, T* grad, const double*, const double*,
                                  const double*) const = 0;
    virtual bool calculate_on_gpu(const T* img, T* grad, const double*, const double*,
                                  const double*, hipStream_t) const = 0;

protected:
    FanBackprojection2DParam param_;
};

// Tables: xpos[nx], ypos[ny], sincostbl[2na] (src/bp_fan_2d.cc:18-44).
class FanBackprojection2DPixDrivenPrep : public FanBackprojection2DPrepare {
public:
    explicit FanBackprojection2DPixDrivenPrep(const FanBackprojection2DParam& param)
        : FanBackprojection2DPrepare(param)
    {
    }
    bool calculate_on_cpu(double* xpos, double* ypos, double* sincostbl) const override;
    bool calculate_on_gpu(double* xpos, double* ypos, double* sincostbl,
                          hipStream_t) const override;
};

// proj [na][ns] -> img [ny][nx] (src/bp_fan_2d.cc:46-91).
template <typename T>
class FanBackprojection2DPixDriven : public FanBackprojection2D<T> {
public:
    explicit FanBackprojection2DPixDriven(const FanBackprojection2DParam& param)
        : FanBackprojection2D<T>(param)
    {
    }
    bool calculate_on_cpu(const T* proj, T* img, const double* xpos, const double* ypos,
                          const double* sincostbl) const override;
    bool calculate_on_gpu(const T* proj, T* img, const double* xpos, const double* ypos,
                          const double* sincostbl, hipStream_t) const override;
};

class FanBackprojection2DPixDrivenGradPrep : public FanBackprojection2DGradPrep {
public:
    explicit FanBackprojection2DPixDrivenGradPrep(const FanBackprojection2DParam& param)
        : FanBackprojection2DGradPrep(param)
    {
    }
    bool calculate_on_cpu(double* xpos, double* ypos, double* sincostbl) const override;
    bool calculate_on_gpu(double* xpos, double* ypos, double* sincostbl,
                          hipStream_t) const override;
};

// img [ny][nx] -> grad [na][ns], the exact transpose (src/bp_fan_2d.cc:124-170).
template <typename T>
class FanBackprojection2DPixDrivenGrad : public FanBackprojection2DGrad<T> {
public:
    explicit FanBackprojection2DPixDrivenGrad(const FanBackprojection2DParam& param)
        : FanBackprojection2DGrad<T>(param)
    {
    }
    bool calculate_on_cpu(const T* img, T* grad, const double* xpos, const double* ypos,
                          const double* sincostbl) const override;
    bool calculate_on_gpu(const T* img, T* grad, const double* xpos, const double* ypos,
                          const double* sincostbl, hipStream_t) const override;
};

}  // namespace ct_recon

#endif
