/*
 * bp_fan_2d_angle.h -- fan-beam pixel-driven backprojection at caller-given
 * source angles (no angular scaling) and its exact transpose: the class API
 * of the reference's include/bp_fan_2d_angle.h:20-223 with hipStream_t GPU
 * entry points.  Conventions as in bp_fan_2d.h.
 */
#ifndef CT_RECON_MI355X_BP_FAN_2D_ANGLE_H_
#define CT_RECON_MI355X_BP_FAN_2D_ANGLE_H_

#include <hip/hip_runtime_api.h>

namespace ct_recon {

struct FanBackprojection2DAngleParam {
    unsigned int ns, na;
    double ds, offset_s;
    unsigned int nx, ny;
    double dx, dy, offset_x, offset_y;
    double dso, dsd, fov;

    FanBackprojection2DAngleParam() {}
    FanBackprojection2DAngleParam(unsigned int ns, unsigned int na, double ds, double offset_s,
                                  unsigned int nx, unsigned int ny, double dx, double dy,
                                  double offset_x, double offset_y, double dso, double dsd,
                                  double fov)
        : ns(ns), na(na), ds(ds), offset_s(offset_s), nx(nx), ny(ny), dx(dx), dy(dy),
          offset_x(offset_x), offset_y(offset_y), dso(dso), dsd(dsd), fov(fov)
    {
    }
};

class FanBackprojection2DAnglePrepare {
public:
    explicit FanBackprojection2DAnglePrepare(const FanBackprojection2DAngleParam& param)
        : param_(param)
    {
    }
    virtual ~FanBackprojection2DAnglePrepare() {}
    virtual bool calculate_on_cpu(const double* angles, double* buffer1, double* buffer2,
                                  double* buffer3) const = 0;
    virtual bool calculate_on_gpu(const double* angles, double* buffer1, double* buffer2,
                                  double* buffer3, hipStream_t) const = 0;

protected:
    FanBackprojection2DAngleParam param_;
};

template <typename T>
class FanBackprojection2DAngle {
public:
    explicit FanBackprojection2DAngle(const FanBackprojection2DAngleParam& param) : param_(param) {}
    virtual ~FanBackprojection2DAngle() {}
    virtual bool calculate_on_cpu(const T* proj, T* img, const double*, const double*,
                                  const double*) const = 0;
    virtual bool calculate_on_gpu(const T* proj, T* img, const double*, const double*,
                                  const double*, hipStream_t) const = 0;

protected:
    FanBackprojection2DAngleParam param_;
};

class FanBackprojection2DAngleGradPrep {
public:
    explicit FanBackprojection2DAngleGradPrep(const FanBackprojection2DAngleParam& param)
        : param_(param)
    {
    }
    virtual ~FanBackprojection2DAngleGradPrep() {}
    virtual bool calculate_on_cpu(const double* angles, double* buffer1, double* buffer2,
                                  double* buffer3) const = 0;
    virtual bool calculate_on_gpu(const double* angles, double* buffer1, double* buffer2,
                                  double* buffer3, hipStream_t) const = 0;

protected:
    FanBackprojection2DAngleParam param_;
};

template <typename T>
class FanBackprojection2DAngleGrad {
public:
    explicit FanBackprojection2DAngleGrad(const FanBackprojection2DAngleParam& param)
        : param_(param)
    {
    }
    virtual ~FanBackprojection2DAngleGrad() {}
    virtual bool calculate_on_cpu(const T* img, T* grad, const double*, const double*,
                                  const double*) const = 0;
    virtual bool calculate_on_gpu(const T* img, T* grad, const double*, const double*,
                                  const double*, hipStream_t) const = 0;

protected:
    FanBackprojection2DAngleParam param_;
};

// angles (host pointer for calculate_on_cpu, device pointer for
// calculate_on_gpu) -> xpos[nx], ypos[ny], sincostbl[2na]
// (src/bp_fan_2d_angle.cc:18-45).
class FanBackprojection2DAnglePixDrivenPrep : public FanBackprojection2DAnglePrepare {
public:
    explicit FanBackprojection2DAnglePixDrivenPrep(const FanBackprojection2DAngleParam& param)
        : FanBackprojection2DAnglePrepare(param)
    {
    }
    bool calculate_on_cpu(const double* angles, double* xpos, double* ypos,
                          double* sincostbl) const override;
    bool calculate_on_gpu(const double* angles, double* xpos, double* ypos, double* sincostbl,
                          hipStream_t) const override;
};

template <typename T>
class FanBackprojection2DAnglePixDriven : public FanBackprojection2DAngle<T> {
public:
    explicit FanBackprojection2DAnglePixDriven(const FanBackprojection2DAngleParam& param)
        : FanBackprojection2DAngle<T>(param)
    {
    }
    bool calculate_on_cpu(const T* proj, T* img, const double* xpos, const double* ypos,
                          const double* sincostbl) const override;
    bool calculate_on_gpu(const T* proj, T* img, const double* xpos, const double* ypos,
                          const double* sincostbl, hipStream_t) const override;
};

class FanBackprojection2DAnglePixDrivenGradPrep : public FanBackprojection2DAngleGradPrep {
public:
    explicit FanBackprojection2DAnglePixDrivenGradPrep(const FanBackprojection2DAngleParam& param)
        : FanBackprojection2DAngleGradPrep(param)
    {
    }
    bool calculate_on_cpu(const double* angles, double* xpos, double* ypos,
                          double* sincostbl) const override;
    bool calculate_on_gpu(const double* angles, double* xpos, double* ypos, double* sincostbl,
                          hipStream_t) const override;
};

// Exact transpose of FanBackprojection2DAnglePixDriven (the reference's own,
// src/bp_fan_2d_angle.cc:124-165, divides by the weights instead).
template <typename T>
class FanBackprojection2DAnglePixDrivenGrad : public FanBackprojection2DAngleGrad<T> {
public:
    explicit FanBackprojection2DAnglePixDrivenGrad(const FanBackprojection2DAngleParam& param)
        : FanBackprojection2DAngleGrad<T>(param)
    {
    }
    bool calculate_on_cpu(const T* img, T* grad, const double* xpos, const double* ypos,
                          const double* sincostbl) const override;
    bool calculate_on_gpu(const T* img, T* grad, const double* xpos, const double* ypos,
                          const double* sincostbl, hipStream_t) const override;
};

}  // namespace ct_recon

#endif
