/*
 * helical_interpolation.h -- helical rebinning (360LI and 180LI): the class
 * API of the reference's include/helical_interpolation.h:17-118 with a
 * hipStream_t GPU entry point.  proj_in is the multi-row helical scan
 * [na][nt][ns] (view, detector row, channel), proj_out the nz fan-beam
 * sinograms [nz][view_per_rot][ns]; both are device pointers for
 * calculate_on_gpu, which synchronises the stream before it returns, and
 * host pointers for calculate_on_cpu.
 *
 * Errors follow the reference: a slice whose views the scan does not cover
 * throws std::runtime_error("View not enough for this slice!") before any
 * work, and a sample with no detector row in either window throws
 * std::runtime_error("No row data available!") (after the launch: the other
 * samples are written, that one is NaN).  A HIP failure returns false with
 * the message in ctr_last_error().  calculate_on_cpu takes host pointers, as
 * the reference's does, and computes on the GPU (buffers staged through
 * device memory), with the same results and errors.
 */
#ifndef CT_RECON_MI355X_HELICAL_INTERPOLATION_H_
#define CT_RECON_MI355X_HELICAL_INTERPOLATION_H_

#include <hip/hip_runtime_api.h>

namespace ct_recon {

struct HelicalInterpolationParam {
    unsigned int ns;            // channels
    unsigned int nt;            // detector rows
    unsigned int na;            // views
    double ds;                  // channel spacing, mm
    double dt;                  // row spacing, mm (unused, as in the reference)
    double offset_s;            // centre channel minus (ns-1)/2
    double offset_t;            // centre row minus (nt-1)/2
    double orbit_start;         // source angle at view 0 (unused)
    double orbit;               // source rotation per view, rad
    double couch_begin;         // source z at view 0
    double couch_mov;           // source z movement per view
    unsigned int view_per_rot;  // views per rotation
    unsigned int nz;            // slices
    double dz;                  // slice spacing
    double z0;                  // position of slice 0
    double thickness;           // slice thickness
    double dso, dsd;            // source-isocentre and source-detector distances

    HelicalInterpolationParam() {}
    HelicalInterpolationParam(unsigned int ns, unsigned int nt, unsigned int na, double ds,
                              double dt, double offset_s, double offset_t, double orbit_start,
                              double orbit, double couch_begin, double couch_mov,
                              unsigned int view_per_rot, unsigned int nz, double dz, double z0,
                              double thickness, double dso, double dsd)
        : ns(ns), nt(nt), na(na), ds(ds), dt(dt), offset_s(offset_s), offset_t(offset_t),
          orbit_start(orbit_start), orbit(orbit), couch_begin(couch_begin),
          couch_mov(couch_mov), view_per_rot(view_per_rot), nz(nz), dz(dz), z0(z0),
          thickness(thickness), dso(dso), dsd(dsd)
    {
    }
};

template <typename T>
class HelicalInterpolation {
public:
    explicit HelicalInterpolation(const HelicalInterpolationParam& param) : param_(param) {}
    virtual ~HelicalInterpolation() {}
    virtual bool calculate_on_cpu(const T* proj_in, T* proj_out) const = 0;
    virtual bool calculate_on_gpu(const T* proj_in, T* proj_out, hipStream_t) const = 0;

protected:
    HelicalInterpolationParam param_;
};

// 360LI: each output view from the same angle one rotation before and after
// the slice (src/helical_interpolation.cc:81-160).
template <typename T>
class HelicalInterpolation360 : public HelicalInterpolation<T> {
public:
    explicit HelicalInterpolation360(const HelicalInterpolationParam& param)
        : HelicalInterpolation<T>(param)
    {
    }
    bool calculate_on_cpu(const T* proj_in, T* proj_out) const override;
    bool calculate_on_gpu(const T* proj_in, T* proj_out, hipStream_t) const override;
};

// 180LI: each output view from its own ray and the opposite (conjugate) ray
// within half a rotation of the slice (src/helical_interpolation.cc:162-269).
template <typename T>
class HelicalInterpolation180 : public HelicalInterpolation<T> {
public:
    explicit HelicalInterpolation180(const HelicalInterpolationParam& param)
        : HelicalInterpolation<T>(param)
    {
    }
    bool calculate_on_cpu(const T* proj_in, T* proj_out) const override;
    bool calculate_on_gpu(const T* proj_in, T* proj_out, hipStream_t) const override;
};

}  // namespace ct_recon

#endif
