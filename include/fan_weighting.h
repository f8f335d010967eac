/*
 * fan_weighting.h -- fan / flat cosine pre-weighting and its transpose: the
 * class API of the reference's include/fan_weighting.h:21-77 with hipStream_t
 * GPU entry points.  Conventions as in bp_fan_2d.h.
 */
#ifndef CT_RECON_MI355X_FAN_WEIGHTING_H_
#define CT_RECON_MI355X_FAN_WEIGHTING_H_

#include <hip/hip_runtime_api.h>

namespace ct_recon {

struct FanWeightingParam {
    unsigned int ns, nrow;
    double ds;
    double offset;  // detector centre minus (ns-1)/2, in channels
    double dso, dsd;
    int type;       // 1 fan (equiangular), 2 flat

    FanWeightingParam() {}
    FanWeightingParam(unsigned int ns, unsigned int nrow, double ds, double offset = 0,
                      double dso = -1, double dsd = -1, int type = 1)
        : ns(ns), nrow(nrow), ds(ds), offset(offset), dso(dso), dsd(dsd), type(type)
    {
    }
};

// out[r][is] = dso |cos(gamma_is)| / dsd * in[r][is] (src/fan_weighting.cc:19-35).
template <typename T>
class FanWeighting {
public:
    explicit FanWeighting(const FanWeightingParam& param) : param_(param) {}
    bool calculate_on_cpu(const T* in, T* out) const;
    bool calculate_on_gpu(const T* in, T* out, hipStream_t) const;

private:
    FanWeightingParam param_;
};

// The same diagonal (src/fan_weighting.cc:39-56).
template <typename T>
class FanWeightingGrad {
public:
    explicit FanWeightingGrad(const FanWeightingParam& param) : param_(param) {}
    bool calculate_on_cpu(const T* in, T* out) const;
    bool calculate_on_gpu(const T* in, T* out, hipStream_t) const;

private:
    FanWeightingParam param_;
};

}  // namespace ct_recon

#endif
