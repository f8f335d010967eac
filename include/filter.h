/*
 * filter.h -- ramp filter and its transpose: the class API of the
 * reference's include/filter.h:21-92 with hipStream_t GPU entry points.
 * Conventions as in bp_fan_2d.h.
 */
#ifndef CT_RECON_MI355X_FILTER_H_
#define CT_RECON_MI355X_FILTER_H_

#include <hip/hip_runtime_api.h>

namespace ct_recon {

struct FilterParam {
    unsigned int ns;    // channels
    unsigned int nrow;  // rows (views) per call
    double ds;          // channel pitch
    double dsd;         // source to detector (fan type only)
    int type;           // 0 par, 1 fan (equiangular), 2 flat
    int window;         // not implemented, as in the reference

    FilterParam() {}
    FilterParam(unsigned int ns, unsigned int nrow, double ds, double dsd = -1, int type = 0,
                int window = 0)
        : ns(ns), nrow(nrow), ds(ds), dsd(dsd), type(type), window(window)
    {
    }
};

// Taps filter[2ns+1], tap n at [ns+n] (src/filter.cc:18-65).
class RampFilterPrep {
public:
    explicit RampFilterPrep(const FilterParam& param) : param_(param) {}
    bool calculate_on_cpu(double* filter) const;
    bool calculate_on_gpu(double* filter, hipStream_t) const;

private:
    FilterParam param_;
};

// out[r][is] = ds * sum_k in[r][k] filter[ns + is - k] (src/filter.cc:67-91).
template <typename T>
class RampFilter {
public:
    explicit RampFilter(const FilterParam& param) : param_(param) {}
    bool calculate_on_cpu(const T* in, const double* filter, T* out) const;
    bool calculate_on_gpu(const T* in, const double* filter, T* out, hipStream_t) const;

private:
    FilterParam param_;
};

// The transpose: correlation, equal to the filter since the taps are
// symmetric (src/filter.cc:96-124).
template <typename T>
class RampFilterGrad {
public:
    explicit RampFilterGrad(const FilterParam& param) : param_(param) {}
    bool calculate_on_cpu(const T* in, const double* filter, T* out) const;
    bool calculate_on_gpu(const T* in, const double* filter, T* out, hipStream_t) const;

private:
    FilterParam param_;
};

}  // namespace ct_recon

#endif
