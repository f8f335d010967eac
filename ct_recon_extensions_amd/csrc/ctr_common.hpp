// Shared definitions of the MI355X CT library: the geometry record, error
// plumbing for the C ABI, and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "include/ctr_api.h"

namespace ctr {

// Last-error string of the calling thread (the C ABI never throws).
void set_error(const std::string& msg);
const char* get_error();

#define CTR_HIP(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ::ctr::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));   \
            return CTR_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

#define CTR_REQUIRE(cond, code, msg)                                               \
    do {                                                                           \
        if (!(cond)) {                                                             \
            ::ctr::set_error(msg);                                                 \
            return (code);                                                         \
        }                                                                          \
    } while (0)

// Launches report asynchronous configuration errors through this check.
inline int check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return CTR_ERR_HIP;
    }
    return CTR_OK;
}

inline int div_up(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

// Centres of the pixel grid and of the detector, as every reference routine
// defines them (e.g. src/bp_fan_2d.cc:22-25, :56-57).
inline double cent_x(const ctr_geom& g) { return static_cast<double>(g.nx - 1) / 2 + g.offset_x; }
inline double cent_y(const ctr_geom& g) { return static_cast<double>(g.ny - 1) / 2 + g.offset_y; }
inline double cent_s(const ctr_geom& g) { return static_cast<double>(g.ns - 1) / 2 + g.offset_s; }

// Scale of bp_fan_2d: |orbit| / round(|na*orbit|/pi) (src/bp_fan_2d.cc:58,83).
double fan_bp_scale(const ctr_geom& g);

}  // namespace ctr
