/*
 * numpy_ctext.h -- the reference's numpy C ABI, served by the MI355X library.
 *
 * Same eleven symbols and signatures as the reference's libnumpy_ctext.so
 * (numpy_ext/common.cu:29-61, numpy_ext/fan_weighting_npext.cu:86-106,
 * numpy_ext/filter_npext.cu:104-124, numpy_ext/bp_fan_2d_angle_npext.cu:136-159),
 * so its ctypes front-end (numpy_ext/np_ctrecon.py:14-61) binds unchanged.
 *
 * Arrays are host float64, C-contiguous; every *_run copies its input to the
 * device, runs the float64 kernels on a stream the library owns, copies the
 * result back and returns when it is ready (in may equal out).
 *
 * Differences from the reference, all on its error paths:
 *   - set_device(d < 0), the reference's CPU mode, computes on GPU 0 (the
 *     arrays are host memory in both modes);
 *   - nothing throws across the C boundary: failures return false / 0;
 *   - fan_bp_2d_angle_create copies `angles` (the reference keeps the
 *     caller's pointer until the first run and leaks its device copy);
 *   - handles are guarded by a mutex.
 */
#ifndef NUMPY_CTEXT_H_
#define NUMPY_CTEXT_H_

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* numpy_ext/common.cu:29-49: select the GPU (and create the stream). */
bool set_device(const int device);
/* numpy_ext/common.cu:51-61: release the stream. */
bool clear();

/* numpy_ext/fan_weighting_npext.cu:86-106 */
int fan_weighting_create(unsigned int ns, unsigned int nrow, double ds, double offset, double dso,
                         double dsd, int type);
bool fan_weighting_run(int handle, double* in, double* out);
bool fan_weighting_destroy(int handle);

/* numpy_ext/filter_npext.cu:104-124 */
int ramp_filter_create(unsigned int ns, unsigned int nrow, double ds, double dsd, int type);
bool ramp_filter_run(int handle, double* in, double* out);
bool ramp_filter_destroy(int handle);

/* numpy_ext/bp_fan_2d_angle_npext.cu:136-159 */
int fan_bp_2d_angle_create(double* angles, unsigned int ns, unsigned int na, double ds,
                           double offset_s, unsigned int nx, unsigned int ny, double dx,
                           double dy, double offset_x, double offset_y, double dso, double dsd,
                           double fov);
bool fan_bp_2d_angle_run(int handle, double* in, double* out);
bool fan_bp_2d_angle_destroy(int handle);

#ifdef __cplusplus
}
#endif

#endif /* NUMPY_CTEXT_H_ */
