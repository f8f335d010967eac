/*
 * ctr_api.h -- C ABI of the MI355X (gfx950) 2-D CT reconstruction library.
 *
 * Plain pointers and sizes only.  Every operator entry point enqueues HIP
 * work on the caller's stream and returns without synchronising; every
 * function returns CTR_OK (0) or an error code, never throws, and leaves a
 * message for ctr_last_error() on the calling thread.  Device pointers are
 * HIP device (or managed) memory; host pointers are ordinary memory.
 *
 * Layouts follow the reference (ltldlzltl/CT_RECON_EXTENSIONS):
 *   sinogram  proj[b][ia][is]   (view-major, channel fastest;
 *                                src/bp_fan_2d.cc:76-77)
 *   image     img[b][iy][ix]    (row 0 is +y; src/bp_fan_2d.cc:22-42)
 * with b the slice of a batch of nbatch slices sharing one geometry, i.e.
 * the NHWC [N,H,W,1] tensors of the reference's TF ops
 * (tensorflow/bp_fan_2d_ops.cu:178-205).
 *
 * Each operator replaces one reference routine; the citation is on its
 * declaration.  The reference's own FFI for this path is the numpy C ABI in
 * include/numpy_ctext.h (re-exported unchanged) and the class API in
 * include/bp_fan_2d.h and siblings (re-declared with hipStream_t); both are
 * thin layers over the functions below.
 */
#ifndef CTR_API_H_
#define CTR_API_H_

#ifdef __cplusplus
extern "C" {
#endif

enum ctr_status {
    CTR_OK = 0,
    CTR_ERR_ARG = 1,          /* null pointer, bad enum, non-positive size  */
    CTR_ERR_SHAPE = 2,        /* geometry the kernels cannot handle         */
    CTR_ERR_HIP = 3,          /* a HIP runtime call failed                   */
    CTR_ERR_UNSUPPORTED = 4,  /* a mode the library does not provide         */
    CTR_ERR_HANDLE = 5,       /* unknown or destroyed handle                 */
    CTR_ERR_ALLOC = 6         /* device allocation failed                    */
};

enum ctr_dtype { CTR_F32 = 0, CTR_F64 = 1 };

/* One geometry record for every operator: the union of the reference's
 * Param structs (include/bp_fan_2d.h:20-59, include/bp_fan_2d_angle.h:20-55,
 * include/bp_par_2d.h:20-55, include/fp_par_2d.h:20-55,
 * include/filter.h:21-36, include/fan_weighting.h:21-39).  Fields a given
 * operator does not use are ignored.  type: ramp filter 0 par / 1 fan
 * (equiangular) / 2 flat; fan weighting 1 fan / 2 flat.  nrow: rows of the
 * ramp filter and fan weighting (views per slice). */
typedef struct ctr_geom {
    unsigned int ns, na;
    double ds, orbit, offset_s, orbit_start;
    unsigned int nx, ny;
    double dx, dy, offset_x, offset_y;
    double dso, dsd, fov;
    int type;
    unsigned int nrow;
} ctr_geom;

/* Message of the last failing call on this thread ("" if none). */
const char* ctr_last_error(void);
/* Library version as 10000*major + 100*minor + patch. */
int ctr_version(void);
/* Device properties the kernels are tuned on (for reports). */
int ctr_device_info(int* cu_count, int* clock_khz, int* lds_bytes_per_cu, char* arch,
                    int arch_len);

/* ---------------------------------------------------------------------
 * FP64 geometry tables, built on the host in the reference's layout and
 * accumulation order (these are the only FP64 step of the path).
 * ------------------------------------------------------------------- */
/* FanBackprojection2DPixDrivenPrep (src/bp_fan_2d.cc:18-44): xpos[nx],
 * ypos[ny], sincostbl[2na] = (sin, cos) of the accumulated angle. */
int ctr_fan_tables(const ctr_geom* g, double* xpos, double* ypos, double* sincostbl);
/* FanBackprojection2DAnglePixDrivenPrep (src/bp_fan_2d_angle.cc:18-45). */
int ctr_fan_angle_tables(const ctr_geom* g, const double* angles, double* xpos, double* ypos,
                         double* sincostbl);
/* ParallelBackprojection2DPixDrivenPrep (src/bp_par_2d.cc:18-49):
 * xcos[ix*na + ia], ysin[iy*na + ia]. */
int ctr_par_tables(const ctr_geom* g, double* xcos, double* ysin);
/* ParallelProjection2DRayDrivenPrepare (src/fp_par_2d.cc:149-208):
 * sincostbl[2na], beginoffset[na][ns][2], usex[na]. */
int ctr_fp_tables(const ctr_geom* g, double* sincostbl, double* beginoffset, int* usex);
/* ParallelProjection2DRayCastingPrepare (src/fp_par_2d.cc:18-89), the
 * reference's default step 0.125: sincostbl[na][2], begins[na][ns][2] (entry
 * point on the FOV circle of radius fov, pixel-index units), nsteps[na][ns]
 * (0 for rays missing the circle). */
int ctr_fp_rc_tables(const ctr_geom* g, double step, double* sincostbl, double* begins,
                     int* nsteps);
/* ParallelProjection2DDisDrivenPrep (src/fp_par_2d.cc:396-457):
 * sincostbl[2na], beginoffset[na][ns][3] = (footprint centre on the first
 * row or column, its shift per row or column, half its width; pixel units),
 * usex[na]. */
int ctr_fp_dd_tables(const ctr_geom* g, double* sincostbl, double* beginoffset, int* usex);
/* ParallelProjection2DRayDrivenGradPrep (src/fp_par_2d.cc:285-330), and
 * ParallelProjection2DDisDrivenGradPrep (src/fp_par_2d.cc:548-599), which
 * computes the same tables: weights[na], pos[na][pos_stride][2], usex[na].
 * The reference uses pos_stride == nx (and so requires nx == ny);
 * pos_stride >= max(nx, ny) lifts that restriction. */
int ctr_fp_grad_tables(const ctr_geom* g, double* weights, double* pos, int* usex,
                       unsigned int pos_stride);
/* RampFilterPrep (src/filter.cc:18-65): filter[2ns+1], tap n at [ns+n]. */
int ctr_ramp_taps(const ctr_geom* g, double* filter);

/* ---------------------------------------------------------------------
 * Operators on device memory (asynchronous on `stream`, a hipStream_t;
 * NULL is the null stream).  dtype selects float32 or float64 buffers;
 * tables are the device copies of the FP64 tables above.
 * ------------------------------------------------------------------- */
/* Pixel-driven fan-beam BP, FanBackprojection2DPixDriven<T>
 * (src/bp_fan_2d.cc:46-91) with scale = |orbit|/round(|na orbit|/pi), and
 * FanBackprojection2DAnglePixDriven<T> (src/bp_fan_2d_angle.cc:47-91) with
 * scale = 1.  proj [nbatch][na][ns] -> img [nbatch][ny][nx]. */
int ctr_fan_bp(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
               const double* xpos, const double* ypos, const double* sincostbl, double scale,
               void* stream);
/* Launch shape of the float32 fan BP kernel, passed per call (no state is
 * kept between calls, so concurrent calls with different shapes do not
 * interact): tile 0 = 16x16, 1 = 8x8, 2 = 32x8 pixels; vchunk = views per
 * LDS round (1..64, capped to the LDS budget; 0 = the most that keep four
 * workgroups per CU); nb = slices per workgroup (1, 2, 4, 8, 16; 0 = from
 * the batch).  All zero is ctr_fan_bp's choice.  Tiles whose detector span
 * exceeds 256 channels switch to 8x8 by themselves.  No counterpart in the
 * reference, whose launch shape is fixed (cuda/bp_fan_2d.cu:219-245). */
typedef struct ctr_fan_bp_launch {
    int tile;
    int vchunk;
    int nb;
} ctr_fan_bp_launch;
/* ctr_fan_bp with an explicit launch shape (NULL = all zero).  The shape is
 * validated before any device work; float64 buffers ignore it. */
int ctr_fan_bp_ex(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
                  const double* xpos, const double* ypos, const double* sincostbl, double scale,
                  const ctr_fan_bp_launch* launch, void* stream);
/* Exact transpose of ctr_fan_bp: img [nbatch][ny][nx] -> proj
 * [nbatch][na][ns] (replaces FanBackprojection2DPixDrivenGrad<T>,
 * src/bp_fan_2d.cc:124-170, and the broken bp_fan_2d_angle transpose).
 * Both types gather, per (view, channel), the pixels that reach it and sum
 * them in registers in a fixed order (no atomics), so the result is
 * bitwise repeatable in either mode; `deterministic` is kept for the
 * reference's signature and changes nothing. */
int ctr_fan_bp_t(const ctr_geom* g, int dtype, const void* img, void* proj, int nbatch,
                 const double* xpos, const double* ypos, const double* sincostbl, double scale,
                 int deterministic, void* stream);
/* Pixel-driven parallel BP, ParallelBackprojection2DPixDriven<T>
 * (src/bp_par_2d.cc:51-90): img = orbit * sum_views lerp(proj, s). */
int ctr_par_bp(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
               const double* xcos, const double* ysin, void* stream);
/* Single-view parallel BP, ParallelSingleViewBp2DPixDriven<T>
 * (src/bp_par_2d_sv.cc:54-92; TF op SingleViewBpParallel2D,
 * tensorflow/bp_par_2d_sv_ops.cu:46-85): proj [nbatch][na][ns] ->
 * img [nbatch][ny][nx][na], each (pixel, view) the view's linear
 * interpolation at s (zero off the detector), no sum and no scale; the
 * tables are ctr_par_tables'.  Bit-identical to the reference's CPU path.
 * The reference's TF launcher steps the output by nx*ny per slice; here it
 * is nx*ny*na, so slices do not overlap. */
int ctr_par_sv_bp(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
                  const double* xcos, const double* ysin, void* stream);
/* Exact transpose of ctr_par_bp (replaces the reference's broken
 * ParallelBackprojection2DPixDrivenGrad<T>, src/bp_par_2d.cc:147-199);
 * deterministic and float64 as for ctr_fan_bp_t. */
int ctr_par_bp_t(const ctr_geom* g, int dtype, const void* img, void* proj, int nbatch,
                 const double* xcos, const double* ysin, int deterministic, void* stream);
/* Ray-driven parallel FP, ParallelProjection2DRayDriven<T>
 * (src/fp_par_2d.cc:210-280): img [nbatch][ny][nx] -> proj [nbatch][na][ns]. */
int ctr_fp(const ctr_geom* g, int dtype, const void* img, void* proj, int nbatch,
           const double* sincostbl, const double* beginoffset, const int* usex, void* stream);
/* Ray-casting parallel FP, ParallelProjection2DRayCasting<T>
 * (src/fp_par_2d.cc:91-144; TF FP op with method "raycasting",
 * tensorflow/fp_par_2d_ops.cu:152-161): img [nbatch][ny][nx] -> proj
 * [nbatch][na][ns] from the tables of ctr_fp_rc_tables; bit-identical to
 * the reference's CPU path.  The reference has no transpose for it. */
int ctr_fp_rc(const ctr_geom* g, int dtype, const void* img, void* proj, int nbatch,
              double step, const double* sincostbl, const double* begins, const int* nsteps,
              void* stream);
/* Gather-form transpose of ctr_fp, ParallelProjection2DRayDrivenGrad<T>
 * (src/fp_par_2d.cc:332-391), channel range clamped to the detector. */
int ctr_fp_t(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
             const double* weights, const double* pos, const int* usex, unsigned int pos_stride,
             void* stream);
/* Distance-driven parallel FP, ParallelProjection2DDisDriven<T>
 * (src/fp_par_2d.cc:459-543; TF FP op with method "disdriven",
 * tensorflow/fp_par_2d_ops.cu:175-186): img [nbatch][ny][nx] -> proj
 * [nbatch][na][ns] from the tables of ctr_fp_dd_tables; bit-identical to the
 * reference's CPU path, including its unsigned-index rule (a row whose
 * footprint starts left of the first pixel adds nothing). */
int ctr_fp_dd(const ctr_geom* g, int dtype, const void* img, void* proj, int nbatch,
              const double* sincostbl, const double* beginoffset, const int* usex,
              void* stream);
/* The gradient the reference defines for it, ParallelProjection2DDisDrivenGrad<T>
 * (src/fp_par_2d.cc:601-655; the TF op's gradient for "disdriven",
 * tensorflow/fp_par_2d_ops.cu:304-315): proj [nbatch][na][ns] -> img
 * [nbatch][ny][nx] from the tables of ctr_fp_grad_tables.  Bit-identical to
 * the reference where it reads inside the detector (channels past ns - 1,
 * which the reference reads beyond the row, are not read).  Not an exact
 * transpose of ctr_fp_dd: the pair is the reference's. */
int ctr_fp_dd_grad(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
                   const double* weights, const double* pos, const int* usex,
                   unsigned int pos_stride, void* stream);
/* ctr_fp_t without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory during the call (code
 * under HIP graph capture, serving loops, harnesses that forbid it).  Same
 * arguments and results as ctr_fp_t; the only memory the transpose needs
 * beyond its buffers (the border pass's FP64 partial sums, float32 on the
 * usual geometries) lives in a static 16 MiB device array per GPU, held
 * until the process ends, so nothing is allocated or freed.  A call needing
 * more (na/16 x nbatch x (2nx + 2ny - 4) doubles) returns CTR_ERR_SHAPE.
 * Calls that overlap on one GPU share that array and race: issue them on
 * one stream, or order them. */
int ctr_fp_t_noalloc(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
                     const double* weights, const double* pos, const int* usex,
                     unsigned int pos_stride, void* stream);
/* Ramp filter, RampFilter<T> / RampFilterGrad<T> (src/filter.cc:67-124):
 * out[r][is] = ds * sum_k in[r][k] h[is-k] over nrow_total rows of ns.
 * filter: device copy of the 2ns+1 FP64 taps.  Optional fused fan weighting
 * of the input (weight_type 1 fan / 2 flat, with g->dso, g->dsd,
 * g->offset_s; 0 = none), FanWeighting<T> (src/fan_weighting.cc:19-35). */
int ctr_ramp(const ctr_geom* g, int dtype, const void* in, void* out, unsigned int nrow_total,
             const double* filter, int weight_type, void* stream);
/* The same ramp filter (float32) as a Toeplitz GEMM on the matrix cores:
 * the even taps other than h0 vanish, so each output parity is a product of
 * the opposite-parity input columns with a Toeplitz matrix of the odd taps.
 * bf16x3 = 0: FP32 MFMA (v_mfma_f32_16x16x4_f32), FP32 accumulation;
 * bf16x3 = 1: BF16 MFMA on hi/lo split operands, hi*hi + hi*lo + lo*hi
 * accumulated in FP32.  Same arguments and fused weighting as ctr_ramp. */
int ctr_ramp_toeplitz(const ctr_geom* g, int dtype, const void* in, void* out,
                      unsigned int nrow_total, const double* filter, int weight_type, int bf16x3,
                      void* stream);
/* Fan / flat cosine weighting and its transpose (the same diagonal),
 * FanWeighting<T> / FanWeightingGrad<T> (src/fan_weighting.cc:19-59). */
int ctr_fan_weighting(const ctr_geom* g, int dtype, const void* in, void* out,
                      unsigned int nrow_total, void* stream);

/* ---------------------------------------------------------------------
 * The reference's execution structure on this GPU, kept as a measured
 * baseline: one launch per slice, one thread per output, FP64 tables read
 * from global memory, FP64 arithmetic and accumulation, the stream
 * synchronised after every launch (tensorflow/bp_fan_2d_ops.h:107-128,
 * cuda/bp_fan_2d.cu:72-110,219-245), with a full grid.
 * ------------------------------------------------------------------- */
int ctr_plain_fan_bp(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
                     const double* xpos, const double* ypos, const double* sincostbl,
                     double scale, void* stream);
/* The same with the reference's launch geometry (GetCudaLaunchConfig and
 * CudaGridRangeX, cuda/cuda_common.h:30-91): threads = min(max_threads,
 * npix), blocks = min(ceil(npix / threads), max_blocks), each thread walking
 * pixels with a grid stride.  The reference uses max_blocks = 32 and
 * max_threads = 512; max_blocks = 0 means a full grid (ctr_plain_fan_bp). */
int ctr_plain_fan_bp_grid(const ctr_geom* g, int dtype, const void* proj, void* img,
                          int nbatch, const double* xpos, const double* ypos,
                          const double* sincostbl, double scale, int max_blocks,
                          int max_threads, void* stream);

/* ---------------------------------------------------------------------
 * Ragged batches of FanBackprojection2DAnglePixDriven<float>
 * (src/bp_fan_2d_angle.cc:47-91): problems with different image sizes,
 * view counts, detectors and source distances in one persistent launch
 * (replaces the reference's serial per-slice loop over one geometry,
 * tensorflow/bp_fan_2d_ops.h:107-128).  A plan fixes the problems (their
 * geometry, device tables and float32 buffers), the pixel tile (0 16x16,
 * 1 8x8, 2 32x8) and the views per LDS round (1..64); running it clears a
 * work counter with a kernel and launches the persistent kernel, either
 * directly or (use_graph != 0) by replaying a HIP graph captured on the
 * first such call.  img = scale * sum over views, scale = 1 for the
 * reference's bp_fan_2d_angle.  Runs of one plan share its work counter,
 * so the library orders them: each run waits (on its stream) for the end of
 * the plan's previous run, whichever stream or host thread issued it. 
 * ------------------------------------------------------------------- */
typedef struct ctr_ragged_problem {
    ctr_geom g;                 /* ns, na, ds, offset_s, nx, ny, dx, dy, offsets, dso, dsd */
    const double* xpos;         /* device tables of ctr_fan_angle_tables */
    const double* ypos;
    const double* sincostbl;
    const float* proj;          /* device [na][ns] */
    float* img;                 /* device [ny][nx] */
    double scale;
} ctr_ragged_problem;

typedef struct ctr_ragged_plan ctr_ragged_plan;

int ctr_ragged_fan_bp_plan(const ctr_ragged_problem* probs, int nprob, int tile, int vchunk,
                           ctr_ragged_plan** plan);
int ctr_ragged_fan_bp_run(ctr_ragged_plan* plan, int use_graph, void* stream);
/* Work items (pixel tiles), persistent grid, block, LDS bytes and the
 * pixel*view count of a plan. */
int ctr_ragged_fan_bp_info(const ctr_ragged_plan* plan, int* nitems, int* grid, int* block,
                           int* lds_bytes, double* pixel_views);
/* One run that also records every workgroup's first and last clock tick;
 * busy_fraction = sum of workgroup busy time / (grid x span).  Synchronises
 * the stream. */
int ctr_ragged_fan_bp_balance(ctr_ragged_plan* plan, void* stream, double* busy_fraction);
int ctr_ragged_fan_bp_destroy(ctr_ragged_plan* plan);
/* Time every tile {16x16, 8x8, 32x8} x view chunk {8, 16, 32, 64} over
 * `reps` runs (after one warm-up) and return the fastest; ms_table
 * (optional, 12 floats, tile-major) receives each mean, -1 where the
 * geometry does not admit the configuration.  Synchronises the stream. */
int ctr_ragged_fan_bp_autotune(const ctr_ragged_problem* probs, int nprob, int reps,
                               void* stream, int* best_tile, int* best_vchunk, float* ms_table);

/* ---------------------------------------------------------------------
 * Helical rebinning: a multi-row helical scan proj_in[b][na][nt][ns]
 * (view, detector row, channel) to nz fan-beam sinograms
 * proj_out[b][nz][view_per_rot][ns], one rotation per slice.
 * ------------------------------------------------------------------- */
/* HelicalInterpolationParam (include/helical_interpolation.h:17-63) field
 * for field.  dt and orbit_start are carried but unused, as in the
 * reference. */
typedef struct ctr_helical_geom {
    unsigned int ns, nt, na;
    double ds, dt, offset_s, offset_t, orbit_start, orbit, couch_begin, couch_mov;
    unsigned int view_per_rot, nz;
    double dz, z0, thickness, dso, dsd;
} ctr_helical_geom;

enum ctr_helical_mode { CTR_HELICAL_360 = 360, CTR_HELICAL_180 = 180 };

/* Host tables of ctr_helical, 3 ns + 4 nz doubles: per channel cos(gamma),
 * then (pi - 2 gamma) / orbit, then cs - s/ds (the 180LI conjugate's view
 * offset and mirrored channel), then per slice (z, centre view, first view
 * of its window, first output view), with z accumulated from z0 in steps of
 * dz as the reference does (src/helical_interpolation.cc:336-365).  A slice whose window of
 * +-1 rotation (360) or of one rotation about its centre view (180) leaves
 * the scan returns CTR_ERR_SHAPE: the reference's "View not enough for this
 * slice!" (:23-24, :165-166). */
int ctr_helical_tables(const ctr_helical_geom* g, int mode, double* tables);
/* HelicalInterpolation360<T> / HelicalInterpolation180<T>::calculate_on_cpu
 * (include/helical_interpolation.h:86-118; src/helical_interpolation.cc:
 * 81-160 and 162-269): bit-identical per element in float32 and float64.
 * tables: device copy of ctr_helical_tables' output.  status (device, may
 * be NULL) is OR-ed with 1 where a sample has no detector row in either
 * window (the reference's "No row data available!"; NaN written there) and
 * with 2 where a 180LI conjugate view would interpolate from a view outside
 * the scan (the reference reads outside its input; NaN written); clear it
 * before the call. */
int ctr_helical(const ctr_helical_geom* g, int mode, int dtype, const void* proj_in,
                void* proj_out, int nbatch, const double* tables, unsigned int* status,
                void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CTR_API_H_ */
