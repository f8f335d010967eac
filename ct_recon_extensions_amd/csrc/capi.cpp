// extern "C" operator entry points: argument checks, dtype dispatch, and the
// hand-off to the HIP launchers.  Nothing here throws.
#include <cmath>
#include <exception>

#include "ctr_common.hpp"

namespace ctr {
int fan_bp_f32(const ctr_geom& g, const float* proj, float* img, int nbatch, const double* xpos,
               const double* ypos, const double* sc, double scale, ctr_fan_bp_launch cfg,
               hipStream_t st);
int fan_bpt_f32(const ctr_geom& g, const float* img, float* proj, int nbatch, const double* xpos,
                const double* ypos, const double* sc, double scale, bool det, void* ws,
                hipStream_t st);
int fan_bp_f64(const ctr_geom& g, const double* proj, double* img, int nbatch,
               const double* xpos, const double* ypos, const double* sc, double scale,
               hipStream_t st);
int fan_bpt_f64(const ctr_geom& g, const double* img, double* proj, int nbatch,
                const double* xpos, const double* ypos, const double* sc, double scale,
                hipStream_t st);
int plain_fan_bp(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
                 const double* xpos, const double* ypos, const double* sc, double scale,
                 int max_blocks, int max_threads, hipStream_t st);
int par_bp(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
           const double* xcos, const double* ysin, hipStream_t st);
int par_bpt(const ctr_geom& g, int dtype, const void* img, void* proj, int nbatch,
            const double* xcos, const double* ysin, bool det, void* ws, hipStream_t st);
int par_svbp(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
             const double* xcos, const double* ysin, hipStream_t st);
int fp_rc(const ctr_geom& g, int dtype, const void* img, void* proj, int nbatch, double step,
          const double* sincostbl, const double* begins, const int* nsteps, hipStream_t st);
int fp_par(const ctr_geom& g, int dtype, const void* img, void* proj, int nbatch,
           const double* sc, const double* bo, const int* usex, hipStream_t st);
int fp_par_t(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
             const double* w, const double* pos, const int* usex, unsigned int pos_stride,
             hipStream_t st);
int fp_dd(const ctr_geom& g, int dtype, const void* img, void* proj, int nbatch,
          const double* sc, const double* bo, const int* usex, hipStream_t st);
int helical(const ctr_helical_geom& g, int mode, int dtype, const void* in, void* out,
            int nbatch, const double* tables, unsigned int* status, hipStream_t st);
int fp_dd_grad(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
               const double* w, const double* pos, const int* usex, unsigned int pos_stride,
               hipStream_t st);
int fp_par_t_noalloc(const ctr_geom& g, int dtype, const void* proj, void* img, int nbatch,
                     const double* w, const double* pos, const int* usex,
                     unsigned int pos_stride, hipStream_t st);
int ramp(const ctr_geom& g, int dtype, const void* in, void* out, unsigned int nrow,
         const double* filter, int weight_type, hipStream_t st);
int fan_weighting(const ctr_geom& g, int dtype, const void* in, void* out, unsigned int nrow,
                  hipStream_t st);
}  // namespace ctr

namespace {

int check_common(const ctr_geom* g, int dtype, const void* a, const void* b, int nbatch,
                 const char* who)
{
    if (!g || !a || !b) {
        ctr::set_error(std::string(who) + ": null pointer argument");
        return CTR_ERR_ARG;
    }
    if (dtype != CTR_F32 && dtype != CTR_F64) {
        ctr::set_error(std::string(who) + ": dtype must be CTR_F32 or CTR_F64");
        return CTR_ERR_ARG;
    }
    if (nbatch < 1) {
        ctr::set_error(std::string(who) + ": nbatch must be >= 1");
        return CTR_ERR_ARG;
    }
    return CTR_OK;
}

template <typename F>
int guarded(const char* who, F&& f)
{
    try {
        return f();
    } catch (const std::exception& e) {
        ctr::set_error(std::string(who) + ": " + e.what());
        return CTR_ERR_ALLOC;
    } catch (...) {
        ctr::set_error(std::string(who) + ": unknown C++ exception");
        return CTR_ERR_ALLOC;
    }
}

}  // namespace

extern "C" {

int ctr_fan_bp_ex(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
                  const double* xpos, const double* ypos, const double* sincostbl, double scale,
                  const ctr_fan_bp_launch* launch, void* stream)
{
    int rc = check_common(g, dtype, proj, img, nbatch, "ctr_fan_bp");
    if (rc) return rc;
    CTR_REQUIRE(xpos && ypos && sincostbl, CTR_ERR_ARG, "ctr_fan_bp: null table");
    const ctr_fan_bp_launch cfg = launch ? *launch : ctr_fan_bp_launch{0, 0, 0};
    CTR_REQUIRE(cfg.tile >= 0 && cfg.tile <= 2, CTR_ERR_ARG,
                "ctr_fan_bp_ex: launch tile must be 0, 1 or 2");
    CTR_REQUIRE(cfg.vchunk >= 0 && cfg.vchunk <= 64, CTR_ERR_ARG,
                "ctr_fan_bp_ex: launch vchunk must be in [0, 64]");
    CTR_REQUIRE(cfg.nb == 0 || cfg.nb == 1 || cfg.nb == 2 || cfg.nb == 4 || cfg.nb == 8 ||
                    cfg.nb == 16,
                CTR_ERR_ARG, "ctr_fan_bp_ex: launch nb must be 0, 1, 2, 4, 8 or 16");
    return guarded("ctr_fan_bp", [&] {
        auto st = static_cast<hipStream_t>(stream);
        if (dtype == CTR_F32)
            return ctr::fan_bp_f32(*g, static_cast<const float*>(proj), static_cast<float*>(img),
                                   nbatch, xpos, ypos, sincostbl, scale, cfg, st);
        return ctr::fan_bp_f64(*g, static_cast<const double*>(proj), static_cast<double*>(img),
                               nbatch, xpos, ypos, sincostbl, scale, st);
    });
}

int ctr_fan_bp(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
               const double* xpos, const double* ypos, const double* sincostbl, double scale,
               void* stream)
{
    return ctr_fan_bp_ex(g, dtype, proj, img, nbatch, xpos, ypos, sincostbl, scale, nullptr,
                         stream);
}

int ctr_fan_bp_t(const ctr_geom* g, int dtype, const void* img, void* proj, int nbatch,
                 const double* xpos, const double* ypos, const double* sincostbl, double scale,
                 int deterministic, void* stream)
{
    int rc = check_common(g, dtype, img, proj, nbatch, "ctr_fan_bp_t");
    if (rc) return rc;
    CTR_REQUIRE(xpos && ypos && sincostbl, CTR_ERR_ARG, "ctr_fan_bp_t: null table");
    return guarded("ctr_fan_bp_t", [&] {
        auto st = static_cast<hipStream_t>(stream);
        if (dtype == CTR_F32)
            return ctr::fan_bpt_f32(*g, static_cast<const float*>(img), static_cast<float*>(proj),
                                    nbatch, xpos, ypos, sincostbl, scale, deterministic != 0,
                                    nullptr, st);
        return ctr::fan_bpt_f64(*g, static_cast<const double*>(img), static_cast<double*>(proj),
                                nbatch, xpos, ypos, sincostbl, scale, st);
    });
}

int ctr_plain_fan_bp_grid(const ctr_geom* g, int dtype, const void* proj, void* img,
                          int nbatch, const double* xpos, const double* ypos,
                          const double* sincostbl, double scale, int max_blocks,
                          int max_threads, void* stream)
{
    int rc = check_common(g, dtype, proj, img, nbatch, "ctr_plain_fan_bp");
    if (rc) return rc;
    CTR_REQUIRE(xpos && ypos && sincostbl, CTR_ERR_ARG, "ctr_plain_fan_bp: null table");
    CTR_REQUIRE(max_blocks >= 0, CTR_ERR_ARG, "ctr_plain_fan_bp_grid: max_blocks must be >= 0");
    CTR_REQUIRE(max_blocks == 0 || (max_threads >= 1 && max_threads <= 1024), CTR_ERR_ARG,
                "ctr_plain_fan_bp_grid: max_threads must be in [1, 1024]");
    return guarded("ctr_plain_fan_bp", [&] {
        return ctr::plain_fan_bp(*g, dtype, proj, img, nbatch, xpos, ypos, sincostbl, scale,
                                 max_blocks, max_threads, static_cast<hipStream_t>(stream));
    });
}

int ctr_plain_fan_bp(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
                     const double* xpos, const double* ypos, const double* sincostbl,
                     double scale, void* stream)
{
    return ctr_plain_fan_bp_grid(g, dtype, proj, img, nbatch, xpos, ypos, sincostbl, scale, 0,
                                 0, stream);
}

int ctr_par_bp(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
               const double* xcos, const double* ysin, void* stream)
{
    int rc = check_common(g, dtype, proj, img, nbatch, "ctr_par_bp");
    if (rc) return rc;
    CTR_REQUIRE(xcos && ysin, CTR_ERR_ARG, "ctr_par_bp: null table");
    return guarded("ctr_par_bp", [&] {
        return ctr::par_bp(*g, dtype, proj, img, nbatch, xcos, ysin,
                           static_cast<hipStream_t>(stream));
    });
}

int ctr_par_sv_bp(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
                  const double* xcos, const double* ysin, void* stream)
{
    int rc = check_common(g, dtype, proj, img, nbatch, "ctr_par_sv_bp");
    if (rc) return rc;
    CTR_REQUIRE(xcos && ysin, CTR_ERR_ARG, "ctr_par_sv_bp: null table");
    return guarded("ctr_par_sv_bp", [&] {
        return ctr::par_svbp(*g, dtype, proj, img, nbatch, xcos, ysin,
                             static_cast<hipStream_t>(stream));
    });
}

int ctr_par_bp_t(const ctr_geom* g, int dtype, const void* img, void* proj, int nbatch,
                 const double* xcos, const double* ysin, int deterministic, void* stream)
{
    int rc = check_common(g, dtype, img, proj, nbatch, "ctr_par_bp_t");
    if (rc) return rc;
    CTR_REQUIRE(xcos && ysin, CTR_ERR_ARG, "ctr_par_bp_t: null table");
    return guarded("ctr_par_bp_t", [&] {
        return ctr::par_bpt(*g, dtype, img, proj, nbatch, xcos, ysin, deterministic != 0,
                            nullptr, static_cast<hipStream_t>(stream));
    });
}

int ctr_fp(const ctr_geom* g, int dtype, const void* img, void* proj, int nbatch,
           const double* sincostbl, const double* beginoffset, const int* usex, void* stream)
{
    int rc = check_common(g, dtype, img, proj, nbatch, "ctr_fp");
    if (rc) return rc;
    CTR_REQUIRE(sincostbl && beginoffset && usex, CTR_ERR_ARG, "ctr_fp: null table");
    return guarded("ctr_fp", [&] {
        return ctr::fp_par(*g, dtype, img, proj, nbatch, sincostbl, beginoffset, usex,
                           static_cast<hipStream_t>(stream));
    });
}

int ctr_fp_rc(const ctr_geom* g, int dtype, const void* img, void* proj, int nbatch,
              double step, const double* sincostbl, const double* begins, const int* nsteps,
              void* stream)
{
    int rc = check_common(g, dtype, img, proj, nbatch, "ctr_fp_rc");
    if (rc) return rc;
    CTR_REQUIRE(sincostbl && begins && nsteps, CTR_ERR_ARG, "ctr_fp_rc: null table");
    return guarded("ctr_fp_rc", [&] {
        return ctr::fp_rc(*g, dtype, img, proj, nbatch, step, sincostbl, begins, nsteps,
                          static_cast<hipStream_t>(stream));
    });
}

int ctr_fp_t(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
             const double* weights, const double* pos, const int* usex, unsigned int pos_stride,
             void* stream)
{
    int rc = check_common(g, dtype, proj, img, nbatch, "ctr_fp_t");
    if (rc) return rc;
    CTR_REQUIRE(weights && pos && usex, CTR_ERR_ARG, "ctr_fp_t: null table");
    const unsigned int need = g->nx > g->ny ? g->nx : g->ny;
    CTR_REQUIRE(pos_stride >= need, CTR_ERR_ARG, "ctr_fp_t: pos_stride must be >= max(nx, ny)");
    return guarded("ctr_fp_t", [&] {
        return ctr::fp_par_t(*g, dtype, proj, img, nbatch, weights, pos, usex, pos_stride,
                             static_cast<hipStream_t>(stream));
    });
}

int ctr_fp_dd(const ctr_geom* g, int dtype, const void* img, void* proj, int nbatch,
              const double* sincostbl, const double* beginoffset, const int* usex, void* stream)
{
    int rc = check_common(g, dtype, img, proj, nbatch, "ctr_fp_dd");
    if (rc) return rc;
    CTR_REQUIRE(sincostbl && beginoffset && usex, CTR_ERR_ARG, "ctr_fp_dd: null table");
    return guarded("ctr_fp_dd", [&] {
        return ctr::fp_dd(*g, dtype, img, proj, nbatch, sincostbl, beginoffset, usex,
                          static_cast<hipStream_t>(stream));
    });
}

int ctr_fp_dd_grad(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
                   const double* weights, const double* pos, const int* usex,
                   unsigned int pos_stride, void* stream)
{
    int rc = check_common(g, dtype, proj, img, nbatch, "ctr_fp_dd_grad");
    if (rc) return rc;
    CTR_REQUIRE(weights && pos && usex, CTR_ERR_ARG, "ctr_fp_dd_grad: null table");
    const unsigned int need = g->nx > g->ny ? g->nx : g->ny;
    CTR_REQUIRE(pos_stride >= need, CTR_ERR_ARG,
                "ctr_fp_dd_grad: pos_stride must be >= max(nx, ny)");
    return guarded("ctr_fp_dd_grad", [&] {
        return ctr::fp_dd_grad(*g, dtype, proj, img, nbatch, weights, pos, usex, pos_stride,
                               static_cast<hipStream_t>(stream));
    });
}

int ctr_helical(const ctr_helical_geom* g, int mode, int dtype, const void* proj_in,
                void* proj_out, int nbatch, const double* tables, unsigned int* status,
                void* stream)
{
    CTR_REQUIRE(g && proj_in && proj_out && tables, CTR_ERR_ARG, "ctr_helical: null argument");
    CTR_REQUIRE(mode == CTR_HELICAL_360 || mode == CTR_HELICAL_180, CTR_ERR_ARG,
                "ctr_helical: mode must be 360 or 180");
    CTR_REQUIRE(dtype == CTR_F32 || dtype == CTR_F64, CTR_ERR_ARG, "ctr_helical: bad dtype");
    CTR_REQUIRE(nbatch >= 1, CTR_ERR_ARG, "ctr_helical: nbatch must be >= 1");
    CTR_REQUIRE(g->ns >= 1 && g->nt >= 1 && g->na >= 1 && g->view_per_rot >= 1 && g->nz >= 1,
                CTR_ERR_ARG, "ctr_helical: ns, nt, na, view_per_rot and nz must be >= 1");
    CTR_REQUIRE(g->ns <= 0x7fffffffu && g->nt <= 0x7fffffffu && g->na <= 0x7fffffffu &&
                    g->view_per_rot <= 0x7fffffffu && g->nz <= 0x7fffffffu,
                CTR_ERR_SHAPE, "ctr_helical: sizes must fit in int");
    auto usable = [](double v) { return std::isfinite(v) && v != 0.0; };
    CTR_REQUIRE(usable(g->ds) && usable(g->dso) && usable(g->dsd) &&
                    (mode == CTR_HELICAL_360 || usable(g->orbit)),
                CTR_ERR_ARG, "ctr_helical: ds, dso, dsd (and orbit for 180LI) must be finite "
                             "and nonzero");
    return guarded("ctr_helical", [&] {
        return ctr::helical(*g, mode, dtype, proj_in, proj_out, nbatch, tables, status,
                            static_cast<hipStream_t>(stream));
    });
}

int ctr_fp_t_noalloc(const ctr_geom* g, int dtype, const void* proj, void* img, int nbatch,
                     const double* weights, const double* pos, const int* usex,
                     unsigned int pos_stride, void* stream)
{
    int rc = check_common(g, dtype, proj, img, nbatch, "ctr_fp_t_noalloc");
    if (rc) return rc;
    CTR_REQUIRE(weights && pos && usex, CTR_ERR_ARG, "ctr_fp_t_noalloc: null table");
    const unsigned int need = g->nx > g->ny ? g->nx : g->ny;
    CTR_REQUIRE(pos_stride >= need, CTR_ERR_ARG,
                "ctr_fp_t_noalloc: pos_stride must be >= max(nx, ny)");
    return guarded("ctr_fp_t_noalloc", [&] {
        return ctr::fp_par_t_noalloc(*g, dtype, proj, img, nbatch, weights, pos, usex,
                                     pos_stride, static_cast<hipStream_t>(stream));
    });
}

int ctr_ramp(const ctr_geom* g, int dtype, const void* in, void* out, unsigned int nrow_total,
             const double* filter, int weight_type, void* stream)
{
    int rc = check_common(g, dtype, in, out, 1, "ctr_ramp");
    if (rc) return rc;
    CTR_REQUIRE(filter, CTR_ERR_ARG, "ctr_ramp: null filter taps");
    CTR_REQUIRE(weight_type >= 0 && weight_type <= 2, CTR_ERR_ARG,
                "ctr_ramp: weight_type must be 0 (none), 1 (fan) or 2 (flat)");
    CTR_REQUIRE(g->ns >= 1, CTR_ERR_ARG, "ctr_ramp: ns must be >= 1");
    if (nrow_total == 0) return CTR_OK;
    return guarded("ctr_ramp", [&] {
        return ctr::ramp(*g, dtype, in, out, nrow_total, filter, weight_type,
                         static_cast<hipStream_t>(stream));
    });
}

int ctr_fan_weighting(const ctr_geom* g, int dtype, const void* in, void* out,
                      unsigned int nrow_total, void* stream)
{
    int rc = check_common(g, dtype, in, out, 1, "ctr_fan_weighting");
    if (rc) return rc;
    CTR_REQUIRE(g->type == 1 || g->type == 2, CTR_ERR_ARG,
                "ctr_fan_weighting: type must be 1 (fan) or 2 (flat)");
    if (nrow_total == 0) return CTR_OK;
    return guarded("ctr_fan_weighting", [&] {
        return ctr::fan_weighting(*g, dtype, in, out, nrow_total, static_cast<hipStream_t>(stream));
    });
}

}  // extern "C"
