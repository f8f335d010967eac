// FP64 geometry tables, built on the host in the reference's layouts and in
// its accumulation order (angles advanced by repeated addition, pixel
// positions likewise), so the device kernels see exactly the values the
// reference's CPU path uses.  This is the only FP64 precomputation of the
// path; the kernels round what they need to FP32.
#include <cmath>
#include <cstring>
#include <string>

#include "ctr_common.hpp"

namespace ctr {

static thread_local std::string t_last_error;

void set_error(const std::string& msg) { t_last_error = msg; }
const char* get_error() { return t_last_error.c_str(); }

}  // namespace ctr

using ctr::cent_s;
using ctr::cent_x;
using ctr::cent_y;

// sin and cos of one angle through glibc's sincos.  GCC at -O2/-O3 (the
// reference's build, CMakeLists.txt:22) turns a sin and a cos of the same
// argument into one sincos call, whose cosine can differ from cos() in the
// last bit (view 464 of 720 over 180 deg is one such angle); calling it
// explicitly keeps these tables equal to the reference's whichever compiler
// builds this file.
static inline void sin_cos(double x, double& s, double& c) { ::sincos(x, &s, &c); }

extern "C" {

const char* ctr_last_error(void) { return ctr::get_error(); }

int ctr_version(void) { return 10000 * 0 + 100 * 3 + 0; }

// Pixel centres x_ix = -centx*dx + ix*dx and y_iy = centy*dy - iy*dy, by
// accumulation (src/bp_fan_2d.cc:27-42).
static void pixel_positions(const ctr_geom& g, double* xpos, double* ypos)
{
    double px = -cent_x(g) * g.dx, py = cent_y(g) * g.dy;
    for (unsigned int i = 0; i < g.nx; ++i) {
        xpos[i] = px;
        px += g.dx;
    }
    for (unsigned int i = 0; i < g.ny; ++i) {
        ypos[i] = py;
        py -= g.dy;
    }
}

int ctr_fan_tables(const ctr_geom* g, double* xpos, double* ypos, double* sincostbl)
{
    CTR_REQUIRE(g && xpos && ypos && sincostbl, CTR_ERR_ARG, "ctr_fan_tables: null argument");
    double angle = g->orbit_start;
    for (unsigned int a = 0; a < g->na; ++a) {
        sin_cos(angle, sincostbl[2 * a], sincostbl[2 * a + 1]);
        angle += g->orbit;
    }
    pixel_positions(*g, xpos, ypos);
    return CTR_OK;
}

int ctr_fan_angle_tables(const ctr_geom* g, const double* angles, double* xpos, double* ypos,
                         double* sincostbl)
{
    CTR_REQUIRE(g && angles && xpos && ypos && sincostbl, CTR_ERR_ARG,
                "ctr_fan_angle_tables: null argument");
    for (unsigned int a = 0; a < g->na; ++a) {
        sin_cos(angles[a], sincostbl[2 * a], sincostbl[2 * a + 1]);
    }
    pixel_positions(*g, xpos, ypos);
    return CTR_OK;
}

// xcos[ix*na + ia] = x*cos(angle), ysin[iy*na + ia] = y*sin(angle)
// (src/bp_par_2d.cc:18-49).
int ctr_par_tables(const ctr_geom* g, double* xcos, double* ysin)
{
    CTR_REQUIRE(g && xcos && ysin, CTR_ERR_ARG, "ctr_par_tables: null argument");
    double angle = g->orbit_start;
    for (unsigned int a = 0; a < g->na; ++a) {
        double sa, ca;
        sin_cos(angle, sa, ca);
        double px = -cent_x(*g) * g->dx;
        for (unsigned int i = 0; i < g->nx; ++i) {
            xcos[a + static_cast<size_t>(i) * g->na] = px * ca;
            px += g->dx;
        }
        double py = cent_y(*g) * g->dy;
        for (unsigned int i = 0; i < g->ny; ++i) {
            ysin[a + static_cast<size_t>(i) * g->na] = py * sa;
            py -= g->dy;
        }
        angle += g->orbit;
    }
    return CTR_OK;
}

// Ray-casting FP tables (src/fp_par_2d.cc:18-89): per view (sin, cos) of
// the accumulated angle; per ray the entry point on the FOV circle (radius
// fov) in pixel-index units and the number of steps of length `step` along
// the marched axis (x when sin > 1/1.414); rays missing the circle: 0 steps.
int ctr_fp_rc_tables(const ctr_geom* g, double step, double* sincostbl, double* begins,
                     int* nsteps)
{
    CTR_REQUIRE(g && sincostbl && begins && nsteps, CTR_ERR_ARG,
                "ctr_fp_rc_tables: null argument");
    CTR_REQUIRE(step > 0, CTR_ERR_ARG, "ctr_fp_rc_tables: step must be positive");
    const double xc = static_cast<double>(g->nx - 1) / 2.0 + g->offset_x;
    const double yc = static_cast<double>(g->ny - 1) / 2.0 + g->offset_y;
    const double fov2 = g->fov * g->fov;
    const double s0 = -(static_cast<double>(g->ns - 1) / 2.0 + g->offset_s) * g->ds;
    double angle = g->orbit_start;
    for (unsigned int a = 0; a < g->na; ++a) {
        double sa, ca;
        sin_cos(angle, sa, ca);
        sincostbl[2 * a] = sa;
        sincostbl[2 * a + 1] = ca;
        const double stx = step * sa, sty = step * ca;
        const bool usex = sa > (1.0 / 1.414);
        double sv = s0;
        for (unsigned int i = 0; i < g->ns; ++i) {
            const size_t r = static_cast<size_t>(a) * g->ns + i;
            const double h2 = fov2 - sv * sv;
            if (h2 > 0) {
                const double h = std::sqrt(h2);
                const double x1 = (sv * ca - h * sa) / g->dx + xc;
                const double y1 = -(sv * sa + h * ca) / g->dy + yc;
                const double x2 = (sv * ca + h * sa) / g->dx + xc;
                const double y2 = -(sv * sa - h * ca) / g->dy + yc;
                nsteps[r] = usex ? static_cast<int>((x2 - x1) / stx)
                                 : static_cast<int>((y2 - y1) / sty);
                begins[2 * r] = x1;
                begins[2 * r + 1] = y1;
            } else {
                nsteps[r] = 0;
                begins[2 * r] = 0;
                begins[2 * r + 1] = 0;
            }
            sv += g->ds;
        }
        angle += g->orbit;
    }
    return CTR_OK;
}

// Ray-driven FP tables (src/fp_par_2d.cc:149-208): the angle is wrapped into
// [0, 2pi) before use, the x-branch is used outside [pi/4,3pi/4) and
// [5pi/4,7pi/4), and each channel's start position advances by repeated
// addition.
int ctr_fp_tables(const ctr_geom* g, double* sincostbl, double* beginoffset, int* usex)
{
    CTR_REQUIRE(g && sincostbl && beginoffset && usex, CTR_ERR_ARG, "ctr_fp_tables: null argument");
    const double cx = cent_x(*g), cy = cent_y(*g), cs = cent_s(*g);
    const double q = M_PI / 4;
    double angle = g->orbit_start;
    double* bo = beginoffset;
    for (unsigned int a = 0; a < g->na; ++a) {
        while (angle < 0) angle += 2 * M_PI;
        while (angle >= 2 * M_PI) angle -= 2 * M_PI;
        double sa, ca;
        sin_cos(angle, sa, ca);
        sincostbl[2 * a] = sa;
        sincostbl[2 * a + 1] = ca;
        const bool ybranch = (angle >= q && angle < 3 * q) || (angle >= 5 * q && angle < 7 * q);
        usex[a] = ybranch ? 0 : 1;
        double step, chstep, begin;
        if (!ybranch) {
            step = g->dy * sa / (ca * g->dx);
            chstep = g->ds / (ca * g->dx);
            begin = cx - cy * step - cs * chstep;
        } else {
            step = g->dx * ca / (sa * g->dy);
            chstep = -g->ds / (sa * g->dy);
            begin = cy - cx * step - cs * chstep;
        }
        for (unsigned int s = 0; s < g->ns; ++s) {
            bo[0] = begin;
            bo[1] = step;
            begin += chstep;
            bo += 2;
        }
        angle += g->orbit;
    }
    return CTR_OK;
}

// Distance-driven FP tables (src/fp_par_2d.cc:396-457): the ray-driven
// tables' angle wrap, branch rule, per-channel begin (accumulated) and step,
// plus half the channel's footprint width in pixel units:
// beginoffset[na][ns][3] = (begin, step, chstep / 2).
int ctr_fp_dd_tables(const ctr_geom* g, double* sincostbl, double* beginoffset, int* usex)
{
    CTR_REQUIRE(g && sincostbl && beginoffset && usex, CTR_ERR_ARG,
                "ctr_fp_dd_tables: null argument");
    const double cx = cent_x(*g), cy = cent_y(*g), cs = cent_s(*g);
    const double q = M_PI / 4;
    double angle = g->orbit_start;
    double* bo = beginoffset;
    for (unsigned int a = 0; a < g->na; ++a) {
        while (angle < 0) angle += 2 * M_PI;
        while (angle >= 2 * M_PI) angle -= 2 * M_PI;
        double sa, ca;
        sin_cos(angle, sa, ca);
        sincostbl[2 * a] = sa;
        sincostbl[2 * a + 1] = ca;
        const bool ybranch = (angle >= q && angle < 3 * q) || (angle >= 5 * q && angle < 7 * q);
        usex[a] = ybranch ? 0 : 1;
        double step, chstep, begin;
        if (!ybranch) {
            step = g->dy * sa / (ca * g->dx);
            chstep = g->ds / (ca * g->dx);
            begin = cx - cy * step - cs * chstep;
        } else {
            step = g->dx * ca / (sa * g->dy);
            chstep = -g->ds / (sa * g->dy);
            begin = cy - cx * step - cs * chstep;
        }
        for (unsigned int s = 0; s < g->ns; ++s) {
            bo[0] = begin;
            bo[1] = step;
            bo[2] = chstep / 2;
            begin += chstep;
            bo += 3;
        }
        angle += g->orbit;
    }
    return CTR_OK;
}

// FP transpose tables (src/fp_par_2d.cc:285-330); no angle wrap, as in the
// reference.  pos is [na][pos_stride][2].
int ctr_fp_grad_tables(const ctr_geom* g, double* weights, double* pos, int* usex,
                       unsigned int pos_stride)
{
    CTR_REQUIRE(g && weights && pos && usex, CTR_ERR_ARG, "ctr_fp_grad_tables: null argument");
    const unsigned int need = g->nx > g->ny ? g->nx : g->ny;
    CTR_REQUIRE(pos_stride >= need || (pos_stride == g->nx && g->nx == g->ny), CTR_ERR_ARG,
                "ctr_fp_grad_tables: pos_stride must be >= max(nx, ny)");
    const double cx = cent_x(*g), cy = cent_y(*g), cs = cent_s(*g);
    const double q = M_PI / 4;
    double angle = g->orbit_start;
    for (unsigned int a = 0; a < g->na; ++a) {
        double sa, ca;
        sin_cos(angle, sa, ca);
        const bool ybranch = (angle >= q && angle < 3 * q) || (angle >= 5 * q && angle < 7 * q);
        usex[a] = ybranch ? 0 : 1;
        double* p = pos + 2 * static_cast<size_t>(a) * pos_stride;
        double o1, o2, begin;
        unsigned int n;
        if (!ybranch) {
            weights[a] = std::fabs(g->dy / ca);
            o1 = g->dx * ca / g->ds;
            o2 = -g->dy * sa / g->ds;
            begin = cs - cx * o1 - cy * o2;
            n = g->ny;
        } else {
            weights[a] = std::fabs(g->dx / sa);
            o1 = -g->dy * sa / g->ds;
            o2 = g->dx * ca / g->ds;
            begin = cs - cy * o1 - cx * o2;
            n = g->nx;
        }
        for (unsigned int k = 0; k < n; ++k) {
            p[0] = begin;
            p[1] = o1;
            begin += o2;
            p += 2;
        }
        angle += g->orbit;
    }
    return CTR_OK;
}

// Ram-Lak taps filter[ns + n], n in [-ns, ns] (src/filter.cc:18-65).
int ctr_ramp_taps(const ctr_geom* g, double* filter)
{
    CTR_REQUIRE(g && filter, CTR_ERR_ARG, "ctr_ramp_taps: null argument");
    const int ns = static_cast<int>(g->ns);
    double* h = filter + ns;
    if (g->type == 0 || g->type == 2) {
        const double inv_ds2 = 1.0 / (g->ds * g->ds);
        const double inv_pi2 = 1.0 / (M_PI * M_PI);
        for (int n = -ns; n <= ns; ++n)
            h[n] = n == 0 ? 0.25 * inv_ds2
                          : (n % 2 == 0 ? 0.0 : -inv_pi2 * inv_ds2 / static_cast<double>(n * n));
    } else if (g->type == 1) {
        const double c = 1.0 / (M_PI * M_PI * g->dsd * g->dsd);
        for (int n = -ns; n <= ns; ++n) {
            if (n == 0) {
                h[n] = 0.25 / (g->ds * g->ds);
            } else if (n % 2 == 0) {
                h[n] = 0.0;
            } else {
                const double sa = std::sin(static_cast<double>(n) * g->ds / g->dsd);
                h[n] = -c / (sa * sa);
            }
        }
    } else {
        ctr::set_error("ctr_ramp_taps: unknown filter type (0 par, 1 fan, 2 flat)");
        return CTR_ERR_ARG;
    }
    return CTR_OK;
}

int ctr_device_info(int* cu_count, int* clock_khz, int* lds_bytes_per_cu, char* arch, int arch_len)
{
    int dev = 0;
    CTR_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    CTR_HIP(hipGetDeviceProperties(&p, dev));
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (clock_khz) *clock_khz = p.clockRate;
    if (lds_bytes_per_cu) *lds_bytes_per_cu = static_cast<int>(p.maxSharedMemoryPerMultiProcessor);
    if (arch && arch_len > 0) {
        std::strncpy(arch, p.gcnArchName, arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    return CTR_OK;
}

int ctr_helical_tables(const ctr_helical_geom* g, int mode, double* tables)
{
    CTR_REQUIRE(g && tables, CTR_ERR_ARG, "ctr_helical_tables: null argument");
    CTR_REQUIRE(mode == CTR_HELICAL_360 || mode == CTR_HELICAL_180, CTR_ERR_ARG,
                "ctr_helical_tables: mode must be 360 or 180");
    CTR_REQUIRE(g->ns >= 1 && g->nt >= 1 && g->na >= 1 && g->view_per_rot >= 1 && g->nz >= 1,
                CTR_ERR_ARG, "ctr_helical_tables: ns, nt, na, view_per_rot and nz must be >= 1");
    CTR_REQUIRE(g->na <= 0x7fffffffu && g->view_per_rot <= 0x3fffffffu, CTR_ERR_SHAPE,
                "ctr_helical_tables: na and view_per_rot must fit in int");
    // per channel (src/helical_interpolation.cc:99-101, 199-201, 215-216):
    // cos(gamma), the 180LI conjugate's view offset (pi - 2 gamma) / orbit,
    // and its mirrored channel cs - s/ds
    const double center_s = (g->ns - 1) / 2.0 + g->offset_s;
    for (unsigned int is = 0; is < g->ns; ++is) {
        const double s = (static_cast<int>(is) - center_s) * g->ds;
        const double gamma = s / g->dsd;
        tables[is] = cos(gamma);
        tables[g->ns + is] = (M_PI - 2 * gamma) / g->orbit;
        tables[2 * g->ns + is] = center_s - s / g->ds;
    }
    // slices: z accumulated from z0 (:338-346), the centre view and the
    // rotation(s) around it (:21-27, :163-168)
    const int vpr = static_cast<int>(g->view_per_rot);
    double* sl = tables + 3 * static_cast<size_t>(g->ns);
    double z = g->z0;
    for (unsigned int iz = 0; iz < g->nz; ++iz, z += g->dz) {
        const double c = std::round((z - g->couch_begin) / g->couch_mov);
        CTR_REQUIRE(std::isfinite(c) && std::fabs(c) < 1e9, CTR_ERR_SHAPE,
                    "ctr_helical_tables: slice " + std::to_string(iz) +
                        " has no centre view (couch_mov 0 or z not finite)");
        const int center_view = static_cast<int>(c);
        int begin, end, first_view;
        if (mode == CTR_HELICAL_360) {
            begin = center_view - vpr;
            end = center_view + vpr - 1;
        } else {
            begin = center_view - vpr / 2;
            end = begin + vpr - 1;
        }
        CTR_REQUIRE(begin >= 0 && end < static_cast<int>(g->na), CTR_ERR_SHAPE,
                    "ctr_helical_tables: View not enough for this slice! (slice " +
                        std::to_string(iz) + ", views " + std::to_string(begin) + ".." +
                        std::to_string(end) + " of " + std::to_string(g->na) + ")");
        if (mode == CTR_HELICAL_360)
            first_view = (begin / vpr + 1) * vpr;
        else
            first_view = static_cast<int>(std::ceil(begin / static_cast<double>(vpr)) * vpr);
        sl[4 * iz] = z;
        sl[4 * iz + 1] = center_view;
        sl[4 * iz + 2] = begin;
        sl[4 * iz + 3] = first_view;
    }
    return CTR_OK;
}

}  // extern "C"
