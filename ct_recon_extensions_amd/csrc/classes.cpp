// The ct_recon:: C++ class API (include/bp_fan_2d.h and siblings) over the
// C ABI.  Prep classes build the reference's FP64 tables on the host (the
// reference's accumulation order) and, for calculate_on_gpu, copy them into
// the caller's device buffers.  Operator classes run one slice on the given
// stream and synchronise it before returning, like the reference's
// calculate_on_gpu (e.g. cuda/bp_fan_2d.cu:229).  The operators'
// calculate_on_cpu takes host buffers, as the reference's does
// (e.g. src/bp_fan_2d.cc:46-91), and computes on the GPU: the buffers are
// staged through device memory and the same kernels run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>
#include <vector>

#include "ctr_common.hpp"
#include "include/bp_fan_2d.h"
#include "include/bp_fan_2d_angle.h"
#include "include/bp_par_2d.h"
#include "include/bp_par_2d_sv.h"
#include "include/fan_weighting.h"
#include "include/filter.h"
#include "include/fp_par_2d.h"
#include "include/helical_interpolation.h"

#include <stdexcept>
#include <string>

namespace ct_recon {

namespace {

template <typename T>
constexpr int dtype_of()
{
    return std::is_same<T, double>::value ? CTR_F64 : CTR_F32;
}

bool ok_sync(int rc, hipStream_t st)
{
    if (rc != CTR_OK) return false;
    const hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        ctr::set_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
        return false;
    }
    return true;
}

template <typename V>
bool to_device(void* dst, const std::vector<V>& src, hipStream_t st)
{
    const hipError_t e =
        hipMemcpyAsync(dst, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
        ctr::set_error(std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
        return false;
    }
    return true;
}

ctr_geom geom(const FanBackprojection2DParam& p)
{
    ctr_geom g{};
    g.ns = p.ns; g.na = p.na; g.ds = p.ds; g.orbit = p.orbit; g.offset_s = p.offset_s;
    g.orbit_start = p.orbit_start; g.nx = p.nx; g.ny = p.ny; g.dx = p.dx; g.dy = p.dy;
    g.offset_x = p.offset_x; g.offset_y = p.offset_y; g.dso = p.dso; g.dsd = p.dsd;
    g.fov = p.fov;
    return g;
}

ctr_geom geom(const FanBackprojection2DAngleParam& p)
{
    ctr_geom g{};
    g.ns = p.ns; g.na = p.na; g.ds = p.ds; g.offset_s = p.offset_s; g.nx = p.nx; g.ny = p.ny;
    g.dx = p.dx; g.dy = p.dy; g.offset_x = p.offset_x; g.offset_y = p.offset_y; g.dso = p.dso;
    g.dsd = p.dsd; g.fov = p.fov;
    return g;
}

template <class P>
ctr_geom geom_par(const P& p)
{
    ctr_geom g{};
    g.ns = p.ns; g.na = p.na; g.ds = p.ds; g.orbit = p.orbit; g.offset_s = p.offset_s;
    g.orbit_start = p.orbit_start; g.nx = p.nx; g.ny = p.ny; g.dx = p.dx; g.dy = p.dy;
    g.offset_x = p.offset_x; g.offset_y = p.offset_y; g.fov = p.fov;
    return g;
}

ctr_geom geom(const FilterParam& p)
{
    ctr_geom g{};
    g.ns = p.ns; g.nrow = p.nrow; g.ds = p.ds; g.dsd = p.dsd; g.type = p.type;
    return g;
}

ctr_geom geom(const FanWeightingParam& p)
{
    ctr_geom g{};
    g.ns = p.ns; g.nrow = p.nrow; g.ds = p.ds; g.offset_s = p.offset; g.dso = p.dso;
    g.dsd = p.dsd; g.type = p.type;
    return g;
}

bool fan_tables_gpu(const ctr_geom& g, double* xpos, double* ypos, double* sc, hipStream_t st)
{
    std::vector<double> x(g.nx), y(g.ny), s(2 * g.na);
    if (ctr_fan_tables(&g, x.data(), y.data(), s.data()) != CTR_OK) return false;
    return to_device(xpos, x, st) && to_device(ypos, y, st) && to_device(sc, s, st) &&
           ok_sync(CTR_OK, st);
}

bool fan_angle_tables_gpu(const ctr_geom& g, const double* angles_dev, double* xpos,
                          double* ypos, double* sc, hipStream_t st)
{
    std::vector<double> a(g.na), x(g.nx), y(g.ny), s(2 * g.na);
    hipError_t e = hipMemcpyAsync(a.data(), angles_dev, g.na * sizeof(double),
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        ctr::set_error(std::string("reading angles: ") + hipGetErrorString(e));
        return false;
    }
    if (ctr_fan_angle_tables(&g, a.data(), x.data(), y.data(), s.data()) != CTR_OK) return false;
    return to_device(xpos, x, st) && to_device(ypos, y, st) && to_device(sc, s, st) &&
           ok_sync(CTR_OK, st);
}

// Temporary device copy of host vectors, released at scope exit.
struct DevTemp {
    void* p = nullptr;
    hipStream_t st;
    explicit DevTemp(hipStream_t s) : st(s) {}
    ~DevTemp()
    {
        if (p) (void)hipFreeAsync(p, st);
    }
    template <typename V>
    bool put(const std::vector<V>& v)
    {
        if (hipMallocAsync(&p, v.size() * sizeof(V), st) != hipSuccess) {
            ctr::set_error("hipMallocAsync failed");
            return false;
        }
        return to_device(p, v, st);
    }
};

// Host-memory calls (the operators' calculate_on_cpu): the library computes
// only on the GPU, so these stage the caller's host buffers through device
// memory on a stream of their own and run the same kernels as
// calculate_on_gpu, whose result they return bit for bit.  Every buffer,
// outputs included, is copied in first, so the operator sees exactly what a
// device caller would hand it.
class Staging {
public:
    Staging()
    {
        const hipError_t e = hipStreamCreate(&st_);
        if (e != hipSuccess) {
            ctr::set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
            st_ = nullptr;
        }
    }
    ~Staging()
    {
        if (st_) (void)hipStreamSynchronize(st_);
        for (void* p : bufs_) (void)hipFree(p);
        if (st_) (void)hipStreamDestroy(st_);
    }
    bool ok() const { return st_ != nullptr && ok_; }
    hipStream_t stream() const { return st_; }
    // device copy of host[0..n); null in, null out (the operator reports it)
    template <typename V>
    V* in(const V* host, size_t n)
    {
        if (!host || !ok()) return nullptr;
        void* d = alloc(n * sizeof(V));
        if (!d) return nullptr;
        const hipError_t e = hipMemcpyAsync(d, host, n * sizeof(V), hipMemcpyHostToDevice, st_);
        if (e != hipSuccess) return fail("hipMemcpyAsync", e);
        return static_cast<V*>(d);
    }
    // copy the device buffer back into host[0..n) (synchronises)
    template <typename V>
    bool out(V* host, const V* dev, size_t n)
    {
        if (!host || !dev || !ok()) return false;
        hipError_t e = hipMemcpyAsync(host, dev, n * sizeof(V), hipMemcpyDeviceToHost, st_);
        if (e == hipSuccess) e = hipStreamSynchronize(st_);
        if (e != hipSuccess) {
            fail("hipMemcpyAsync", e);
            return false;
        }
        return true;
    }

private:
    void* alloc(size_t bytes)
    {
        void* d = nullptr;
        const hipError_t e = hipMalloc(&d, bytes ? bytes : 1);
        if (e != hipSuccess) return fail("hipMalloc", e);
        bufs_.push_back(d);
        return d;
    }
    std::nullptr_t fail(const char* what, hipError_t e)
    {
        ctr::set_error(std::string(what) + ": " + hipGetErrorString(e));
        ok_ = false;
        return nullptr;
    }
    hipStream_t st_ = nullptr;
    bool ok_ = true;
    std::vector<void*> bufs_;
};

// Runs `gpu` (the class's calculate_on_gpu on staged pointers) and copies the
// output back.
template <typename V, typename F>
bool staged(Staging& s, V* host_out, V* dev_out, size_t n_out, F&& gpu)
{
    return s.ok() && dev_out && gpu() && s.out(host_out, dev_out, n_out);
}

}  // namespace

// ------------------------------------------------------------- bp_fan_2d
bool FanBackprojection2DPixDrivenPrep::calculate_on_cpu(double* xpos, double* ypos,
                                                        double* sc) const
{
    const ctr_geom g = geom(param_);
    return ctr_fan_tables(&g, xpos, ypos, sc) == CTR_OK;
}

bool FanBackprojection2DPixDrivenPrep::calculate_on_gpu(double* xpos, double* ypos, double* sc,
                                                        hipStream_t st) const
{
    return fan_tables_gpu(geom(param_), xpos, ypos, sc, st);
}

bool FanBackprojection2DPixDrivenGradPrep::calculate_on_cpu(double* xpos, double* ypos,
                                                            double* sc) const
{
    const ctr_geom g = geom(param_);
    return ctr_fan_tables(&g, xpos, ypos, sc) == CTR_OK;
}

bool FanBackprojection2DPixDrivenGradPrep::calculate_on_gpu(double* xpos, double* ypos,
                                                            double* sc, hipStream_t st) const
{
    return fan_tables_gpu(geom(param_), xpos, ypos, sc, st);
}

template <typename T>
bool FanBackprojection2DPixDriven<T>::calculate_on_cpu(const T* proj, T* img, const double* xpos,
                                                       const double* ypos,
                                                       const double* sc) const
{
    const auto& p = this->param_;
    const size_t ni = static_cast<size_t>(p.nx) * p.ny;
    Staging s;
    T* d = s.in(img, ni);
    return staged(s, img, d, ni, [&] {
        return calculate_on_gpu(s.in(proj, static_cast<size_t>(p.na) * p.ns), d,
                                s.in(xpos, p.nx), s.in(ypos, p.ny), s.in(sc, 2 * p.na),
                                s.stream());
    });
}

template <typename T>
bool FanBackprojection2DPixDriven<T>::calculate_on_gpu(const T* proj, T* img, const double* xpos,
                                                       const double* ypos, const double* sc,
                                                       hipStream_t st) const
{
    const ctr_geom g = geom(this->param_);
    return ok_sync(ctr_fan_bp(&g, dtype_of<T>(), proj, img, 1, xpos, ypos, sc,
                              ctr::fan_bp_scale(g), st),
                   st);
}

template <typename T>
bool FanBackprojection2DPixDrivenGrad<T>::calculate_on_cpu(const T* img, T* grad,
                                                           const double* xpos, const double* ypos,
                                                           const double* sc) const
{
    const auto& p = this->param_;
    const size_t np = static_cast<size_t>(p.na) * p.ns;
    Staging s;
    T* d = s.in(grad, np);
    return staged(s, grad, d, np, [&] {
        return calculate_on_gpu(s.in(img, static_cast<size_t>(p.nx) * p.ny), d,
                                s.in(xpos, p.nx), s.in(ypos, p.ny), s.in(sc, 2 * p.na),
                                s.stream());
    });
}

template <typename T>
bool FanBackprojection2DPixDrivenGrad<T>::calculate_on_gpu(const T* img, T* grad,
                                                           const double* xpos, const double* ypos,
                                                           const double* sc, hipStream_t st) const
{
    const ctr_geom g = geom(this->param_);
    return ok_sync(ctr_fan_bp_t(&g, dtype_of<T>(), img, grad, 1, xpos, ypos, sc,
                                ctr::fan_bp_scale(g), 0, st),
                   st);
}

template class FanBackprojection2DPixDriven<float>;
template class FanBackprojection2DPixDriven<double>;
template class FanBackprojection2DPixDrivenGrad<float>;
template class FanBackprojection2DPixDrivenGrad<double>;

// ------------------------------------------------------- bp_fan_2d_angle
bool FanBackprojection2DAnglePixDrivenPrep::calculate_on_cpu(const double* angles, double* xpos,
                                                             double* ypos, double* sc) const
{
    const ctr_geom g = geom(param_);
    return ctr_fan_angle_tables(&g, angles, xpos, ypos, sc) == CTR_OK;
}

bool FanBackprojection2DAnglePixDrivenPrep::calculate_on_gpu(const double* angles, double* xpos,
                                                             double* ypos, double* sc,
                                                             hipStream_t st) const
{
    return fan_angle_tables_gpu(geom(param_), angles, xpos, ypos, sc, st);
}

bool FanBackprojection2DAnglePixDrivenGradPrep::calculate_on_cpu(const double* angles,
                                                                 double* xpos, double* ypos,
                                                                 double* sc) const
{
    const ctr_geom g = geom(param_);
    return ctr_fan_angle_tables(&g, angles, xpos, ypos, sc) == CTR_OK;
}

bool FanBackprojection2DAnglePixDrivenGradPrep::calculate_on_gpu(const double* angles,
                                                                 double* xpos, double* ypos,
                                                                 double* sc,
                                                                 hipStream_t st) const
{
    return fan_angle_tables_gpu(geom(param_), angles, xpos, ypos, sc, st);
}

template <typename T>
bool FanBackprojection2DAnglePixDriven<T>::calculate_on_cpu(const T* proj, T* img,
                                                            const double* xpos,
                                                            const double* ypos,
                                                            const double* sc) const
{
    const auto& p = this->param_;
    const size_t ni = static_cast<size_t>(p.nx) * p.ny;
    Staging s;
    T* d = s.in(img, ni);
    return staged(s, img, d, ni, [&] {
        return calculate_on_gpu(s.in(proj, static_cast<size_t>(p.na) * p.ns), d,
                                s.in(xpos, p.nx), s.in(ypos, p.ny), s.in(sc, 2 * p.na),
                                s.stream());
    });
}

template <typename T>
bool FanBackprojection2DAnglePixDriven<T>::calculate_on_gpu(const T* proj, T* img,
                                                            const double* xpos,
                                                            const double* ypos,
                                                            const double* sc,
                                                            hipStream_t st) const
{
    const ctr_geom g = geom(this->param_);
    return ok_sync(ctr_fan_bp(&g, dtype_of<T>(), proj, img, 1, xpos, ypos, sc, 1.0, st), st);
}

template <typename T>
bool FanBackprojection2DAnglePixDrivenGrad<T>::calculate_on_cpu(const T* img, T* grad,
                                                                const double* xpos,
                                                                const double* ypos,
                                                                const double* sc) const
{
    const auto& p = this->param_;
    const size_t np = static_cast<size_t>(p.na) * p.ns;
    Staging s;
    T* d = s.in(grad, np);
    return staged(s, grad, d, np, [&] {
        return calculate_on_gpu(s.in(img, static_cast<size_t>(p.nx) * p.ny), d,
                                s.in(xpos, p.nx), s.in(ypos, p.ny), s.in(sc, 2 * p.na),
                                s.stream());
    });
}

template <typename T>
bool FanBackprojection2DAnglePixDrivenGrad<T>::calculate_on_gpu(const T* img, T* grad,
                                                                const double* xpos,
                                                                const double* ypos,
                                                                const double* sc,
                                                                hipStream_t st) const
{
    const ctr_geom g = geom(this->param_);
    return ok_sync(ctr_fan_bp_t(&g, dtype_of<T>(), img, grad, 1, xpos, ypos, sc, 1.0, 0, st),
                   st);
}

template class FanBackprojection2DAnglePixDriven<float>;
template class FanBackprojection2DAnglePixDriven<double>;
template class FanBackprojection2DAnglePixDrivenGrad<float>;
template class FanBackprojection2DAnglePixDrivenGrad<double>;

// -------------------------------------------------------------- bp_par_2d
bool ParallelBackprojection2DPixDrivenPrep::calculate_on_cpu(double* xcos, double* ysin,
                                                             int*) const
{
    const ctr_geom g = geom_par(param_);
    return ctr_par_tables(&g, xcos, ysin) == CTR_OK;
}

bool ParallelBackprojection2DPixDrivenPrep::calculate_on_gpu(double* xcos, double* ysin, int*,
                                                             hipStream_t st) const
{
    const ctr_geom g = geom_par(param_);
    std::vector<double> xc(static_cast<size_t>(g.na) * g.nx), ys(static_cast<size_t>(g.na) * g.ny);
    if (ctr_par_tables(&g, xc.data(), ys.data()) != CTR_OK) return false;
    return to_device(xcos, xc, st) && to_device(ysin, ys, st) && ok_sync(CTR_OK, st);
}

// begins[ia*ns + is], offsets[ia], usex[ia] as src/bp_par_2d.cc:95-145 builds them.
bool ParallelBackprojection2DPixDrivenGradPrep::calculate_on_cpu(double* begins, double* offsets,
                                                                 bool* usex) const
{
    const ctr_geom g = geom_par(param_);
    const double cs = ctr::cent_s(g), cx = ctr::cent_x(g), cy = ctr::cent_y(g);
    const double q = M_PI / 4;
    double angle = g.orbit_start;
    for (unsigned int a = 0; a < g.na; ++a) {
        // one sincos call, as the reference's g++ -O3 build makes of its sin
        // and cos of the angle (the cosine can differ from cos() by an ulp)
        double sa, ca;
        ::sincos(angle, &sa, &ca);
        while (angle < 0) angle += 2 * M_PI;
        while (angle >= 2 * M_PI) angle -= 2 * M_PI;
        const bool bx = !((angle >= q && angle < 3 * q) || (angle >= 5 * q && angle < 7 * q));
        usex[a] = bx;
        double o1, o2, begin;
        if (bx) {
            o1 = g.ds / (ca * g.dx);
            o2 = g.dy * sa / (ca * g.dx);
            begin = cx - cy * o2 - cs * o1;
        } else {
            o1 = -g.ds / (sa * g.dy);
            o2 = g.dx * ca / (sa * g.dy);
            begin = cy - cx * o2 - cs * o1;
        }
        offsets[a] = o2;
        for (unsigned int s = 0; s < g.ns; ++s) {
            begins[static_cast<size_t>(a) * g.ns + s] = begin;
            begin += o1;
        }
        angle += g.orbit;
    }
    return true;
}

bool ParallelBackprojection2DPixDrivenGradPrep::calculate_on_gpu(double* begins, double* offsets,
                                                                 bool* usex,
                                                                 hipStream_t st) const
{
    const ctr_geom g = geom_par(param_);
    std::vector<double> b(static_cast<size_t>(g.na) * g.ns), o(g.na);
    std::vector<unsigned char> u(g.na);
    if (!calculate_on_cpu(b.data(), o.data(), reinterpret_cast<bool*>(u.data()))) return false;
    return to_device(begins, b, st) && to_device(offsets, o, st) && to_device(usex, u, st) &&
           ok_sync(CTR_OK, st);
}

bool ParallelSingleViewBp2DPixDrivenPrep::calculate_on_cpu(double* xcos, double* ysin,
                                                           int*) const
{
    const ctr_geom g = geom_par(param_);
    return ctr_par_tables(&g, xcos, ysin) == CTR_OK;
}

bool ParallelSingleViewBp2DPixDrivenPrep::calculate_on_gpu(double* xcos, double* ysin, int*,
                                                           hipStream_t st) const
{
    const ctr_geom g = geom_par(param_);
    std::vector<double> xc(static_cast<size_t>(g.na) * g.nx), ys(static_cast<size_t>(g.na) * g.ny);
    if (ctr_par_tables(&g, xc.data(), ys.data()) != CTR_OK) return false;
    return to_device(xcos, xc, st) && to_device(ysin, ys, st) && ok_sync(CTR_OK, st);
}

template <typename T>
bool ParallelSingleViewBp2DPixDriven<T>::calculate_on_cpu(const T* proj, T* img,
                                                          const double* xcos,
                                                          const double* ysin,
                                                          const int*) const
{
    const auto& p = this->param_;
    const size_t ni = static_cast<size_t>(p.nx) * p.ny * p.na;
    Staging s;
    T* d = s.in(img, ni);
    return staged(s, img, d, ni, [&] {
        return calculate_on_gpu(s.in(proj, static_cast<size_t>(p.na) * p.ns), d,
                                s.in(xcos, static_cast<size_t>(p.na) * p.nx),
                                s.in(ysin, static_cast<size_t>(p.na) * p.ny), nullptr,
                                s.stream());
    });
}

template <typename T>
bool ParallelSingleViewBp2DPixDriven<T>::calculate_on_gpu(const T* proj, T* img,
                                                          const double* xcos,
                                                          const double* ysin, const int*,
                                                          hipStream_t st) const
{
    const ctr_geom g = geom_par(this->param_);
    return ok_sync(ctr_par_sv_bp(&g, dtype_of<T>(), proj, img, 1, xcos, ysin, st), st);
}

template class ParallelSingleViewBp2DPixDriven<float>;
template class ParallelSingleViewBp2DPixDriven<double>;

template <typename T>
bool ParallelBackprojection2DPixDriven<T>::calculate_on_cpu(const T* proj, T* img,
                                                            const double* xcos,
                                                            const double* ysin,
                                                            const int*) const
{
    const auto& p = this->param_;
    const size_t ni = static_cast<size_t>(p.nx) * p.ny;
    Staging s;
    T* d = s.in(img, ni);
    return staged(s, img, d, ni, [&] {
        return calculate_on_gpu(s.in(proj, static_cast<size_t>(p.na) * p.ns), d,
                                s.in(xcos, static_cast<size_t>(p.na) * p.nx),
                                s.in(ysin, static_cast<size_t>(p.na) * p.ny), nullptr,
                                s.stream());
    });
}

template <typename T>
bool ParallelBackprojection2DPixDriven<T>::calculate_on_gpu(const T* proj, T* img,
                                                            const double* xcos,
                                                            const double* ysin, const int*,
                                                            hipStream_t st) const
{
    const ctr_geom g = geom_par(this->param_);
    return ok_sync(ctr_par_bp(&g, dtype_of<T>(), proj, img, 1, xcos, ysin, st), st);
}

// The GPU transpose builds its own tables, so the Grad Prep's are not staged.
template <typename T>
bool ParallelBackprojection2DPixDrivenGrad<T>::calculate_on_cpu(const T* img, T* grad,
                                                                const double*, const double*,
                                                                const bool*) const
{
    const auto& p = this->param_;
    const size_t np = static_cast<size_t>(p.na) * p.ns;
    Staging s;
    T* d = s.in(grad, np);
    return staged(s, grad, d, np, [&] {
        return calculate_on_gpu(s.in(img, static_cast<size_t>(p.nx) * p.ny), d, nullptr,
                                nullptr, nullptr, s.stream());
    });
}

// The exact transpose needs the forward tables, built here per call.
template <typename T>
bool ParallelBackprojection2DPixDrivenGrad<T>::calculate_on_gpu(const T* img, T* grad,
                                                                const double*, const double*,
                                                                const bool*,
                                                                hipStream_t st) const
{
    const ctr_geom g = geom_par(this->param_);
    std::vector<double> xc(static_cast<size_t>(g.na) * g.nx), ys(static_cast<size_t>(g.na) * g.ny);
    if (ctr_par_tables(&g, xc.data(), ys.data()) != CTR_OK) return false;
    DevTemp dx(st), dy(st);
    if (!dx.put(xc) || !dy.put(ys)) return false;
    return ok_sync(ctr_par_bp_t(&g, dtype_of<T>(), img, grad, 1, static_cast<double*>(dx.p),
                                static_cast<double*>(dy.p), 0, st),
                   st);
}

template class ParallelBackprojection2DPixDriven<float>;
template class ParallelBackprojection2DPixDriven<double>;
template class ParallelBackprojection2DPixDrivenGrad<float>;
template class ParallelBackprojection2DPixDrivenGrad<double>;

// -------------------------------------------------------------- fp_par_2d
bool ParallelProjection2DRayDrivenPrepare::calculate_on_cpu(double* sc, double* bo,
                                                            int* usex) const
{
    const ctr_geom g = geom_par(param_);
    return ctr_fp_tables(&g, sc, bo, usex) == CTR_OK;
}

bool ParallelProjection2DRayDrivenPrepare::calculate_on_gpu(double* sc, double* bo, int* usex,
                                                            hipStream_t st) const
{
    const ctr_geom g = geom_par(param_);
    std::vector<double> s(2 * g.na), b(2 * static_cast<size_t>(g.na) * g.ns);
    std::vector<int> u(g.na);
    if (ctr_fp_tables(&g, s.data(), b.data(), u.data()) != CTR_OK) return false;
    return to_device(sc, s, st) && to_device(bo, b, st) && to_device(usex, u, st) &&
           ok_sync(CTR_OK, st);
}

bool ParallelProjection2DRayDrivenGradPrep::calculate_on_cpu(double* w, double* pos,
                                                             int* usex) const
{
    const ctr_geom g = geom_par(param_);
    return ctr_fp_grad_tables(&g, w, pos, usex, g.nx) == CTR_OK;
}

bool ParallelProjection2DRayDrivenGradPrep::calculate_on_gpu(double* w, double* pos, int* usex,
                                                             hipStream_t st) const
{
    const ctr_geom g = geom_par(param_);
    std::vector<double> wv(g.na), pv(2 * static_cast<size_t>(g.na) * g.nx);
    std::vector<int> u(g.na);
    if (ctr_fp_grad_tables(&g, wv.data(), pv.data(), u.data(), g.nx) != CTR_OK) return false;
    return to_device(w, wv, st) && to_device(pos, pv, st) && to_device(usex, u, st) &&
           ok_sync(CTR_OK, st);
}

template <typename T>
bool ParallelProjection2DRayDriven<T>::calculate_on_cpu(const T* img, T* proj, const double* sc,
                                                        const double* bo,
                                                        const int* usex) const
{
    const auto& p = this->param_;
    const size_t np = static_cast<size_t>(p.na) * p.ns;
    Staging s;
    T* d = s.in(proj, np);
    return staged(s, proj, d, np, [&] {
        return calculate_on_gpu(s.in(img, static_cast<size_t>(p.nx) * p.ny), d,
                                s.in(sc, 2 * p.na), s.in(bo, 2 * np), s.in(usex, p.na),
                                s.stream());
    });
}

template <typename T>
bool ParallelProjection2DRayDriven<T>::calculate_on_gpu(const T* img, T* proj, const double* sc,
                                                        const double* bo, const int* usex,
                                                        hipStream_t st) const
{
    const ctr_geom g = geom_par(this->param_);
    return ok_sync(ctr_fp(&g, dtype_of<T>(), img, proj, 1, sc, bo, usex, st), st);
}

template <typename T>
bool ParallelProjection2DRayDrivenGrad<T>::calculate_on_cpu(const T* proj, T* img,
                                                            const double* w, const double* pos,
                                                            const int* usex) const
{
    const auto& p = this->param_;
    const size_t ni = static_cast<size_t>(p.nx) * p.ny;
    Staging s;
    T* d = s.in(img, ni);
    return staged(s, img, d, ni, [&] {
        return calculate_on_gpu(s.in(proj, static_cast<size_t>(p.na) * p.ns), d, s.in(w, p.na),
                                s.in(pos, 2 * static_cast<size_t>(p.na) * p.nx),
                                s.in(usex, p.na), s.stream());
    });
}

template <typename T>
bool ParallelProjection2DRayDrivenGrad<T>::calculate_on_gpu(const T* proj, T* img,
                                                            const double* w, const double* pos,
                                                            const int* usex,
                                                            hipStream_t st) const
{
    const ctr_geom g = geom_par(this->param_);
    return ok_sync(ctr_fp_t(&g, dtype_of<T>(), proj, img, 1, w, pos, usex, g.nx, st), st);
}

template class ParallelProjection2DRayDriven<float>;
template class ParallelProjection2DRayDriven<double>;
template class ParallelProjection2DRayDrivenGrad<float>;
template class ParallelProjection2DRayDrivenGrad<double>;

// ------------------------------------------------------ fp_par_2d raycasting
bool ParallelProjection2DRayCastingPrepare::calculate_on_cpu(double* sc, double* begins,
                                                             int* nsteps) const
{
    const ctr_geom g = geom_par(param_);
    return ctr_fp_rc_tables(&g, step_size_, sc, begins, nsteps) == CTR_OK;
}

bool ParallelProjection2DRayCastingPrepare::calculate_on_gpu(double* sc, double* begins,
                                                             int* nsteps, hipStream_t st) const
{
    const ctr_geom g = geom_par(param_);
    std::vector<double> s(2 * g.na), b(2 * static_cast<size_t>(g.na) * g.ns);
    std::vector<int> n(static_cast<size_t>(g.na) * g.ns);
    if (ctr_fp_rc_tables(&g, step_size_, s.data(), b.data(), n.data()) != CTR_OK) return false;
    return to_device(sc, s, st) && to_device(begins, b, st) && to_device(nsteps, n, st) &&
           ok_sync(CTR_OK, st);
}

template <typename T>
bool ParallelProjection2DRayCasting<T>::calculate_on_cpu(const T* img, T* proj, const double* sc,
                                                         const double* begins,
                                                         const int* nsteps) const
{
    const auto& p = this->param_;
    const size_t np = static_cast<size_t>(p.na) * p.ns;
    Staging s;
    T* d = s.in(proj, np);
    return staged(s, proj, d, np, [&] {
        return calculate_on_gpu(s.in(img, static_cast<size_t>(p.nx) * p.ny), d,
                                s.in(sc, 2 * p.na), s.in(begins, 2 * np), s.in(nsteps, np),
                                s.stream());
    });
}

template <typename T>
bool ParallelProjection2DRayCasting<T>::calculate_on_gpu(const T* img, T* proj, const double* sc,
                                                         const double* begins,
                                                         const int* nsteps, hipStream_t st) const
{
    const ctr_geom g = geom_par(this->param_);
    return ok_sync(ctr_fp_rc(&g, dtype_of<T>(), img, proj, 1, step_size_, sc, begins, nsteps, st),
                   st);
}

template class ParallelProjection2DRayCasting<float>;
template class ParallelProjection2DRayCasting<double>;

// ---------------------------------------------------- fp_par_2d disdriven
bool ParallelProjection2DDisDrivenPrep::calculate_on_cpu(double* sc, double* bo, int* usex) const
{
    const ctr_geom g = geom_par(param_);
    return ctr_fp_dd_tables(&g, sc, bo, usex) == CTR_OK;
}

bool ParallelProjection2DDisDrivenPrep::calculate_on_gpu(double* sc, double* bo, int* usex,
                                                         hipStream_t st) const
{
    const ctr_geom g = geom_par(param_);
    std::vector<double> s(2 * g.na), b(3 * static_cast<size_t>(g.na) * g.ns);
    std::vector<int> u(g.na);
    if (ctr_fp_dd_tables(&g, s.data(), b.data(), u.data()) != CTR_OK) return false;
    return to_device(sc, s, st) && to_device(bo, b, st) && to_device(usex, u, st) &&
           ok_sync(CTR_OK, st);
}

bool ParallelProjection2DDisDrivenGradPrep::calculate_on_cpu(double* w, double* pos,
                                                             int* usex) const
{
    const ctr_geom g = geom_par(param_);
    return ctr_fp_grad_tables(&g, w, pos, usex, g.nx) == CTR_OK;
}

bool ParallelProjection2DDisDrivenGradPrep::calculate_on_gpu(double* w, double* pos, int* usex,
                                                             hipStream_t st) const
{
    const ctr_geom g = geom_par(param_);
    std::vector<double> wv(g.na), pv(2 * static_cast<size_t>(g.na) * g.nx);
    std::vector<int> u(g.na);
    if (ctr_fp_grad_tables(&g, wv.data(), pv.data(), u.data(), g.nx) != CTR_OK) return false;
    return to_device(w, wv, st) && to_device(pos, pv, st) && to_device(usex, u, st) &&
           ok_sync(CTR_OK, st);
}

template <typename T>
bool ParallelProjection2DDisDriven<T>::calculate_on_cpu(const T* img, T* proj, const double* sc,
                                                        const double* bo,
                                                        const int* usex) const
{
    const auto& p = this->param_;
    const size_t np = static_cast<size_t>(p.na) * p.ns;
    Staging s;
    T* d = s.in(proj, np);
    return staged(s, proj, d, np, [&] {
        return calculate_on_gpu(s.in(img, static_cast<size_t>(p.nx) * p.ny), d,
                                s.in(sc, 2 * p.na), s.in(bo, 3 * np), s.in(usex, p.na),
                                s.stream());
    });
}

template <typename T>
bool ParallelProjection2DDisDriven<T>::calculate_on_gpu(const T* img, T* proj, const double* sc,
                                                        const double* bo, const int* usex,
                                                        hipStream_t st) const
{
    const ctr_geom g = geom_par(this->param_);
    return ok_sync(ctr_fp_dd(&g, dtype_of<T>(), img, proj, 1, sc, bo, usex, st), st);
}

template <typename T>
bool ParallelProjection2DDisDrivenGrad<T>::calculate_on_cpu(const T* proj, T* img,
                                                            const double* w, const double* pos,
                                                            const int* usex) const
{
    const auto& p = this->param_;
    const size_t ni = static_cast<size_t>(p.nx) * p.ny;
    Staging s;
    T* d = s.in(img, ni);
    return staged(s, img, d, ni, [&] {
        return calculate_on_gpu(s.in(proj, static_cast<size_t>(p.na) * p.ns), d, s.in(w, p.na),
                                s.in(pos, 2 * static_cast<size_t>(p.na) * p.nx),
                                s.in(usex, p.na), s.stream());
    });
}

template <typename T>
bool ParallelProjection2DDisDrivenGrad<T>::calculate_on_gpu(const T* proj, T* img,
                                                            const double* w, const double* pos,
                                                            const int* usex, hipStream_t st) const
{
    const ctr_geom g = geom_par(this->param_);
    return ok_sync(ctr_fp_dd_grad(&g, dtype_of<T>(), proj, img, 1, w, pos, usex, g.nx, st), st);
}

template class ParallelProjection2DDisDriven<float>;
template class ParallelProjection2DDisDriven<double>;
template class ParallelProjection2DDisDrivenGrad<float>;
template class ParallelProjection2DDisDrivenGrad<double>;

// ----------------------------------------------------------------- filter
bool RampFilterPrep::calculate_on_cpu(double* filter) const
{
    const ctr_geom g = geom(param_);
    return ctr_ramp_taps(&g, filter) == CTR_OK;
}

bool RampFilterPrep::calculate_on_gpu(double* filter, hipStream_t st) const
{
    const ctr_geom g = geom(param_);
    std::vector<double> h(2 * static_cast<size_t>(g.ns) + 1);
    if (ctr_ramp_taps(&g, h.data()) != CTR_OK) return false;
    return to_device(filter, h, st) && ok_sync(CTR_OK, st);
}

template <typename T>
bool RampFilter<T>::calculate_on_cpu(const T* in, const double* filter, T* out) const
{
    const size_t n = static_cast<size_t>(param_.nrow) * param_.ns;
    Staging s;
    T* d = s.in(out, n);
    return staged(s, out, d, n, [&] {
        return calculate_on_gpu(s.in(in, n), s.in(filter, 2 * static_cast<size_t>(param_.ns) + 1),
                                d, s.stream());
    });
}

template <typename T>
bool RampFilter<T>::calculate_on_gpu(const T* in, const double* filter, T* out,
                                     hipStream_t st) const
{
    const ctr_geom g = geom(param_);
    return ok_sync(ctr_ramp(&g, dtype_of<T>(), in, out, g.nrow, filter, 0, st), st);
}

template <typename T>
bool RampFilterGrad<T>::calculate_on_cpu(const T* in, const double* filter, T* out) const
{
    const size_t n = static_cast<size_t>(param_.nrow) * param_.ns;
    Staging s;
    T* d = s.in(out, n);
    return staged(s, out, d, n, [&] {
        return calculate_on_gpu(s.in(in, n), s.in(filter, 2 * static_cast<size_t>(param_.ns) + 1),
                                d, s.stream());
    });
}

template <typename T>
bool RampFilterGrad<T>::calculate_on_gpu(const T* in, const double* filter, T* out,
                                         hipStream_t st) const
{
    const ctr_geom g = geom(param_);
    return ok_sync(ctr_ramp(&g, dtype_of<T>(), in, out, g.nrow, filter, 0, st), st);
}

template class RampFilter<float>;
template class RampFilter<double>;
template class RampFilterGrad<float>;
template class RampFilterGrad<double>;

// --------------------------------------------------------- fan_weighting
template <typename T>
bool FanWeighting<T>::calculate_on_cpu(const T* in, T* out) const
{
    const size_t n = static_cast<size_t>(param_.nrow) * param_.ns;
    Staging s;
    T* d = s.in(out, n);
    return staged(s, out, d, n, [&] { return calculate_on_gpu(s.in(in, n), d, s.stream()); });
}

template <typename T>
bool FanWeighting<T>::calculate_on_gpu(const T* in, T* out, hipStream_t st) const
{
    const ctr_geom g = geom(param_);
    return ok_sync(ctr_fan_weighting(&g, dtype_of<T>(), in, out, g.nrow, st), st);
}

template <typename T>
bool FanWeightingGrad<T>::calculate_on_cpu(const T* in, T* out) const
{
    const size_t n = static_cast<size_t>(param_.nrow) * param_.ns;
    Staging s;
    T* d = s.in(out, n);
    return staged(s, out, d, n, [&] { return calculate_on_gpu(s.in(in, n), d, s.stream()); });
}

template <typename T>
bool FanWeightingGrad<T>::calculate_on_gpu(const T* in, T* out, hipStream_t st) const
{
    const ctr_geom g = geom(param_);
    return ok_sync(ctr_fan_weighting(&g, dtype_of<T>(), in, out, g.nrow, st), st);
}

template class FanWeighting<float>;
template class FanWeighting<double>;
template class FanWeightingGrad<float>;
template class FanWeightingGrad<double>;

// ----------------------------------------------------- helical rebinning
namespace {

ctr_helical_geom geom(const HelicalInterpolationParam& p)
{
    ctr_helical_geom g{};
    g.ns = p.ns; g.nt = p.nt; g.na = p.na; g.ds = p.ds; g.dt = p.dt;
    g.offset_s = p.offset_s; g.offset_t = p.offset_t; g.orbit_start = p.orbit_start;
    g.orbit = p.orbit; g.couch_begin = p.couch_begin; g.couch_mov = p.couch_mov;
    g.view_per_rot = p.view_per_rot; g.nz = p.nz; g.dz = p.dz; g.z0 = p.z0;
    g.thickness = p.thickness; g.dso = p.dso; g.dsd = p.dsd;
    return g;
}

// Tables on the host, one small device block for them and the status word,
// the kernel, then the reference's exceptions from the status.
template <typename T>
bool helical_gpu(const HelicalInterpolationParam& p, int mode, const T* in, T* out,
                 hipStream_t st)
{
    const ctr_helical_geom g = geom(p);
    std::vector<double> tab(3 * static_cast<size_t>(p.ns) + 4 * static_cast<size_t>(p.nz));
    const int rc = ctr_helical_tables(&g, mode, tab.data());
    if (rc == CTR_ERR_SHAPE &&
        std::string(ctr::get_error()).find("View not enough") != std::string::npos)
        throw std::runtime_error("View not enough for this slice!");
    if (rc != CTR_OK) return false;
    const size_t tbytes = tab.size() * sizeof(double);
    void* dev = nullptr;
    hipError_t e = hipMalloc(&dev, tbytes + sizeof(unsigned int));
    if (e != hipSuccess) {
        ctr::set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return false;
    }
    struct Release {
        void* p;
        ~Release() { (void)hipFree(p); }
    } release{dev};
    unsigned int* status = reinterpret_cast<unsigned int*>(static_cast<char*>(dev) + tbytes);
    e = hipMemcpyAsync(dev, tab.data(), tbytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(status, 0, sizeof(unsigned int), st);
    if (e != hipSuccess) {
        ctr::set_error(std::string("helical tables upload: ") + hipGetErrorString(e));
        return false;
    }
    if (!ok_sync(ctr_helical(&g, mode, dtype_of<T>(), in, out, 1,
                             static_cast<const double*>(dev), status, st),
                 st))
        return false;
    unsigned int code = 0;
    e = hipMemcpy(&code, status, sizeof(code), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        ctr::set_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
        return false;
    }
    if (code & 1u) throw std::runtime_error("No row data available!");
    if (code & 2u)
        throw std::runtime_error("180LI conjugate view outside the scan");
    return true;
}

}  // namespace

template <typename T>
bool HelicalInterpolation360<T>::calculate_on_cpu(const T* in, T* out) const
{
    const auto& p = this->param_;
    const size_t no = static_cast<size_t>(p.nz) * p.view_per_rot * p.ns;
    Staging s;
    T* d = s.in(out, no);
    return staged(s, out, d, no, [&] {
        return calculate_on_gpu(s.in(in, static_cast<size_t>(p.na) * p.nt * p.ns), d,
                                s.stream());
    });
}

template <typename T>
bool HelicalInterpolation360<T>::calculate_on_gpu(const T* in, T* out, hipStream_t st) const
{
    return helical_gpu<T>(this->param_, CTR_HELICAL_360, in, out, st);
}

template <typename T>
bool HelicalInterpolation180<T>::calculate_on_cpu(const T* in, T* out) const
{
    const auto& p = this->param_;
    const size_t no = static_cast<size_t>(p.nz) * p.view_per_rot * p.ns;
    Staging s;
    T* d = s.in(out, no);
    return staged(s, out, d, no, [&] {
        return calculate_on_gpu(s.in(in, static_cast<size_t>(p.na) * p.nt * p.ns), d,
                                s.stream());
    });
}

template <typename T>
bool HelicalInterpolation180<T>::calculate_on_gpu(const T* in, T* out, hipStream_t st) const
{
    return helical_gpu<T>(this->param_, CTR_HELICAL_180, in, out, st);
}

template class HelicalInterpolation360<float>;
template class HelicalInterpolation360<double>;
template class HelicalInterpolation180<float>;
template class HelicalInterpolation180<double>;

}  // namespace ct_recon
