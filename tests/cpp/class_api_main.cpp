// Drives the ct_recon:: class API the way a user of the reference's C++
// library does (Prep -> operator -> Grad, device buffers, a stream), for
// tests/test_class_api.py.  Usage:
//   class_api_main cpu <dir>   tables from every Prep::calculate_on_cpu
//   class_api_main gpu <dir>   every operator on the GPU, float and double
//                              (helical rebinning included), then every
//                              operator again through calculate_on_cpu
// Inputs are read from <dir>/*.in.f64 and outputs written to <dir>/*.f64 as
// raw little-endian doubles; geometries are fixed and mirrored in the test.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "include/bp_fan_2d.h"
#include "include/bp_fan_2d_angle.h"
#include "include/bp_par_2d.h"
#include "include/bp_par_2d_sv.h"
#include "include/fan_weighting.h"
#include "include/filter.h"
#include "include/fp_par_2d.h"
#include "include/helical_interpolation.h"

#include <stdexcept>

using namespace ct_recon;

static std::string g_dir;

static std::vector<double> load(const std::string& name, size_t n)
{
    std::vector<double> v(n);
    FILE* f = std::fopen((g_dir + "/" + name + ".in.f64").c_str(), "rb");
    if (!f || std::fread(v.data(), 8, n, f) != n) {
        std::fprintf(stderr, "cannot read %s\n", name.c_str());
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

template <typename V>
static void save(const std::string& name, const std::vector<V>& v)
{
    std::vector<double> d(v.begin(), v.end());
    FILE* f = std::fopen((g_dir + "/" + name + ".f64").c_str(), "wb");
    std::fwrite(d.data(), 8, d.size(), f);
    std::fclose(f);
}

#define CHECK(x)                                                            \
    do {                                                                    \
        if (!(x)) {                                                         \
            std::fprintf(stderr, "FAILED: %s (line %d)\n", #x, __LINE__);   \
            return 1;                                                       \
        }                                                                   \
    } while (0)

template <typename T>
struct Dev {
    T* p = nullptr;
    size_t n;
    explicit Dev(size_t n) : n(n) { (void)hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T)); }
    Dev(const std::vector<double>& h) : Dev(h.size())
    {
        std::vector<T> t(h.begin(), h.end());
        (void)hipMemcpy(p, t.data(), n * sizeof(T), hipMemcpyHostToDevice);
    }
    ~Dev() { (void)hipFree(p); }
    std::vector<T> get() const
    {
        std::vector<T> h(n);
        (void)hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost);
        return h;
    }
};

// Geometries (mirrored in tests/test_class_api.py).
static const FanBackprojection2DParam kFan(41, 48, 1.7, 2 * M_PI / 48, 0.4, 0.3, 24, 20, 1.3, 1.5,
                                          0.3, -0.2, 60.0, 110.0, -1.0);
static const FanBackprojection2DAngleParam kFanA(41, 48, 1.7, 0.4, 24, 20, 1.3, 1.5, 0.3, -0.2,
                                                 60.0, 110.0, -1.0);
static const ParallelBackprojection2DParam kPar(41, 36, 0.8, M_PI / 36, -0.35, 0.1, 22, 26, 0.7,
                                                0.9, -0.4, 0.25, -1.0);
static const ParallelSingleViewBp2DParam kSv(41, 36, 0.8, M_PI / 36, -0.35, 0.1, 22, 26, 0.7,
                                              0.9, -0.4, 0.25, -1.0);
static const ParallelProjection2DParam kFp(41, 40, 1.0, M_PI / 40, 0.3, 0.0, 24, 24, 1.0, 1.0,
                                           0.2, -0.1, -1.0);
static const ParallelProjection2DParam kFpRc(41, 40, 1.0, M_PI / 40, 0.3, 0.0, 24, 24, 1.0, 1.0,
                                             0.2, -0.1, 11.0);
static const FilterParam kFilt(33, 5, 0.9, 40.0, 1, 0);
static const FanWeightingParam kFw(37, 4, 1.3, 0.6, 595.0, 1085.6, 1);

static int cpu()
{
    std::vector<double> x(kFan.nx), y(kFan.ny), sc(2 * kFan.na);
    CHECK(FanBackprojection2DPixDrivenPrep(kFan).calculate_on_cpu(x.data(), y.data(), sc.data()));
    save("fan_xpos", x); save("fan_ypos", y); save("fan_sincos", sc);
    std::vector<double> xc(kPar.na * kPar.nx), ys(kPar.na * kPar.ny);
    CHECK(ParallelBackprojection2DPixDrivenPrep(kPar).calculate_on_cpu(xc.data(), ys.data(), nullptr));
    save("par_xcos", xc); save("par_ysin", ys);
    std::vector<double> sxc(kSv.na * kSv.nx), sys(kSv.na * kSv.ny);
    CHECK(ParallelSingleViewBp2DPixDrivenPrep(kSv).calculate_on_cpu(sxc.data(), sys.data(), nullptr));
    save("svbp_xcos", sxc); save("svbp_ysin", sys);
    std::vector<double> s2(2 * kFp.na), bo(2 * kFp.na * kFp.ns);
    std::vector<int> u(kFp.na);
    CHECK(ParallelProjection2DRayDrivenPrepare(kFp).calculate_on_cpu(s2.data(), bo.data(), u.data()));
    save("fp_sincos", s2); save("fp_beginoffset", bo); save("fp_usex", u);
    std::vector<double> w(kFp.na), pos(2 * kFp.na * kFp.nx);
    CHECK(ParallelProjection2DRayDrivenGradPrep(kFp).calculate_on_cpu(w.data(), pos.data(), u.data()));
    save("fpg_weights", w); save("fpg_pos", pos); save("fpg_usex", u);
    std::vector<double> d3(3 * kFp.na * kFp.ns);
    CHECK(ParallelProjection2DDisDrivenPrep(kFp).calculate_on_cpu(s2.data(), d3.data(), u.data()));
    save("fpdd_sincos", s2); save("fpdd_beginoffset", d3); save("fpdd_usex", u);
    std::vector<double> rb(2 * kFp.na * kFp.ns);
    std::vector<int> rn(kFp.na * kFp.ns);
    CHECK(ParallelProjection2DRayCastingPrepare(kFpRc).calculate_on_cpu(s2.data(), rb.data(), rn.data()));
    save("fprc_begins", rb); save("fprc_nsteps", rn);
    std::vector<double> h(2 * kFilt.ns + 1);
    CHECK(RampFilterPrep(kFilt).calculate_on_cpu(h.data()));
    save("ramp_taps", h);
    std::puts("cpu ok");
    return 0;
}

template <typename T>
static int gpu_typed(const char* suf, hipStream_t st)
{
    const std::string s = suf;
    {   // fan BP and its transpose
        Dev<double> x(kFan.nx), y(kFan.ny), sc(2 * kFan.na);
        CHECK(FanBackprojection2DPixDrivenPrep(kFan).calculate_on_gpu(x.p, y.p, sc.p, st));
        Dev<T> proj(load("fan_proj", kFan.na * kFan.ns)), img(kFan.nx * kFan.ny);
        CHECK(FanBackprojection2DPixDriven<T>(kFan).calculate_on_gpu(proj.p, img.p, x.p, y.p, sc.p, st));
        save("fan_bp_" + s, img.get());
        Dev<double> gx(kFan.nx), gy(kFan.ny), gsc(2 * kFan.na);
        CHECK(FanBackprojection2DPixDrivenGradPrep(kFan).calculate_on_gpu(gx.p, gy.p, gsc.p, st));
        Dev<T> im2(load("fan_img", kFan.nx * kFan.ny)), grad(kFan.na * kFan.ns);
        CHECK(FanBackprojection2DPixDrivenGrad<T>(kFan).calculate_on_gpu(im2.p, grad.p, gx.p, gy.p, gsc.p, st));
        save("fan_bpt_" + s, grad.get());
    }
    {   // fan BP at given angles
        Dev<double> ang(load("fana_angles", kFanA.na));
        Dev<double> x(kFanA.nx), y(kFanA.ny), sc(2 * kFanA.na);
        CHECK(FanBackprojection2DAnglePixDrivenPrep(kFanA).calculate_on_gpu(ang.p, x.p, y.p, sc.p, st));
        Dev<T> proj(load("fan_proj", kFanA.na * kFanA.ns)), img(kFanA.nx * kFanA.ny);
        CHECK(FanBackprojection2DAnglePixDriven<T>(kFanA).calculate_on_gpu(proj.p, img.p, x.p, y.p, sc.p, st));
        save("fana_bp_" + s, img.get());
    }
    {   // parallel BP and its transpose
        Dev<double> xc(kPar.na * kPar.nx), ys(kPar.na * kPar.ny);
        CHECK(ParallelBackprojection2DPixDrivenPrep(kPar).calculate_on_gpu(xc.p, ys.p, nullptr, st));
        Dev<T> proj(load("par_proj", kPar.na * kPar.ns)), img(kPar.nx * kPar.ny);
        CHECK(ParallelBackprojection2DPixDriven<T>(kPar).calculate_on_gpu(proj.p, img.p, xc.p, ys.p, nullptr, st));
        save("par_bp_" + s, img.get());
        Dev<double> b(kPar.na * kPar.ns), o(kPar.na);
        Dev<bool> u(kPar.na);
        CHECK(ParallelBackprojection2DPixDrivenGradPrep(kPar).calculate_on_gpu(b.p, o.p, u.p, st));
        Dev<T> im2(load("par_img", kPar.nx * kPar.ny)), grad(kPar.na * kPar.ns);
        CHECK(ParallelBackprojection2DPixDrivenGrad<T>(kPar).calculate_on_gpu(im2.p, grad.p, b.p, o.p, u.p, st));
        save("par_bpt_" + s, grad.get());
    }
    {   // single-view parallel BP (the parallel geometry)
        Dev<double> xc(kSv.na * kSv.nx), ys(kSv.na * kSv.ny);
        CHECK(ParallelSingleViewBp2DPixDrivenPrep(kSv).calculate_on_gpu(xc.p, ys.p, nullptr, st));
        Dev<T> proj(load("par_proj", kSv.na * kSv.ns)), img(kSv.nx * kSv.ny * kSv.na);
        CHECK(ParallelSingleViewBp2DPixDriven<T>(kSv).calculate_on_gpu(proj.p, img.p, xc.p, ys.p, nullptr, st));
        save("svbp_" + s, img.get());
    }
    {   // ray-driven FP and its transpose
        Dev<double> sc(2 * kFp.na), bo(2 * kFp.na * kFp.ns);
        Dev<int> u(kFp.na);
        CHECK(ParallelProjection2DRayDrivenPrepare(kFp).calculate_on_gpu(sc.p, bo.p, u.p, st));
        Dev<T> img(load("fp_img", kFp.nx * kFp.ny)), proj(kFp.na * kFp.ns);
        CHECK(ParallelProjection2DRayDriven<T>(kFp).calculate_on_gpu(img.p, proj.p, sc.p, bo.p, u.p, st));
        save("fp_" + s, proj.get());
        Dev<double> w(kFp.na), pos(2 * kFp.na * kFp.nx);
        Dev<int> ug(kFp.na);
        CHECK(ParallelProjection2DRayDrivenGradPrep(kFp).calculate_on_gpu(w.p, pos.p, ug.p, st));
        Dev<T> pr2(load("fp_proj", kFp.na * kFp.ns)), img2(kFp.nx * kFp.ny);
        CHECK(ParallelProjection2DRayDrivenGrad<T>(kFp).calculate_on_gpu(pr2.p, img2.p, w.p, pos.p, ug.p, st));
        save("fpt_" + s, img2.get());
    }
    {   // distance-driven FP and the reference's gradient for it
        Dev<double> sc(2 * kFp.na), bo(3 * kFp.na * kFp.ns);
        Dev<int> u(kFp.na);
        CHECK(ParallelProjection2DDisDrivenPrep(kFp).calculate_on_gpu(sc.p, bo.p, u.p, st));
        Dev<T> img(load("fp_img", kFp.nx * kFp.ny)), proj(kFp.na * kFp.ns);
        CHECK(ParallelProjection2DDisDriven<T>(kFp).calculate_on_gpu(img.p, proj.p, sc.p, bo.p, u.p, st));
        save("fpdd_" + s, proj.get());
        Dev<double> w(kFp.na), pos(2 * kFp.na * kFp.nx);
        Dev<int> ug(kFp.na);
        CHECK(ParallelProjection2DDisDrivenGradPrep(kFp).calculate_on_gpu(w.p, pos.p, ug.p, st));
        Dev<T> pr2(load("fp_proj", kFp.na * kFp.ns)), img2(kFp.nx * kFp.ny);
        CHECK(ParallelProjection2DDisDrivenGrad<T>(kFp).calculate_on_gpu(pr2.p, img2.p, w.p, pos.p, ug.p, st));
        save("fpddg_" + s, img2.get());
    }
    {   // ray-casting FP (FOV circle of radius 11 inside the 24 x 24 image)
        Dev<double> sc(2 * kFpRc.na), bg(2 * kFpRc.na * kFpRc.ns);
        Dev<int> ns(kFpRc.na * kFpRc.ns);
        CHECK(ParallelProjection2DRayCastingPrepare(kFpRc).calculate_on_gpu(sc.p, bg.p, ns.p, st));
        Dev<T> img(load("fp_img", kFpRc.nx * kFpRc.ny)), proj(kFpRc.na * kFpRc.ns);
        CHECK(ParallelProjection2DRayCasting<T>(kFpRc).calculate_on_gpu(img.p, proj.p, sc.p, bg.p, ns.p, st));
        save("fprc_" + s, proj.get());
    }
    {   // ramp filter, its transpose, fan weighting
        Dev<double> h(2 * kFilt.ns + 1);
        CHECK(RampFilterPrep(kFilt).calculate_on_gpu(h.p, st));
        Dev<T> in(load("ramp_in", kFilt.nrow * kFilt.ns)), out(kFilt.nrow * kFilt.ns);
        CHECK(RampFilter<T>(kFilt).calculate_on_gpu(in.p, h.p, out.p, st));
        save("ramp_" + s, out.get());
        CHECK(RampFilterGrad<T>(kFilt).calculate_on_gpu(in.p, h.p, out.p, st));
        save("rampg_" + s, out.get());
        Dev<T> wi(load("fw_in", kFw.nrow * kFw.ns)), wo(kFw.nrow * kFw.ns);
        CHECK(FanWeighting<T>(kFw).calculate_on_gpu(wi.p, wo.p, st));
        save("fw_" + s, wo.get());
        CHECK(FanWeightingGrad<T>(kFw).calculate_on_gpu(wi.p, wo.p, st));
        save("fwg_" + s, wo.get());
    }
    return 0;
}

// Every operator through calculate_on_cpu: host tables from the Prep
// classes, host inputs and outputs; saved as <name>_host_<suffix> for a
// bitwise comparison with the device runs above.
template <typename T>
static std::vector<T> host_in(const std::string& name, size_t n)
{
    const std::vector<double> d = load(name, n);
    return std::vector<T>(d.begin(), d.end());
}

template <typename T>
static int host_typed(const char* suf)
{
    const std::string s = std::string("_host_") + suf;
    {
        std::vector<double> x(kFan.nx), y(kFan.ny), sc(2 * kFan.na);
        CHECK(FanBackprojection2DPixDrivenPrep(kFan).calculate_on_cpu(x.data(), y.data(), sc.data()));
        std::vector<T> proj = host_in<T>("fan_proj", kFan.na * kFan.ns), img(kFan.nx * kFan.ny);
        CHECK(FanBackprojection2DPixDriven<T>(kFan).calculate_on_cpu(proj.data(), img.data(), x.data(), y.data(), sc.data()));
        save("fan_bp" + s, img);
        std::vector<T> im2 = host_in<T>("fan_img", kFan.nx * kFan.ny), grad(kFan.na * kFan.ns);
        CHECK(FanBackprojection2DPixDrivenGrad<T>(kFan).calculate_on_cpu(im2.data(), grad.data(), x.data(), y.data(), sc.data()));
        save("fan_bpt" + s, grad);
    }
    {
        const std::vector<double> ang = load("fana_angles", kFanA.na);
        std::vector<double> x(kFanA.nx), y(kFanA.ny), sc(2 * kFanA.na);
        CHECK(FanBackprojection2DAnglePixDrivenPrep(kFanA).calculate_on_cpu(ang.data(), x.data(), y.data(), sc.data()));
        std::vector<T> proj = host_in<T>("fan_proj", kFanA.na * kFanA.ns), img(kFanA.nx * kFanA.ny);
        CHECK(FanBackprojection2DAnglePixDriven<T>(kFanA).calculate_on_cpu(proj.data(), img.data(), x.data(), y.data(), sc.data()));
        save("fana_bp" + s, img);
    }
    {
        std::vector<double> xc(kPar.na * kPar.nx), ys(kPar.na * kPar.ny);
        CHECK(ParallelBackprojection2DPixDrivenPrep(kPar).calculate_on_cpu(xc.data(), ys.data(), nullptr));
        std::vector<T> proj = host_in<T>("par_proj", kPar.na * kPar.ns), img(kPar.nx * kPar.ny);
        CHECK(ParallelBackprojection2DPixDriven<T>(kPar).calculate_on_cpu(proj.data(), img.data(), xc.data(), ys.data(), nullptr));
        save("par_bp" + s, img);
        std::vector<T> im2 = host_in<T>("par_img", kPar.nx * kPar.ny), grad(kPar.na * kPar.ns);
        CHECK(ParallelBackprojection2DPixDrivenGrad<T>(kPar).calculate_on_cpu(im2.data(), grad.data(), nullptr, nullptr, nullptr));
        save("par_bpt" + s, grad);
        std::vector<T> sv(kSv.nx * kSv.ny * kSv.na);
        CHECK(ParallelSingleViewBp2DPixDriven<T>(kSv).calculate_on_cpu(proj.data(), sv.data(), xc.data(), ys.data(), nullptr));
        save("svbp" + s, sv);
    }
    {
        std::vector<double> sc(2 * kFp.na), bo(2 * kFp.na * kFp.ns), w(kFp.na), pos(2 * kFp.na * kFp.nx);
        std::vector<int> u(kFp.na), ug(kFp.na);
        CHECK(ParallelProjection2DRayDrivenPrepare(kFp).calculate_on_cpu(sc.data(), bo.data(), u.data()));
        CHECK(ParallelProjection2DRayDrivenGradPrep(kFp).calculate_on_cpu(w.data(), pos.data(), ug.data()));
        std::vector<T> img = host_in<T>("fp_img", kFp.nx * kFp.ny), proj(kFp.na * kFp.ns);
        CHECK(ParallelProjection2DRayDriven<T>(kFp).calculate_on_cpu(img.data(), proj.data(), sc.data(), bo.data(), u.data()));
        save("fp" + s, proj);
        std::vector<T> pr2 = host_in<T>("fp_proj", kFp.na * kFp.ns), img2(kFp.nx * kFp.ny);
        CHECK(ParallelProjection2DRayDrivenGrad<T>(kFp).calculate_on_cpu(pr2.data(), img2.data(), w.data(), pos.data(), ug.data()));
        save("fpt" + s, img2);
        std::vector<double> bo3(3 * kFp.na * kFp.ns);
        CHECK(ParallelProjection2DDisDrivenPrep(kFp).calculate_on_cpu(sc.data(), bo3.data(), u.data()));
        CHECK(ParallelProjection2DDisDriven<T>(kFp).calculate_on_cpu(img.data(), proj.data(), sc.data(), bo3.data(), u.data()));
        save("fpdd" + s, proj);
        CHECK(ParallelProjection2DDisDrivenGradPrep(kFp).calculate_on_cpu(w.data(), pos.data(), ug.data()));
        CHECK(ParallelProjection2DDisDrivenGrad<T>(kFp).calculate_on_cpu(pr2.data(), img2.data(), w.data(), pos.data(), ug.data()));
        save("fpddg" + s, img2);
        std::vector<double> rb(2 * kFpRc.na * kFpRc.ns);
        std::vector<int> rn(kFpRc.na * kFpRc.ns);
        CHECK(ParallelProjection2DRayCastingPrepare(kFpRc).calculate_on_cpu(sc.data(), rb.data(), rn.data()));
        CHECK(ParallelProjection2DRayCasting<T>(kFpRc).calculate_on_cpu(img.data(), proj.data(), sc.data(), rb.data(), rn.data()));
        save("fprc" + s, proj);
    }
    {
        std::vector<double> h(2 * kFilt.ns + 1);
        CHECK(RampFilterPrep(kFilt).calculate_on_cpu(h.data()));
        std::vector<T> in = host_in<T>("ramp_in", kFilt.nrow * kFilt.ns), out(kFilt.nrow * kFilt.ns);
        CHECK(RampFilter<T>(kFilt).calculate_on_cpu(in.data(), h.data(), out.data()));
        save("ramp" + s, out);
        CHECK(RampFilterGrad<T>(kFilt).calculate_on_cpu(in.data(), h.data(), out.data()));
        save("rampg" + s, out);
        std::vector<T> wi = host_in<T>("fw_in", kFw.nrow * kFw.ns), wo(kFw.nrow * kFw.ns);
        CHECK(FanWeighting<T>(kFw).calculate_on_cpu(wi.data(), wo.data()));
        save("fw" + s, wo);
        CHECK(FanWeightingGrad<T>(kFw).calculate_on_cpu(wi.data(), wo.data()));
        save("fwg" + s, wo);
    }
    // a null input is an error, reported as false
    std::vector<T> o1(kFw.nrow * kFw.ns);
    CHECK(!FanWeighting<T>(kFw).calculate_on_cpu(nullptr, o1.data()));
    return 0;
}

// Helical rebinning; the geometry comes from <dir>/heli_geom.in.f64 (the
// 18 fields in constructor order), the scan from heli_proj.in.f64.
static HelicalInterpolationParam heli_geom(const std::vector<double>& v)
{
    return HelicalInterpolationParam(
        static_cast<unsigned>(v[0]), static_cast<unsigned>(v[1]), static_cast<unsigned>(v[2]),
        v[3], v[4], v[5], v[6], v[7], v[8], v[9], v[10], static_cast<unsigned>(v[11]),
        static_cast<unsigned>(v[12]), v[13], v[14], v[15], v[16], v[17]);
}

template <typename T>
static int heli_typed(const char* suf, hipStream_t st)
{
    const std::string s = suf;
    const HelicalInterpolationParam p = heli_geom(load("heli_geom", 18));
    Dev<T> in(load("heli_proj", static_cast<size_t>(p.na) * p.nt * p.ns));
    Dev<T> out(static_cast<size_t>(p.nz) * p.view_per_rot * p.ns);
    CHECK(HelicalInterpolation360<T>(p).calculate_on_gpu(in.p, out.p, st));
    save("heli360_" + s, out.get());
    CHECK(HelicalInterpolation180<T>(p).calculate_on_gpu(in.p, out.p, st));
    save("heli180_" + s, out.get());
    {   // the same through host pointers
        std::vector<T> hin = host_in<T>("heli_proj", static_cast<size_t>(p.na) * p.nt * p.ns);
        std::vector<T> hout(static_cast<size_t>(p.nz) * p.view_per_rot * p.ns);
        CHECK(HelicalInterpolation360<T>(p).calculate_on_cpu(hin.data(), hout.data()));
        save("heli360_host_" + s, hout);
        CHECK(HelicalInterpolation180<T>(p).calculate_on_cpu(hin.data(), hout.data()));
        save("heli180_host_" + s, hout);
    }
    // a slice one view too close to the scan start: the reference's exception
    HelicalInterpolationParam q = p;
    q.z0 = p.couch_begin + (p.view_per_rot - 1) * p.couch_mov;
    bool threw = false;
    try {
        HelicalInterpolation360<T>(q).calculate_on_gpu(in.p, out.p, st);
    } catch (const std::runtime_error& e) {
        threw = std::strcmp(e.what(), "View not enough for this slice!") == 0;
    }
    CHECK(threw);
    return 0;
}

static int gpu()
{
    hipStream_t st;
    CHECK(hipStreamCreate(&st) == hipSuccess);
    if (gpu_typed<float>("f32", st) || gpu_typed<double>("f64", st)) return 1;
    if (heli_typed<float>("f32", st) || heli_typed<double>("f64", st)) return 1;
    if (host_typed<float>("f32") || host_typed<double>("f64")) return 1;
    CHECK(hipStreamDestroy(st) == hipSuccess);
    std::puts("gpu ok");
    return 0;
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s cpu|gpu <dir>\n", argv[0]);
        return 2;
    }
    g_dir = argv[2];
    return std::strcmp(argv[1], "cpu") == 0 ? cpu() : gpu();
}
