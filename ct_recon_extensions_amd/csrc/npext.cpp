// The reference's numpy C ABI (include/numpy_ctext.h) over the MI355X
// library: integer handles to operator objects that own their device
// tables and staging buffers; each run is H2D copy -> float64 kernel ->
// D2H copy on the library's stream, synchronised before returning.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "include/ctr_api.h"
#include "include/numpy_ctext.h"

namespace {

struct DeviceBuffer {
    void* p = nullptr;
    ~DeviceBuffer()
    {
        if (p) (void)hipFree(p);
    }
    bool alloc(size_t bytes) { return p || hipMalloc(&p, bytes) == hipSuccess; }
    template <typename T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
};

struct State {
    std::mutex mu;
    int device = -1;
    hipStream_t stream = nullptr;
    int next = 0;
};

// Heap-allocated and never destroyed: device buffers must not be freed by
// static destructors after the HIP runtime has shut down.
State& state()
{
    static State* s = new State;
    return *s;
}

bool upload(DeviceBuffer& d, const void* h, size_t bytes, hipStream_t st)
{
    return d.alloc(bytes) && hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, st) == hipSuccess;
}

// One operator object per handle.
struct Op {
    virtual ~Op() = default;
    virtual bool run(const double* in, double* out, hipStream_t st) = 0;
};

struct FanWeightingOp : Op {
    ctr_geom g{};
    DeviceBuffer din, dout;
    bool run(const double* in, double* out, hipStream_t st) override
    {
        const size_t n = static_cast<size_t>(g.ns) * g.nrow * sizeof(double);
        if (!upload(din, in, n, st) || !dout.alloc(n)) return false;
        if (ctr_fan_weighting(&g, CTR_F64, din.p, dout.p, g.nrow, st) != CTR_OK) return false;
        return hipMemcpyAsync(out, dout.p, n, hipMemcpyDeviceToHost, st) == hipSuccess &&
               hipStreamSynchronize(st) == hipSuccess;
    }
};

struct RampFilterOp : Op {
    ctr_geom g{};
    DeviceBuffer taps, din, dout;
    bool run(const double* in, double* out, hipStream_t st) override
    {
        if (!taps.p) {
            std::vector<double> h(2 * g.ns + 1);
            if (ctr_ramp_taps(&g, h.data()) != CTR_OK) return false;
            if (!upload(taps, h.data(), h.size() * sizeof(double), st)) return false;
        }
        const size_t n = static_cast<size_t>(g.ns) * g.nrow * sizeof(double);
        if (!upload(din, in, n, st) || !dout.alloc(n)) return false;
        if (ctr_ramp(&g, CTR_F64, din.p, dout.p, g.nrow, taps.as<double>(), 0, st) != CTR_OK)
            return false;
        return hipMemcpyAsync(out, dout.p, n, hipMemcpyDeviceToHost, st) == hipSuccess &&
               hipStreamSynchronize(st) == hipSuccess;
    }
};

struct FanBpAngleOp : Op {
    ctr_geom g{};
    std::vector<double> angles;
    DeviceBuffer xpos, ypos, sc, din, dout;
    bool run(const double* in, double* out, hipStream_t st) override
    {
        if (!sc.p) {
            std::vector<double> x(g.nx), y(g.ny), s(2 * g.na);
            if (ctr_fan_angle_tables(&g, angles.data(), x.data(), y.data(), s.data()) != CTR_OK)
                return false;
            if (!upload(xpos, x.data(), x.size() * 8, st) || !upload(ypos, y.data(), y.size() * 8, st) ||
                !upload(sc, s.data(), s.size() * 8, st))
                return false;
        }
        const size_t nin = static_cast<size_t>(g.ns) * g.na * sizeof(double);
        const size_t nout = static_cast<size_t>(g.nx) * g.ny * sizeof(double);
        if (!upload(din, in, nin, st) || !dout.alloc(nout)) return false;
        if (ctr_fan_bp(&g, CTR_F64, din.p, dout.p, 1, xpos.as<double>(), ypos.as<double>(),
                       sc.as<double>(), 1.0, st) != CTR_OK)
            return false;
        return hipMemcpyAsync(out, dout.p, nout, hipMemcpyDeviceToHost, st) == hipSuccess &&
               hipStreamSynchronize(st) == hipSuccess;
    }
};

std::map<int, std::unique_ptr<Op>>& ops()
{
    static auto* m = new std::map<int, std::unique_ptr<Op>>;
    return *m;
}

int add(std::unique_ptr<Op> op)
{
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    const int h = ++s.next;
    ops()[h] = std::move(op);
    return h;
}

bool run(int handle, const double* in, double* out)
{
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.device < 0 || !in || !out) return false;
    auto it = ops().find(handle);
    if (it == ops().end()) return false;
    if (hipSetDevice(s.device) != hipSuccess) return false;
    return it->second->run(in, out, s.stream);
}

bool erase(int handle)
{
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    return ops().erase(handle) == 1;
}

}  // namespace

extern "C" {

// device < 0 is the reference's CPU mode (numpy_ext/common.cu:29-49): the
// arrays are host memory either way, and this library computes them on GPU 0.
bool set_device(const int device_in)
{
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    const int device = device_in < 0 ? 0 : device_in;
    if (hipSetDevice(device) != hipSuccess) return false;
    if (s.stream && s.device != device) {
        (void)hipStreamDestroy(s.stream);
        s.stream = nullptr;
    }
    if (!s.stream && hipStreamCreate(&s.stream) != hipSuccess) return false;
    s.device = device;
    return true;
}

bool clear()
{
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    ops().clear();
    if (s.stream) (void)hipStreamDestroy(s.stream);
    s.stream = nullptr;
    s.device = -1;
    return true;
}

int fan_weighting_create(unsigned int ns, unsigned int nrow, double ds, double offset, double dso,
                         double dsd, int type)
{
    if (ns == 0 || nrow == 0 || (type != 1 && type != 2)) return 0;
    auto op = std::make_unique<FanWeightingOp>();
    op->g.ns = ns;
    op->g.nrow = nrow;
    op->g.ds = ds;
    op->g.offset_s = offset;
    op->g.dso = dso;
    op->g.dsd = dsd;
    op->g.type = type;
    return add(std::move(op));
}

bool fan_weighting_run(int handle, double* in, double* out) { return run(handle, in, out); }
bool fan_weighting_destroy(int handle) { return erase(handle); }

int ramp_filter_create(unsigned int ns, unsigned int nrow, double ds, double dsd, int type)
{
    if (ns == 0 || nrow == 0 || type < 0 || type > 2) return 0;
    auto op = std::make_unique<RampFilterOp>();
    op->g.ns = ns;
    op->g.nrow = nrow;
    op->g.ds = ds;
    op->g.dsd = dsd;
    op->g.type = type;
    return add(std::move(op));
}

bool ramp_filter_run(int handle, double* in, double* out) { return run(handle, in, out); }
bool ramp_filter_destroy(int handle) { return erase(handle); }

int fan_bp_2d_angle_create(double* angles, unsigned int ns, unsigned int na, double ds,
                           double offset_s, unsigned int nx, unsigned int ny, double dx,
                           double dy, double offset_x, double offset_y, double dso, double dsd,
                           double fov)
{
    if (!angles || ns < 2 || na == 0 || nx == 0 || ny == 0) return 0;
    auto op = std::make_unique<FanBpAngleOp>();
    op->g.ns = ns;
    op->g.na = na;
    op->g.ds = ds;
    op->g.offset_s = offset_s;
    op->g.nx = nx;
    op->g.ny = ny;
    op->g.dx = dx;
    op->g.dy = dy;
    op->g.offset_x = offset_x;
    op->g.offset_y = offset_y;
    op->g.dso = dso;
    op->g.dsd = dsd;
    op->g.fov = fov;
    op->angles.assign(angles, angles + na);
    return add(std::move(op));
}

bool fan_bp_2d_angle_run(int handle, double* in, double* out) { return run(handle, in, out); }
bool fan_bp_2d_angle_destroy(int handle) { return erase(handle); }

}  // extern "C"
