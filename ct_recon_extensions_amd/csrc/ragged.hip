// Ragged batches of fan-beam backprojections with arbitrary view angles
// (bp_fan_2d_angle, src/bp_fan_2d_angle.cc:47-91) in one persistent launch.
//
// Every problem of the batch has its own image size, view count, detector
// and source geometry.  The work is cut into (problem, pixel tile) items,
// ordered by cost (views x tile pixels, exact-mode tiles weighted up), and a
// grid of exactly as many workgroups as the chip holds at once pulls items
// from a global counter: longest-first list scheduling, so the tail is one
// small item long.  Each item runs the shared tile engine (bp_engine.hpp)
// with the problem's own span width and series/exact policy, picked per item
// by a uniform branch; the tile shape and the views per LDS round are
// launch-wide and come from the autotuner.
//
// The counter is cleared by a one-thread kernel launched just before the
// persistent kernel (no memset node), so the pair can be captured into a
// HIP graph and replayed; a plan owns that graph.
#include "ctr_common.hpp"
#include "bp_engine.hpp"
#include "fan_geom.hpp"
#include "tile_common.hpp"

#include <algorithm>
#include <cmath>
#include <exception>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

namespace ctr {

// One problem as the persistent kernel reads it (uniform loads).
struct RaggedDesc {
    FanProblem P;
    FanTables T;
    const float* proj;
    float* img;
    double scale;
    int tiles_x, tiles_y;
    int lmax, vchunk;
    int variant;  // 4 * exact + span class (0: 64, 1: 128, 2: 256 channels)
    int pad;
};

struct RaggedItem {
    int prob, tile;
};

__global__ void ragged_clear_kernel(unsigned int* counter) { counter[0] = 0u; }

template <int TW, int TH>
__device__ __forceinline__ void ragged_run_item(const RaggedDesc& D, int tile, unsigned char* smem)
{
#define CTR_RAGGED_CASE(V_, E_, LROW_)                                                         \
    case V_: {                                                                                 \
        BpLaunch<FanGeo<E_>> L{D.P, 1, D.tiles_x, D.tiles_y, D.lmax, D.vchunk, D.scale, D.proj, \
                               D.img};                                                         \
        bp_tile<FanGeo<E_>, TW, TH, 1, LROW_>(D.T, L, tile, 0, smem);                          \
        break;                                                                                 \
    }
    switch (D.variant) {
        CTR_RAGGED_CASE(0, false, 64)
        CTR_RAGGED_CASE(1, false, 128)
        CTR_RAGGED_CASE(2, false, 256)
        CTR_RAGGED_CASE(4, true, 64)
        CTR_RAGGED_CASE(5, true, 128)
        CTR_RAGGED_CASE(6, true, 256)
        default: break;
    }
#undef CTR_RAGGED_CASE
}

// Persistent kernel: every workgroup loops until the item counter passes
// nitems (each workgroup exits after at most one failed fetch, so the grid
// always drains).  stamps (optional): per-workgroup first and last
// wall-clock ticks, for the load-balance figure.
template <int TW, int TH>
__global__ __launch_bounds__((TileShape<TW, TH>::kBlock))
__attribute__((amdgpu_waves_per_eu(4, 4))) void ragged_fan_bp_kernel(
    const RaggedDesc* __restrict__ descs, const RaggedItem* __restrict__ items, int nitems,
    unsigned int* counter, unsigned long long* stamps)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_item;
    const unsigned long long t0 = wall_clock64();
    for (;;) {
        if (threadIdx.x == 0) s_item = static_cast<int>(atomicAdd(counter, 1u));
        __syncthreads();
        const int it = uniform_i(s_item);
        if (it >= nitems) break;
        const RaggedItem item = items[it];
        const int prob = uniform_i(item.prob), tile = uniform_i(item.tile);
        ragged_run_item<TW, TH>(descs[prob], tile, smem);
    }
    if (stamps && threadIdx.x == 0) {
        stamps[2 * blockIdx.x] = t0;
        stamps[2 * blockIdx.x + 1] = wall_clock64();
    }
}

namespace {

constexpr int kTileW[3] = {16, 8, 32}, kTileH[3] = {16, 8, 8};

// Staged span classes.  A tile whose span exceeds 256 channels is not
// admitted: pixel offsets from the tile anchor then exceed ~128 channels and
// their FP32 rounding (~2e-7 relative) would cost more than 1e-5 channel of
// position accuracy -- the rule the single-geometry launcher applies by
// switching to 8x8 tiles.
int span_class(int lmax)
{
    if (lmax <= 64) return 0;
    if (lmax <= 128) return 1;
    if (lmax <= 256) return 2;
    return -1;
}

template <int TW, int TH>
const void* ragged_kernel_ptr()
{
    return reinterpret_cast<const void*>(ragged_fan_bp_kernel<TW, TH>);
}

const void* kernel_of(int tile)
{
    switch (tile) {
        case 1: return ragged_kernel_ptr<8, 8>();
        case 2: return ragged_kernel_ptr<32, 8>();
        default: return ragged_kernel_ptr<16, 16>();
    }
}

}  // namespace

}  // namespace ctr

// Opaque plan of the C ABI.
struct ctr_ragged_plan {
    int device = 0;
    int tile = 0, vchunk = 0;
    int nprob = 0, nitems = 0;
    int grid = 0, block = 0;
    size_t lds = 0;
    double pixel_views = 0;
    ctr::RaggedDesc* d_descs = nullptr;
    ctr::RaggedItem* d_items = nullptr;
    unsigned int* d_counter = nullptr;
    unsigned long long* d_stamps = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // Runs of a plan share its work counter.  Each run waits for this event
    // (the end of the previous run) on its own stream and records it again,
    // so runs from different streams or host threads execute one after the
    // other instead of clearing the counter under each other; `mu` makes the
    // wait / enqueue / record sequence and the first graph capture atomic.
    hipEvent_t done = nullptr;
    std::mutex mu;
    ~ctr_ragged_plan()
    {
        if (done) (void)hipEventDestroy(done);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (d_descs) (void)hipFree(d_descs);
        if (d_items) (void)hipFree(d_items);
        if (d_counter) (void)hipFree(d_counter);
        if (d_stamps) (void)hipFree(d_stamps);
    }
};

namespace ctr {
namespace {

int enqueue(ctr_ragged_plan* p, hipStream_t st, unsigned long long* stamps)
{
    hipLaunchKernelGGL(ragged_clear_kernel, dim3(1), dim3(1), 0, st, p->d_counter);
    int rc = check_launch("ragged_clear_kernel");
    if (rc) return rc;
    void* args[] = {&p->d_descs, &p->d_items, &p->nitems, &p->d_counter, &stamps};
    CTR_HIP(hipLaunchKernel(kernel_of(p->tile), dim3(p->grid), dim3(p->block), args, p->lds, st));
    return check_launch("ragged_fan_bp_kernel");
}

}  // namespace
}  // namespace ctr

extern "C" {

int ctr_ragged_fan_bp_plan(const ctr_ragged_problem* probs, int nprob, int tile, int vchunk,
                           ctr_ragged_plan** out)
{
    using namespace ctr;
    try {
        CTR_REQUIRE(out, CTR_ERR_ARG, "ragged fan BP: null plan pointer");
        *out = nullptr;
        CTR_REQUIRE(probs && nprob > 0, CTR_ERR_ARG, "ragged fan BP: no problems");
        CTR_REQUIRE(tile >= 0 && tile <= 2, CTR_ERR_ARG,
                    "ragged fan BP: tile must be 0 (16x16), 1 (8x8) or 2 (32x8)");
        CTR_REQUIRE(vchunk >= 1 && vchunk <= kAnchorChunk, CTR_ERR_ARG,
                    "ragged fan BP: view chunk must be in [1, 64]");
        const int TW = kTileW[tile], TH = kTileH[tile];
        const int block = std::min(256, TW * TH);
        // LDS per workgroup such that four waves per SIMD stay resident.
        const int wg_per_cu = 16 / (block / 64);
        const size_t lds_budget = 160 * 1024 / wg_per_cu;
        const size_t anchor_bytes = kAnchorChunk * sizeof(FanAnchor);

        std::vector<RaggedDesc> descs(nprob);
        std::vector<double> item_cost(nprob);
        size_t lds = anchor_bytes + 2 * 64 * sizeof(float);
        double pv = 0;
        long long nitems = 0;
        for (int i = 0; i < nprob; ++i) {
            const ctr_ragged_problem& q = probs[i];
            CTR_REQUIRE(q.xpos && q.ypos && q.sincostbl && q.proj && q.img, CTR_ERR_ARG,
                        "ragged fan BP: problem " + std::to_string(i) + " has a null pointer");
            RaggedDesc& D = descs[i];
            int lmax = 0;
            bool exact = false;
            const int rc = fan_problem(q.g, TW, TH, D.P, lmax, exact);
            if (rc) {
                set_error("ragged fan BP: problem " + std::to_string(i) + ": " + get_error());
                return rc;
            }
            const int cls = span_class(lmax);
            CTR_REQUIRE(cls >= 0, CTR_ERR_SHAPE,
                        "ragged fan BP: problem " + std::to_string(i) +
                            ": a pixel tile spans more than 256 channels; use a smaller tile");
            const int lrow = 64 << cls;
            int vc = static_cast<int>((lds_budget - anchor_bytes) / (2 * lrow * sizeof(float)));
            vc = std::max(1, std::min(vchunk, vc));
            D.T = FanTables{q.xpos, q.ypos, q.sincostbl};
            D.proj = q.proj;
            D.img = q.img;
            D.scale = q.scale;
            D.tiles_x = div_up(q.g.nx, TW);
            D.tiles_y = div_up(q.g.ny, TH);
            D.lmax = lmax;
            D.vchunk = vc;
            D.variant = (exact ? 4 : 0) + cls;
            lds = std::max(lds, anchor_bytes + 2 * static_cast<size_t>(vc) * lrow * sizeof(float));
            // exact tiles evaluate atanf per pixel-view: roughly 1.6x the work
            item_cost[i] = static_cast<double>(q.g.na) * TW * TH * (exact ? 1.6 : 1.0);
            pv += static_cast<double>(q.g.na) * q.g.nx * q.g.ny;
            nitems += static_cast<long long>(D.tiles_x) * D.tiles_y;
        }
        CTR_REQUIRE(nitems < (1LL << 31), CTR_ERR_SHAPE, "ragged fan BP: too many tiles");
        // Longest-first: problems by decreasing tile cost, tiles in order.
        std::vector<int> order(nprob);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return item_cost[a] > item_cost[b]; });
        std::vector<RaggedItem> items;
        items.reserve(static_cast<size_t>(nitems));
        for (int i : order)
            for (int t = 0; t < descs[i].tiles_x * descs[i].tiles_y; ++t) items.push_back({i, t});

        std::unique_ptr<ctr_ragged_plan> p(new ctr_ragged_plan);
        CTR_HIP(hipGetDevice(&p->device));
        p->tile = tile;
        p->vchunk = vchunk;
        p->nprob = nprob;
        p->nitems = static_cast<int>(items.size());
        p->block = block;
        p->lds = lds;
        p->pixel_views = pv;
        const void* kern = kernel_of(tile);
        if (lds > 64 * 1024)
            CTR_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
        int cus = 0, per_cu = 0;
        CTR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device));
        CTR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds));
        CTR_REQUIRE(per_cu >= 1, CTR_ERR_SHAPE, "ragged fan BP: kernel does not fit on a CU");
        p->grid = std::max(1, std::min(p->nitems, cus * per_cu));
        CTR_HIP(hipEventCreateWithFlags(&p->done, hipEventDisableTiming));
        CTR_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_descs), nprob * sizeof(RaggedDesc)));
        CTR_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_items), items.size() * sizeof(RaggedItem)));
        CTR_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_counter), 64));
        CTR_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_stamps),
                          2 * sizeof(unsigned long long) * p->grid));
        CTR_HIP(hipMemcpy(p->d_descs, descs.data(), nprob * sizeof(RaggedDesc),
                          hipMemcpyHostToDevice));
        CTR_HIP(hipMemcpy(p->d_items, items.data(), items.size() * sizeof(RaggedItem),
                          hipMemcpyHostToDevice));
        *out = p.release();
        return CTR_OK;
    } catch (const std::exception& e) {
        ctr::set_error(std::string("ragged fan BP: ") + e.what());
        return CTR_ERR_ALLOC;
    }
}

int ctr_ragged_fan_bp_run(ctr_ragged_plan* plan, int use_graph, void* stream_)
{
    using namespace ctr;
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    CTR_REQUIRE(plan, CTR_ERR_ARG, "ragged fan BP: null plan");
    std::lock_guard<std::mutex> lock(plan->mu);
    CTR_HIP(hipStreamWaitEvent(stream, plan->done, 0));
    if (!use_graph) {
        const int rc = enqueue(plan, stream, nullptr);
        if (rc) return rc;
        CTR_HIP(hipEventRecord(plan->done, stream));
        return CTR_OK;
    }
    if (!plan->exec) {
        // Capture on a private stream (the caller's may be the legacy null
        // stream, which cannot be captured); replay on the caller's.
        hipStream_t cap = nullptr;
        CTR_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        int rc = CTR_OK;
        if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            set_error("ragged fan BP: hipStreamBeginCapture failed");
            rc = CTR_ERR_HIP;
        } else {
            rc = enqueue(plan, cap, nullptr);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(cap, &g);
            if (rc == CTR_OK && e != hipSuccess) {
                set_error(std::string("ragged fan BP: hipStreamEndCapture: ") + hipGetErrorString(e));
                rc = CTR_ERR_HIP;
            }
            if (rc == CTR_OK) {
                plan->graph = g;
                if (hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0) != hipSuccess) {
                    set_error("ragged fan BP: hipGraphInstantiate failed");
                    plan->exec = nullptr;
                    rc = CTR_ERR_HIP;
                }
            } else if (g) {
                (void)hipGraphDestroy(g);
            }
        }
        (void)hipStreamDestroy(cap);
        if (rc) return rc;
    }
    CTR_HIP(hipGraphLaunch(plan->exec, stream));
    CTR_HIP(hipEventRecord(plan->done, stream));
    return CTR_OK;
}

int ctr_ragged_fan_bp_info(const ctr_ragged_plan* plan, int* nitems, int* grid, int* block,
                           int* lds_bytes, double* pixel_views)
{
    using namespace ctr;
    CTR_REQUIRE(plan, CTR_ERR_ARG, "ragged fan BP: null plan");
    if (nitems) *nitems = plan->nitems;
    if (grid) *grid = plan->grid;
    if (block) *block = plan->block;
    if (lds_bytes) *lds_bytes = static_cast<int>(plan->lds);
    if (pixel_views) *pixel_views = plan->pixel_views;
    return CTR_OK;
}

int ctr_ragged_fan_bp_balance(ctr_ragged_plan* plan, void* stream_, double* busy_fraction)
{
    using namespace ctr;
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    CTR_REQUIRE(plan && busy_fraction, CTR_ERR_ARG, "ragged fan BP: null argument");
    std::lock_guard<std::mutex> lock(plan->mu);
    CTR_HIP(hipStreamWaitEvent(stream, plan->done, 0));
    int rc = enqueue(plan, stream, plan->d_stamps);
    if (rc) return rc;
    CTR_HIP(hipEventRecord(plan->done, stream));
    std::vector<unsigned long long> st(2 * static_cast<size_t>(plan->grid));
    CTR_HIP(hipMemcpyAsync(st.data(), plan->d_stamps, st.size() * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, stream));
    CTR_HIP(hipStreamSynchronize(stream));
    unsigned long long lo = ~0ull, hi = 0;
    double busy = 0;
    for (int i = 0; i < plan->grid; ++i) {
        lo = std::min(lo, st[2 * i]);
        hi = std::max(hi, st[2 * i + 1]);
        busy += static_cast<double>(st[2 * i + 1] - st[2 * i]);
    }
    *busy_fraction = hi > lo ? busy / (static_cast<double>(hi - lo) * plan->grid) : 1.0;
    return CTR_OK;
}

int ctr_ragged_fan_bp_destroy(ctr_ragged_plan* plan)
{
    delete plan;
    return CTR_OK;
}

int ctr_ragged_fan_bp_autotune(const ctr_ragged_problem* probs, int nprob, int reps,
                               void* stream_, int* best_tile, int* best_vchunk, float* ms_table)
{
    using namespace ctr;
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    CTR_REQUIRE(best_tile && best_vchunk && reps >= 1, CTR_ERR_ARG,
                "ragged fan BP autotune: bad arguments");
    static const int chunks[4] = {8, 16, 32, 64};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CTR_HIP(hipEventCreate(&e0));
    CTR_HIP(hipEventCreate(&e1));
    float best = -1.0f;
    int rc = CTR_OK;
    for (int t = 0; t < 3 && rc == CTR_OK; ++t)
        for (int c = 0; c < 4 && rc == CTR_OK; ++c) {
            float ms = -1.0f;
            ctr_ragged_plan* p = nullptr;
            if (ctr_ragged_fan_bp_plan(probs, nprob, t, chunks[c], &p) == CTR_OK) {
                rc = enqueue(p, stream, nullptr);  // warm-up
                if (rc == CTR_OK && hipEventRecord(e0, stream) != hipSuccess) rc = CTR_ERR_HIP;
                for (int r = 0; r < reps && rc == CTR_OK; ++r) rc = enqueue(p, stream, nullptr);
                if (rc == CTR_OK && (hipEventRecord(e1, stream) != hipSuccess ||
                                     hipEventSynchronize(e1) != hipSuccess ||
                                     hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                    rc = CTR_ERR_HIP;
                ms /= reps;
                delete p;
            }
            // a configuration the geometry does not admit stays at -1
            if (ms_table) ms_table[4 * t + c] = rc == CTR_OK ? ms : -1.0f;
            if (rc == CTR_OK && ms > 0 && (best < 0 || ms < best)) {
                best = ms;
                *best_tile = t;
                *best_vchunk = chunks[c];
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    CTR_REQUIRE(best > 0, CTR_ERR_SHAPE, "ragged fan BP autotune: no configuration fits");
    return CTR_OK;
}

}  // extern "C"
