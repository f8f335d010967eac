#!/usr/bin/env python3
"""Benchmark of the fan-beam FBP hot path on one MI355X.

Workload (one "step"): a batch of 16 seeded-ellipse 512x512 slices, equiangular
fan, 720 views over 360 deg, 736 channels at 1.30 mm arc pitch, SID 595 mm,
SDD 1085.6 mm, pixel 0.7 mm; analytic ellipse line integrals plus Poisson noise
at 1e5 photons; the step is fan weighting fused into the Ram-Lak ramp filter,
then bp_fan_2d, all on inputs already resident in HBM.

Prints one JSON line: whole-batch throughput in Gpixel*view/s (one
pixel*view = one interpolate-accumulate of one pixel from one view), the
roofline of the dominant kernel (the BP) timed with HIP events on the stream
it runs on, the reference's own CPU path (oracle/_ref, else the C
restatement) timed on this host's cores on a bounded sample, the relative L2
error against the reference's float and FP64 CPU results, and the plain
per-slice GPU path (the reference's execution structure) for comparison.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks
PEAK_HBM_GBS = 8000.0
CU, LANES, CLK = 256, 64, 2.4e9
PEAK_VALU_SLOTS = CU * LANES * CLK  # FP32 VALU lane-issues per second (3.93e13)

# Algorithmic FP32 VALU issue slots of the fan BP, counting a quarter-rate
# transcendental (v_rcp_f32) as 4 slots.  Geometry, once per pixel*view and
# shared by every slice the launch carries: numerator 2, denominator 2, rcp 4,
# v = num*rcp 1, v^2 1, atan series 2, s = s_c + ... 2, weight 4, fraction 1
# (v_fract_f32; the int convert truncates, s >= 1 in the staged span), int
# convert 1, two tap weights 2, tap address 2 = 24 (the index clamp and the
# per-pixel edge test run only for tiles that reach past the detector).
# Per slice: two FMAs into the accumulator = 2.
FAN_GEOM_SLOTS = 24
FAN_SLICE_SLOTS = 2
# Per-slice model with no geometry sharing (atan2 + divide per pixel*view,
# the textbook count for one slice at a time).
UNBATCHED_FAN_SLOTS = 28

LSQ_SEED = 20261203
RAGGED_SEED = 20261204
PRECISION_SEED = 20261205

CFG = dict(nbatch=16, n=512, dx=0.7, na=720, ns=736, ds=1.30, dso=595.0, dsd=1085.6,
           i0=1e5, ellipses=10, seed=20261202)


def make_inputs(cfg):
    from ct_recon_extensions_amd import phantom as P
    orbit = 2 * math.pi / cfg["na"]
    radius = cfg["n"] * cfg["dx"] / 2
    sinos = np.empty((cfg["nbatch"], cfg["na"], cfg["ns"]), np.float32)
    for b in range(cfg["nbatch"]):
        ell = P.random_ellipses(cfg["ellipses"], cfg["seed"] + b, radius=radius)
        p = P.fan_sinogram(ell, cfg["na"], cfg["ns"], cfg["ds"], orbit, cfg["dso"], cfg["dsd"])
        sinos[b] = P.poisson_noise(p, cfg["i0"], cfg["seed"] + 1000 + b)
    return sinos


def kw_fan(cfg):
    return dict(img_shape=[cfg["n"], cfg["n"]], img_space=[cfg["dx"], cfg["dx"]],
                proj_shape=[cfg["na"], cfg["ns"]], channel_space=cfg["ds"],
                orbit=2 * math.pi / cfg["na"], dso=cfg["dso"], dsd=cfg["dsd"])


def kw_filter(cfg):
    return dict(proj_shape=[cfg["na"], cfg["ns"]], channel_space=cfg["ds"], dso=cfg["dso"],
                dsd=cfg["dsd"], tp="fan")


def time_events(fn, reps, torch):
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps  # ms


def cpu_baseline(cfg, sinos, threads, budget_s=30.0):
    """Reference CPU FBP (fan weighting -> ramp -> bp_fan_2d), one slice per
    thread, on `threads` host threads (ctypes releases the GIL).  Slice 0 is
    run both in float (the reference's float path) and in double (FP64
    oracle) for the error figures."""
    from oracle import oracle as O
    kind = "ref" if O.have_ref() else "oracle"
    g = O.geom(ns=cfg["ns"], na=cfg["na"], ds=cfg["ds"], orbit=2 * math.pi / cfg["na"],
               nx=cfg["n"], ny=cfg["n"], dx=cfg["dx"], dy=cfg["dx"], dso=cfg["dso"],
               dsd=cfg["dsd"], type=1, nrow=cfg["na"])
    tables = O.fan_prep(g, kind)
    taps = O.ramp_prep(g, kind)

    def fbp(x):
        return O.fan_bp(g, O.ramp(g, O.fanw(g, x, kind=kind), taps, kind=kind), tables, kind=kind)

    # single thread, one slice: the reference's own execution (a serial loop)
    t0 = time.perf_counter()
    fbp(sinos[0])
    single_s = time.perf_counter() - t0
    jobs = [sinos[0], sinos[0].astype(np.float64)]
    jobs += [sinos[b % len(sinos)] for b in range(1, threads - 1)]
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        outs = list(ex.map(fbp, jobs))
    wall = time.perf_counter() - t0
    pv = len(jobs) * cfg["n"] ** 2 * cfg["na"]
    return dict(kind="reference" if kind == "ref" else "port", wall_s=wall, slices=len(jobs),
                pixel_views=pv, ref_f32=outs[0], ref_f64=outs[1], single_s=single_s)


TRAFFIC_PROFILE = "profiles/fan_bp_16x512sq_720v.md"


def profiled_traffic():
    """HBM bytes per BP launch recorded by the counter runs of --profile-bp."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_PROFILE)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


def profile_kernels(torch, F, which):
    """One kernel family alone (for rocprofv3 --pmc): three launches on random
    inputs of the workload's shape, then a copy of a 256 MiB buffer that was
    evicted from L2 and MALL by writing 1 GiB elsewhere: its FETCH_SIZE /
    WRITE_SIZE against 256 MiB calibrate the counters."""
    cfg = CFG
    g = torch.Generator(device="cuda").manual_seed(0)
    if which == "lsq":
        from ct_recon_extensions_amd import recon
        scan = recon.ParallelScan()
        fp, fpt = recon.projector(scan)
        x = torch.randn((8, scan.n, scan.n), generator=g, device="cuda")
        for _ in range(3):
            img = fpt(fp(x))
    elif which == "rc":
        c = SVBP_CFG
        x = torch.randn((1, c["n"], c["n"]), generator=g, device="cuda")
        for _ in range(3):
            img = F.parallel_projection_2d(x, **kw_rc())
    elif which == "dd":
        c = SVBP_CFG
        x = torch.randn((8, c["n"], c["n"]), generator=g, device="cuda")
        y = torch.randn((8, c["na"], c["ns"]), generator=g, device="cuda")
        for _ in range(3):
            img = F.parallel_projection_2d(x, **kw_dd())
            img = F.parallel_projection_2d_grad(y, **kw_dd())
    elif which == "svbp":
        c = SVBP_CFG
        proj = torch.randn((1, c["na"], c["ns"]), generator=g, device="cuda")
        for _ in range(3):
            img = F.parallel_single_view_bp_2d(proj, **kw_svbp())
    elif which == "bpt":
        x = torch.randn((cfg["nbatch"], cfg["n"], cfg["n"]), generator=g, device="cuda")
        for det in (False, True):
            for _ in range(3):
                img = F.fan_backprojection_2d_grad(x, **kw_fan(cfg), deterministic=det)
    elif which == "ragged":
        from ct_recon_extensions_amd import ragged as R
        params, angles = R.random_fan_problems(64, seed=RAGGED_SEED)
        rb = R.RaggedFanBackprojection2DAngle(params, angles, tile=0, vchunk=32)
        rb.proj_flat.normal_(generator=g)
        for _ in range(3):
            rb(graph=False)
        img = rb.img_flat
    elif which == "helical":
        from ct_recon_extensions_amd.geometry import HelicalInterpolationParam
        p = HelicalInterpolationParam(**helical_scan(HELICAL_NZ, HELICAL_NA))
        x = torch.randn((p.na, p.nt, p.ns), generator=g, device="cuda")
        for mode in (360, 180):
            for _ in range(3):
                img = F.helical_interpolation(x, p, mode, check=False)
    elif which == "filter5":
        from ct_recon_extensions_amd import phantom as P
        x, kw, _ = P.shepp_logan_fan_batch(32, PRECISION_SEED, "cuda")
        for _ in range(3):
            img = F.fan_filter(x, **kw, impl="bf16x3")
    else:
        proj = torch.randn((cfg["nbatch"], cfg["na"], cfg["ns"]), generator=g, device="cuda")
        for _ in range(3):
            img = (F.fan_backprojection_2d(proj, **kw_fan(cfg)) if which == "bp"
                   else F.fan_filter(proj, **kw_filter(cfg)))
    torch.cuda.synchronize()
    src = torch.randn(64 * 1024 * 1024, generator=g, device="cuda")
    evict = torch.zeros(256 * 1024 * 1024, device="cuda")
    torch.cuda.synchronize()
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"profile": which, "checksum": float(img.double().abs().sum()),
                      "copy_bytes": src.numel() * 4, "evict_bytes": evict.numel() * 4}))


C1_SEED = 20261206
SVBP_SEED = 20261207
# Single-view parallel BP workload: one 512^2 slice, parallel beam, 720 views
# over 180 deg, 729 channels (ds = dx = dy = 1): 755 MB of output per launch.
SVBP_CFG = dict(n=512, na=720, ns=729)


def kw_svbp():
    c = SVBP_CFG
    return dict(img_shape=[c["n"], c["n"]], img_space=[1.0, 1.0], proj_shape=[c["na"], c["ns"]],
                channel_space=1.0, orbit=math.pi / c["na"])


RC_SEED = 20261208
# Ray-casting FP workload: the single-view BP's geometry with the FOV circle
# inscribed in the image (fov = 256), the reference's step 0.125.
PEAK_FP64_TFLOPS = 78.6
RC_FLOPS_PER_SAMPLE = 16  # bilinear lerp 14 (u, v, 1-u, 1-v, 6 mul, 3 add, sum) + 2 advances


def kw_rc():
    return dict(kw_svbp(), fov=256.0, method="raycasting")


def bench_raycasting(torch, F, reps, with_cpu=True):
    """Ray-casting parallel FP (src/fp_par_2d.cc:18-144) of one seeded random
    512^2 image: GPU time, samples/s and the FP64 rate of its bilinear
    samples, and the reference CPU path on one thread for the same slice."""
    from ct_recon_extensions_amd.geometry import ParallelProjection2DParam, fp_rc_tables
    c = SVBP_CFG
    n, na, ns = c["n"], c["na"], c["ns"]
    g = torch.Generator(device="cuda").manual_seed(RC_SEED)
    img = torch.randn((1, n, n), generator=g, device="cuda")
    out = F.parallel_projection_2d(img, **kw_rc())
    torch.cuda.synchronize()
    ms = time_events(lambda: F.parallel_projection_2d(img, **kw_rc()), reps, torch)
    p = ParallelProjection2DParam(ns, na, 1.0, math.pi / na, 0.0, 0.0, n, n, 1.0, 1.0, 0.0, 0.0,
                                  256.0)
    samples = int(fp_rc_tables(p)[2].astype(np.int64).sum())
    tflops = samples * RC_FLOPS_PER_SAMPLE / (ms * 1e-3) / 1e12
    res = {"workload": "1 x 512^2, parallel 720 views/180deg, 729 ch, ds=dx=1, fov 256, step "
                       "0.125; seeded N(0,1) image", "seed": RC_SEED, "gpu_ms": round(ms, 4),
           "samples": samples, "gsamples_per_s": round(samples / (ms * 1e-3) / 1e9, 2),
           "fp64_tflops": round(tflops, 3), "fp64_frac": round(tflops / PEAK_FP64_TFLOPS, 4),
           "valu_issue": _valu_issue(RC_VALU, ms, "profiles/fp_raycasting_512sq_720v_729ch.md")}
    if with_cpu:
        from oracle import oracle as O
        kind = "ref" if O.have_ref() else "oracle"
        geo = O.geom(ns=ns, na=na, ds=1.0, orbit=math.pi / na, nx=n, ny=n, dx=1.0, dy=1.0,
                     fov=256.0)
        t = O.fp_rc_prep(geo, kind=kind)
        host = img[0].cpu().numpy()
        t0 = time.perf_counter()
        ref = O.fp_rc(geo, host, t, kind=kind)
        res.update(cpu_kind="reference" if kind == "ref" else "port",
                   cpu_1thread_ms=round((time.perf_counter() - t0) * 1e3, 1),
                   bit_identical=bool(np.array_equal(out[0].cpu().numpy(), ref)))
    return res


DD_SEED = 20261210


def kw_dd():
    return dict(kw_svbp(), method="disdriven")


# Executed VALU wave-instructions per launch of the distance-driven kernels
# on this workload (SQ_INSTS_VALU, profiles/fp_dd_8x512sq_720v_729ch.md).
# The reference's float/double mix and operation order are kept for bit
# identity, so issue, not HBM, bounds both kernels.
DD_VALU_FP = 1.21e9
DD_VALU_GRAD = 9.97e8


def _dd_roofline(kernel, valu_insts, alg_bytes, ms):
    slots = valu_insts * 64
    return {"kernel": kernel, "bound": "valu",
            "achieved": round(slots / (ms * 1e-3) / 1e12, 3),
            "peak": round(PEAK_VALU_SLOTS / 1e12, 3), "unit": "Tslot/s (FP32 VALU lane-issues)",
            "frac": round(slots / (ms * 1e-3) / PEAK_VALU_SLOTS, 4),
            "model": f"executed VALU instructions per launch ({valu_insts:.3g}, rocprof "
                     "SQ_INSTS_VALU) x 64 lanes: issue utilisation, not an algorithmic count",
            "hbm": {"achieved": round(alg_bytes / (ms * 1e-3) / 1e9, 1), "peak": PEAK_HBM_GBS,
                    "unit": "GB/s",
                    "frac": round(alg_bytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                    "algorithmic_bytes": alg_bytes}}


def _valu_issue(valu_insts, ms, source):
    """Issue utilisation from a kernel's profiled instruction count: executed
    VALU wave-instructions x 64 lanes over the measured time."""
    slots = valu_insts * 64
    return {"executed_valu_insts": valu_insts, "source": source,
            "achieved": round(slots / (ms * 1e-3) / 1e12, 3),
            "peak": round(PEAK_VALU_SLOTS / 1e12, 3), "unit": "Tslot/s (FP32 VALU lane-issues)",
            "frac": round(slots / (ms * 1e-3) / PEAK_VALU_SLOTS, 4)}


# Executed VALU wave-instructions per launch (rocprof SQ_INSTS_VALU) of the
# ray-casting FP and the helical kernels on their workloads.
RC_VALU = 1.37e9
HELI_VALU = {360: 1.39e8, 180: 3.78e8}


def bench_fp_dd(torch, F, reps, with_cpu=True):
    """Distance-driven parallel FP and the gradient the reference defines for
    it (src/fp_par_2d.cc:396-655) on the least-squares workload's geometry
    (8 x 512^2, 720 views over 180 deg, 729 channels, ds = dx = 1): GPU time
    per op call for the 8-slice batch, against VALU issue (the executed
    instruction count of the kernel's profile) with the HBM fraction of its
    algorithmic bytes beside it (image and sinogram once each, plus the
    tables), and the reference CPU path on one thread for slice 0
    (bit-identity checked there)."""
    c = SVBP_CFG
    n, na, ns, nb = c["n"], c["na"], c["ns"], 8
    g = torch.Generator(device="cuda").manual_seed(DD_SEED)
    img = torch.randn((nb, n, n), generator=g, device="cuda")
    proj = torch.randn((nb, na, ns), generator=g, device="cuda")
    out = F.parallel_projection_2d(img, **kw_dd())
    gout = F.parallel_projection_2d_grad(proj, **kw_dd())
    torch.cuda.synchronize()
    ms_fp = time_events(lambda: F.parallel_projection_2d(img, **kw_dd()), reps, torch)
    ms_gr = time_events(lambda: F.parallel_projection_2d_grad(proj, **kw_dd()), reps, torch)
    alg_fp = nb * (n * n + na * ns) * 4 + na * ns * 3 * 8
    alg_gr = nb * (n * n + na * ns) * 4 + na * n * 2 * 8
    res = {"workload": "8 x 512^2, parallel 720 views/180deg, 729 ch, ds=dx=1, distance-driven; "
                       "seeded N(0,1) image and sinogram", "seed": DD_SEED,
           "fp_ms": round(ms_fp, 4), "grad_ms": round(ms_gr, 4),
           "roofline_fp": _dd_roofline("fp_dd_kernel", DD_VALU_FP, alg_fp, ms_fp),
           "roofline_grad": _dd_roofline("fp_dd_grad_kernel", DD_VALU_GRAD, alg_gr, ms_gr)}
    if with_cpu:
        from oracle import oracle as O
        kind = "ref" if O.have_ref() else "oracle"
        geo = O.geom(ns=ns, na=na, ds=1.0, orbit=math.pi / na, nx=n, ny=n, dx=1.0, dy=1.0)
        t = O.fp_dd_prep(geo, kind=kind)
        tg = O.fp_dd_grad_prep(geo, kind=kind)
        x0, y0 = img[0].cpu().numpy(), proj[0].cpu().numpy()
        t0 = time.perf_counter()
        ref = O.fp_dd(geo, x0, t, kind=kind)
        t1 = time.perf_counter()
        gref = O.fp_dd_grad(geo, y0, tg, kind=kind)
        t2 = time.perf_counter()
        res.update(cpu_kind="reference" if kind == "ref" else "port",
                   cpu_1thread_fp_ms_per_slice=round((t1 - t0) * 1e3, 1),
                   cpu_1thread_grad_ms_per_slice=round((t2 - t1) * 1e3, 1),
                   bit_identical_slice0=bool(np.array_equal(out[0].cpu().numpy(), ref) and
                                             np.array_equal(gout[0].cpu().numpy(), gref)))
    return res


def bench_svbp(torch, F, reps, with_cpu=True):
    """Single-view parallel BP (src/bp_par_2d_sv.cc:54-92) of one seeded
    random sinogram: GPU time per launch against the HBM roofline (the output,
    nx*ny*na floats, dominates the algorithmic bytes), and the reference CPU
    path on one thread for the same slice."""
    c = SVBP_CFG
    g = torch.Generator(device="cuda").manual_seed(SVBP_SEED)
    proj = torch.randn((1, c["na"], c["ns"]), generator=g, device="cuda")
    out = F.parallel_single_view_bp_2d(proj, **kw_svbp())
    torch.cuda.synchronize()
    ms = time_events(lambda: F.parallel_single_view_bp_2d(proj, **kw_svbp()), reps, torch)
    n, na, ns = c["n"], c["na"], c["ns"]
    alg = n * n * na * 4 + na * ns * 4 + 2 * n * na * 8
    res = {"workload": "1 x 512^2, parallel 720 views/180deg, 729 ch, ds=dx=1; seeded N(0,1) "
                       "sinogram", "seed": SVBP_SEED, "gpu_ms": round(ms, 4),
           "algorithmic_bytes": alg, "gb_s": round(alg / (ms * 1e-3) / 1e9, 1),
           "hbm_frac": round(alg / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    if with_cpu:
        from oracle import oracle as O
        kind = "ref" if O.have_ref() else "oracle"
        geo = O.geom(ns=ns, na=na, ds=1.0, orbit=math.pi / na, nx=n, ny=n, dx=1.0, dy=1.0)
        host = proj[0].cpu().numpy()
        t = O.par_svbp_prep(geo, kind=kind)
        t0 = time.perf_counter()
        ref = O.par_svbp(geo, host, t, kind=kind)
        cpu_s = time.perf_counter() - t0
        res.update(cpu_kind="reference" if kind == "ref" else "port", cpu_1thread_ms=round(
            cpu_s * 1e3, 1), bit_identical=bool(np.array_equal(out[0].cpu().numpy(), ref)))
    return res


def bench_c1(torch, F, threads, with_cpu=True):
    """Plain parallel-beam FBP of a 256^2 seeded 10-ellipse phantom: ray-driven
    FP to 360 views over 180 deg, 367 channels (ds = dx = 1), Ram-Lak, then
    pixel-driven BP -- on the GPU, and the reference CPU path on one thread
    and on `threads` threads (one copy per thread)."""
    from ct_recon_extensions_amd import phantom as P
    from oracle import oracle as O
    n, na, ns = 256, 360, 367
    img = P.rasterize(P.random_ellipses(10, C1_SEED, radius=n / 2), n, n, 1.0, 1.0)
    img = img.astype(np.float32)
    kw = dict(img_shape=[n, n], img_space=[1.0, 1.0], proj_shape=[na, ns], channel_space=1.0,
              orbit=math.pi / na)
    x = torch.from_numpy(img)[None].cuda()

    def gpu():
        s = F.parallel_projection_2d(x, **kw)
        f = F.ramp_filter(s, ns, na, 1.0, type="par")
        return F.parallel_backprojection_2d(f, **kw)

    out = gpu()
    torch.cuda.synchronize()
    ms = time_events(gpu, 20, torch)
    res = {"workload": "256^2 phantom, FP 360 views/180deg, 367 ch, Ram-Lak, parallel BP",
           "seed": C1_SEED, "gpu_ms": round(ms, 4)}
    if with_cpu:
        kind = "ref" if O.have_ref() else "oracle"
        g = O.geom(ns=ns, na=na, ds=1.0, orbit=math.pi / na, nx=n, ny=n, dx=1.0, dy=1.0,
                   type=0, nrow=na)
        tf, tp, taps = O.fp_prep(g, kind), O.par_prep(g, kind), O.ramp_prep(g, kind)

        def cpu(v):
            return O.par_bp(g, O.ramp(g, O.fp(g, v, tf, kind=kind), taps, kind=kind), tp,
                            kind=kind)

        t0 = time.perf_counter()
        ref = cpu(img)
        one = time.perf_counter() - t0
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            list(ex.map(cpu, [img] * threads))
        allc = time.perf_counter() - t0
        got = out[0].double().cpu().numpy()
        res.update({"cpu_kind": "reference" if kind == "ref" else "port",
                    "cpu_1thread_ms": round(one * 1e3, 2),
                    "cpu_all_threads_ms_per_slice": round(allc * 1e3 / threads, 2),
                    "cpu_threads": threads,
                    "rel_l2_vs_reference_f32": float(np.linalg.norm(got - ref) /
                                                     np.linalg.norm(ref))})
    return res


def bench_lsq(torch, steps):
    """Least squares min 0.5||FP x - y||^2 on 8 seeded-ellipse 512^2 slices,
    parallel beam, 720 views over 180 deg, 729 channels: `steps` gradient
    steps x <- x - FP^T(FP x - y)/||FP||^2 through autograd, timed."""
    from ct_recon_extensions_amd import phantom as P
    from ct_recon_extensions_amd import recon
    scan = recon.ParallelScan()
    fp, fpt = recon.projector(scan)
    truth = np.stack([P.rasterize(P.random_ellipses(10, LSQ_SEED + b, radius=scan.n / 2),
                                  scan.n, scan.n, scan.dx, scan.dx) for b in range(8)])
    xt = torch.from_numpy(truth.astype(np.float32)).cuda()
    y = fp(xt)
    lam = recon.operator_norm_sq(scan, "cuda", iters=20)
    g = torch.Generator(device="cuda").manual_seed(LSQ_SEED)
    xa = torch.randn(xt.shape, generator=g, device="cuda")
    ya = torch.randn(y.shape, generator=g, device="cuda")
    ax, aty = fp(xa), fpt(ya)
    gap = abs(float((ax.double() * ya.double()).sum()) - float((xa.double() * aty.double()).sum()))
    gap /= float(ax.double().norm()) * float(ya.double().norm())
    recon.gradient_descent(torch.zeros_like(xt), y, scan, steps=3, alpha=1.0 / lam)  # warm-up
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    x, losses = recon.gradient_descent(torch.zeros_like(xt), y, scan, steps=steps,
                                       alpha=1.0 / lam)
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / steps
    ms_fp = time_events(lambda: fp(x), 10, torch)
    ms_fpt = time_events(lambda: fpt(y), 10, torch)
    ls = losses.cpu().numpy()
    return {
        "workload": "8 x 512^2 seeded ellipses, parallel 720 views/180deg, 729 ch, ds=dx=1; "
                    "x <- x - FP^T(FP x - y)/||FP||^2 via autograd",
        "steps": steps, "ms_per_grad_step": round(ms, 4), "fp_ms": round(ms_fp, 4),
        "fpt_ms": round(ms_fpt, 4), "adjoint_gap": gap,
        "loss_first": float(ls[0]), "loss_last": float(ls[-1]),
        "rel_err_last": float((x - xt).norm() / xt.norm()),
        "fp_ray_samples_per_s": 8 * scan.na * scan.ns * scan.n / (ms_fp * 1e-3),
        "fpt_pixel_views_per_s": 8 * scan.na * scan.n ** 2 / (ms_fpt * 1e-3),
        "roofline_fp": _valu_roofline(
            "fp_lds_kernel", 8 * scan.na * scan.ns * scan.n,
            FP_GEOM_SLOTS / FP_SLICES_PER_WG + FP_SLICE_SLOTS, ms_fp,
            f"ray sample ({FP_GEOM_SLOTS} per ray-row shared by {FP_SLICES_PER_WG} slices "
            f"+ {FP_SLICE_SLOTS} per slice)"),
        "roofline_fpt": _valu_roofline(
            "bp_forward_kernel<FptGeo> + border passes", 8 * scan.na * scan.n ** 2,
            FPT_GEOM_SLOTS / FPT_SLICES_PER_WG + FPT_SLICE_SLOTS, ms_fpt,
            f"pixel*view*slice ({FPT_GEOM_SLOTS} per pixel*view shared by "
            f"{FPT_SLICES_PER_WG} slices + {FPT_SLICE_SLOTS} per slice)"),
    }


def _valu_roofline(kernel, units, slots_per_unit, ms, unit_name):
    achieved = units * slots_per_unit / (ms * 1e-3)
    return {"kernel": kernel, "bound": "valu", "achieved": round(achieved / 1e12, 3),
            "peak": round(PEAK_VALU_SLOTS / 1e12, 3), "unit": "Tslot/s (FP32 VALU lane-issues)",
            "frac": round(achieved / PEAK_VALU_SLOTS, 4),
            "model": f"{slots_per_unit:g} slots per {unit_name} x {units:.4g} per launch"}


def bench_ragged(torch, reps):
    """64 seeded fan problems (n in {128,256,384,512}, 180..1440 irregular
    views, 256..1024 channels, SID/SDD 0.5..0.6) backprojected in one
    persistent launch replayed from a HIP graph; tile and view chunk from the
    autotuner."""
    from ct_recon_extensions_amd import ragged as R
    params, angles = R.random_fan_problems(64, seed=RAGGED_SEED)
    rb = R.RaggedFanBackprojection2DAngle(params, angles, autotune=True, tune_reps=3)
    rb.proj_flat.normal_(generator=torch.Generator(device="cuda").manual_seed(RAGGED_SEED))
    rb(graph=True)
    torch.cuda.synchronize()
    ms = time_events(lambda: rb(graph=True), reps, torch)
    ms_eager = time_events(lambda: rb(graph=False), reps, torch)
    busy = rb.balance()
    out = {
        "workload": "64 fan problems, bp_fan_2d_angle, one persistent graph-replayed launch",
        "seed": RAGGED_SEED, "pixel_views": rb.pixel_views, "ms_per_batch": round(ms, 4),
        "ms_per_batch_eager": round(ms_eager, 4),
        "gpixel_views_per_s": round(rb.pixel_views / (ms * 1e-3) / 1e9, 2),
        "tile": R.TILES[rb.tile], "vchunk": rb.vchunk, "autotune_ms": rb.tune_table,
        "work_items": rb.nitems, "grid": rb.grid, "busy_fraction": round(busy, 4),
        "unbatched_28slot_model_frac": round(
            rb.pixel_views / (ms * 1e-3) * UNBATCHED_FAN_SLOTS / PEAK_VALU_SLOTS, 4),
        # one slice per problem: the geometry of every pixel*view is its own
        "roofline": _valu_roofline("ragged_fan_bp_kernel", rb.pixel_views,
                                   FAN_GEOM_SLOTS + FAN_SLICE_SLOTS, ms, "pixel*view"),
    }
    rb.close()
    return out


def bench_filter_precision(torch, reps):
    """Fan-weighted Ram-Lak filter on the 32-slice precision workload by the
    BF16x3 MFMA Toeplitz GEMM, the FP32 MFMA GEMM and the direct FP32 kernel,
    each against the FP64 filter per slice (relative L2), with times and the
    matrix-core roofline of the GEMMs."""
    import ct_recon_extensions_amd as F
    from ct_recon_extensions_amd import phantom as P
    x, kw, i0s = P.shepp_logan_fan_batch(32, PRECISION_SEED, "cuda")
    ref = F.fan_filter(x.double(), **kw)
    nb, na, ns = x.shape
    res = {"workload": "32 x 1440 x 1472 Shepp-Logan-type fan sinograms, Poisson I0 1e4..1e6, "
                       "fan weighting + Ram-Lak", "seed": PRECISION_SEED,
           "i0": [round(v) for v in i0s], "ms": {}, "rel_l2_per_slice": {},
           "worst_rel_l2": {}}
    for impl in ("bf16x3", "mfma", "direct"):
        y = F.fan_filter(x, **kw, impl=impl)
        d = (y.double() - ref).reshape(nb, -1).norm(dim=1) / ref.reshape(nb, -1).norm(dim=1)
        errs = [float(v) for v in d.cpu()]
        res["rel_l2_per_slice"][impl] = [float(f"{e:.3e}") for e in errs]
        res["worst_rel_l2"][impl] = max(errs)
        res["ms"][impl] = round(time_events(lambda: F.fan_filter(x, **kw, impl=impl), reps,
                                            torch), 4)
    # Parity-split Toeplitz GEMM: M = rows, N = K = ns, half the dense MACs.
    macs = nb * na * ns * ns / 2
    hbm = 2 * x.numel() * 4
    for impl, kern, mult, peak in (("bf16x3", "ramp_persist_bf16x3_kernel<8, 12, 48>", 3, PEAK_MFMA_BF16_TFLOPS),
                                   ("mfma", "ramp_mfma_kernel", 1, PEAK_FP32_MFMA_TFLOPS)):
        tf = mult * 2 * macs / (res["ms"][impl] * 1e9)
        res["roofline_" + ("fp32_mfma" if impl == "mfma" else impl)] = {
            "kernel": kern, "bound": "mfma", "achieved": round(tf, 2), "peak": peak,
            "unit": "TFLOP/s", "frac": round(tf / peak, 4), "flops": mult * 2 * macs,
            "hbm": {"achieved": round(hbm / (res["ms"][impl] * 1e6), 1), "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": round(hbm / (res["ms"][impl] * 1e6) / PEAK_HBM_GBS, 4),
                    "algorithmic_bytes": hbm}}
    res["hbm_bytes"] = hbm
    res["_x"] = (x, kw)
    return res


BWD_SEED = 20261209
# Algorithmic VALU issue slots of the ray-driven FP and of its transpose,
# split like the fan BP's into geometry shared by the slices one workgroup
# carries and work per slice.
# FP, per ray and image row: the FP64 position step and the inclusion test
# (2 + 2, FP64 at half rate), the tap index and fraction (2) = 6 per ray-row,
# shared by the 4 slices of a workgroup; per slice the lerp into the sum, 2.
# FP^T on the engine, per pixel and view: the channel position (a linear
# form, 2), its floor and fraction (2), the two tap weights and the edge
# rule (4) = 8, shared by the 8 slices of a workgroup; per slice 2.
FP_GEOM_SLOTS, FP_SLICE_SLOTS, FP_SLICES_PER_WG = 6, 2, 4
FPT_GEOM_SLOTS, FPT_SLICE_SLOTS, FPT_SLICES_PER_WG = 8, 2, 8
PEAK_MFMA_BF16_TFLOPS = 2500.0
PEAK_FP32_MFMA_TFLOPS = 157.3


def bench_fbp_backward(torch, F, proj, reps):
    """The differentiable fan FBP on the 16-slice workload: fan weighting +
    ramp filter + bp_fan_2d forward and its backward through autograd (the
    BP transpose, then the filter and the weighting transposed), per step;
    the BP transpose alone in both modes (LDS float scatter, and the
    order-independent fixed-point mode); and the adjoint gap of the pair
    on seeded Gaussian inputs."""
    cfg = CFG
    B, n, na, ns = cfg["nbatch"], cfg["n"], cfg["na"], cfg["ns"]
    kwf, kwb = kw_filter(cfg), kw_fan(cfg)
    g = torch.Generator(device="cuda").manual_seed(BWD_SEED)
    gimg = torch.randn((B, n, n), generator=g, device="cuda")
    x = proj.detach().clone().requires_grad_(True)

    def fwd_bwd():
        x.grad = None
        F.fan_backprojection_2d(F.fan_filter(x, **kwf), **kwb).backward(gimg)

    fwd_bwd()
    torch.cuda.synchronize()
    ms_fb = time_events(fwd_bwd, reps, torch)
    ms_bpt = time_events(lambda: F.fan_backprojection_2d_grad(gimg, **kwb), reps, torch)
    ms_bpt_det = time_events(
        lambda: F.fan_backprojection_2d_grad(gimg, **kwb, deterministic=True), reps, torch)
    pr = torch.randn((B, na, ns), generator=g, device="cuda")
    ap = F.fan_backprojection_2d(pr, **kwb).double()
    aty = F.fan_backprojection_2d_grad(gimg, **kwb).double()
    gap = abs(float((ap * gimg.double()).sum()) - float((pr.double() * aty).sum()))
    gap /= float(ap.norm()) * float(gimg.double().norm())
    a1 = F.fan_backprojection_2d_grad(gimg, **kwb, deterministic=True)
    a2 = F.fan_backprojection_2d_grad(gimg, **kwb, deterministic=True)
    pv = B * n * n * na
    # the transpose's algorithmic work is the forward's interpolation, term by
    # term: geometry per pixel*view once per 16-slice group, 2 per slice
    groups = math.ceil(B / 16)
    slots = n * n * na * FAN_GEOM_SLOTS * groups + pv * FAN_SLICE_SLOTS
    return {
        "workload": "16 x 512^2 fan FBP (720 views, 736 ch): forward + backward via autograd; "
                    "backward = BP^T, ramp filter, fan weighting",
        "seed": BWD_SEED, "ms_fwd_bwd_step": round(ms_fb, 4),
        "bpt_ms": round(ms_bpt, 4), "bpt_det_ms": round(ms_bpt_det, 4),
        "bpt_gpixel_views_per_s": round(pv / (ms_bpt * 1e-3) / 1e9, 1),
        "bpt_det_gpixel_views_per_s": round(pv / (ms_bpt_det * 1e-3) / 1e9, 1),
        "roofline_bpt": {"kernel": "bp_gather_t_kernel<FanGeo<false>, 16, 4, 128>",
                         "bound": "valu",
                         "achieved": round(slots / (ms_bpt * 1e-3) / 1e12, 3),
                         "peak": round(PEAK_VALU_SLOTS / 1e12, 3),
                         "unit": "Tslot/s (FP32 VALU lane-issues)",
                         "frac": round(slots / (ms_bpt * 1e-3) / PEAK_VALU_SLOTS, 4),
                         "slots_per_launch": slots},
        "adjoint_gap": gap,
        "det_bitwise_repeatable": bool(torch.equal(a1, a2)),
    }


HELICAL_SEED = 20261208


def helical_scan(nz, na):
    """Helical workload: 736 channels of 1 mm and 32 rows at SID 595 / SDD
    1085.6 mm (17.5 mm collimation at the isocentre), 1152 views per
    rotation, pitch 1, nz slices of 1 mm thickness at 1 mm spacing starting
    1160 views into the scan."""
    m = 32 * 595.0 / 1085.6 / 1152
    return dict(ns=736, nt=32, na=na, ds=1.0, dt=1.0, offset_s=1.25, offset_t=0.0,
                orbit_start=0.0, orbit=2 * math.pi / 1152, couch_begin=0.0, couch_mov=m,
                view_per_rot=1152, nz=nz, dz=1.0, z0=1160 * m, thickness=1.0, dso=595.0,
                dsd=1085.6)


HELICAL_NZ, HELICAL_NA = 96, 8640


def bench_helical(torch, F, reps, with_cpu=True):
    """Helical rebinning (src/helical_interpolation.cc:81-269), 360LI and
    180LI, of one seeded N(0,1) scan of 8640 views x 32 rows x 736 channels
    (814 MB, float32) into 96 fan sinograms of 1152 views: GPU time per
    launch against the HBM roofline, whose compulsory bytes are the views
    the slices' windows span (every row of each) plus the output; and the
    reference CPU path on one thread for 2 of the slices."""
    from ct_recon_extensions_amd.geometry import HelicalInterpolationParam
    geo = helical_scan(HELICAL_NZ, HELICAL_NA)
    p = HelicalInterpolationParam(**geo)
    g = torch.Generator(device="cuda").manual_seed(HELICAL_SEED)
    x = torch.randn((p.na, p.nt, p.ns), generator=g, device="cuda")
    res = {"workload": f"helical scan {p.na} views x {p.nt} rows x {p.ns} ch (1152 views/rot, "
                       f"pitch 1, 17.5 mm collimation) -> {p.nz} slices of 1 mm; seeded N(0,1)",
           "seed": HELICAL_SEED}
    m = geo["couch_mov"]
    cv = [round((geo["z0"] + k * geo["dz"] - geo["couch_begin"]) / m) for k in range(p.nz)]
    out_bytes = p.nz * p.view_per_rot * p.ns * 4
    samples = p.nz * p.view_per_rot * p.ns
    for mode in (360, 180):
        out = F.helical_interpolation(x, p, mode)
        torch.cuda.synchronize()
        ms = time_events(lambda: F.helical_interpolation(x, p, mode, check=False), reps, torch)
        if mode == 360:
            v0, v1 = cv[0] - p.view_per_rot, cv[-1] + p.view_per_rot - 1
        else:
            v0 = cv[0] - p.view_per_rot // 2 - 1
            v1 = cv[-1] - p.view_per_rot // 2 + p.view_per_rot - 1
        in_bytes = (v1 - v0 + 1) * p.nt * p.ns * 4
        bytes_ = in_bytes + out_bytes
        gbs = bytes_ / (ms * 1e-3) / 1e9
        r = {"gpu_ms": round(ms, 4), "msamples_per_s": round(samples / (ms * 1e-3) / 1e6, 1),
             "roofline": {"kernel": f"helical_kernel<float, {mode}>", "bound": "hbm",
                          "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": round(gbs / PEAK_HBM_GBS, 4),
                          "compulsory_bytes": bytes_},
             "valu_issue": _valu_issue(HELI_VALU[mode], ms,
                                       "profiles/helical_rebin_8640v_32x736.md"),
             "nan_free": bool(torch.isfinite(out).all())}
        if with_cpu:
            from oracle import oracle as O
            kind, label = _cpu_kind()
            gs = dict(geo, nz=2)
            hs = O.heli(**gs)
            host = x.cpu().numpy()
            t0 = time.perf_counter()
            ref, code = O.helical(hs, host, mode, kind=kind)
            sec = time.perf_counter() - t0
            r["cpu_baseline"] = {
                "value": round(2 * p.view_per_rot * p.ns / sec / 1e6, 2), "unit": "Msample/s",
                "cores": 1, "kind": label,
                "sample": f"the first 2 of the {p.nz} slices on one thread, {sec:.2f}s"}
            r["bit_identical_first_2_slices"] = bool(
                code == 1 and np.array_equal(out[:2].cpu().numpy(), ref))
        res[f"li{mode}"] = r
    return res


def _cpu_kind():
    from oracle import oracle as O
    return ("ref", "reference") if O.have_ref() else ("oracle", "port")


def cpu_lsq(threads):
    """The reference CPU path of one least-squares gradient step's operators,
    ray-driven FP then its transpose (src/fp_par_2d.cc:210-391), on one
    512^2 slice of the workload: on one thread, and one slice per thread on
    `threads` threads; reported per 8-slice step."""
    from ct_recon_extensions_amd import phantom as P
    from oracle import oracle as O
    kind, label = _cpu_kind()
    n, na, ns = 512, 720, 729
    g = O.geom(ns=ns, na=na, ds=1.0, orbit=math.pi / na, nx=n, ny=n, dx=1.0, dy=1.0)
    tf, tg = O.fp_prep(g, kind), O.fp_grad_prep(g, kind)
    x = P.rasterize(P.random_ellipses(10, LSQ_SEED, radius=n / 2), n, n, 1.0, 1.0)
    x = x.astype(np.float32)

    def one(v):
        return O.fp_grad(g, O.fp(g, v, tf, kind=kind), tg, kind=kind)

    t0 = time.perf_counter()
    one(x)
    single = time.perf_counter() - t0
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(one, [x] * threads))
    wall = time.perf_counter() - t0
    return {"value": round(8 * wall / threads * 1e3, 1), "unit": "ms per 8-slice FP + FP^T",
            "cores": threads, "kind": label,
            "sample": f"{threads} slices of 512^2 x 720 views x 729 ch, FP then FP^T, one per "
                      f"thread, {wall:.1f}s wall",
            "single_thread_value": round(8 * single * 1e3, 1),
            "single_thread_sample": f"1 slice on 1 thread, {single:.1f}s"}


def cpu_ragged(threads):
    """The reference CPU bp_fan_2d_angle (src/bp_fan_2d_angle.cc:47-91) on
    the first `threads` problems of the seeded ragged batch, one per thread,
    and on problem 0 alone on one thread; the whole batch's time is
    extrapolated from the all-thread rate."""
    from ct_recon_extensions_amd import ragged as R
    from oracle import oracle as O
    kind, label = _cpu_kind()
    params, angles = R.random_fan_problems(64, seed=RAGGED_SEED)
    rng = np.random.default_rng(RAGGED_SEED)
    jobs = []
    for p, a in list(zip(params, angles))[:threads]:
        g = O.geom(ns=p.ns, na=p.na, ds=p.ds, offset_s=p.offset_s, nx=p.nx, ny=p.ny, dx=p.dx,
                   dy=p.dy, offset_x=p.offset_x, offset_y=p.offset_y, dso=p.dso, dsd=p.dsd)
        jobs.append((g, O.fan_angle_prep(g, a, kind),
                     rng.standard_normal((p.na, p.ns)).astype(np.float32), p.na * p.nx * p.ny))

    def one(job):
        g, t, pr, _ = job
        return O.fan_bp(g, pr, t, scaled=False, kind=kind)

    t0 = time.perf_counter()
    one(jobs[0])
    single = time.perf_counter() - t0
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(one, jobs))
    wall = time.perf_counter() - t0
    pv = sum(j[3] for j in jobs)
    total = sum(p.na * p.nx * p.ny for p in params)
    rate = pv / wall
    return {"value": round(rate / 1e9, 4), "unit": "Gpixel*view/s", "cores": threads,
            "kind": label,
            "sample": f"problems 0..{threads - 1} of the batch ({pv:.3g} pixel*views), one per "
                      f"thread, {wall:.1f}s wall",
            "batch_ms_extrapolated": round(total / rate * 1e3, 1),
            "single_thread_value": round(jobs[0][3] / single / 1e9, 4),
            "single_thread_sample": f"problem 0 on 1 thread, {single:.1f}s"}


def cpu_filter(threads, x, kw, rows=360):
    """The reference CPU fan weighting + ramp filter in FP64
    (src/fan_weighting.cc:19-35, src/filter.cc:67-91, every one of the 2ns+1
    taps per output as it loops) on `rows` views of one slice per thread, and
    on one thread; reported per 32 x 1440 x 1472 batch."""
    from oracle import oracle as O
    kind, label = _cpu_kind()
    nb, na, ns = x.shape
    g = O.geom(ns=ns, nrow=rows, ds=kw["channel_space"], dso=kw["dso"], dsd=kw["dsd"], type=1)
    taps = O.ramp_prep(g, kind)
    host = [x[b, :rows].double().cpu().numpy() for b in range(min(threads, nb))]

    def one(v):
        return O.ramp(g, O.fanw(g, v, kind=kind), taps, kind=kind)

    t0 = time.perf_counter()
    one(host[0])
    single = time.perf_counter() - t0
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(one, [host[i % len(host)] for i in range(threads)]))
    wall = time.perf_counter() - t0
    per_batch = nb * na / rows
    return {"value": round(wall / threads * per_batch * 1e3, 1),
            "unit": "ms per 32 x 1440 x 1472 batch (FP64)", "cores": threads, "kind": label,
            "sample": f"{rows} rows of one slice per thread on {threads} threads, {wall:.1f}s wall",
            "single_thread_value": round(single * per_batch * 1e3, 1),
            "single_thread_sample": f"{rows} rows on 1 thread, {single:.1f}s"}


def device_spec():
    """The GPU as this run sees it: torch's device properties, and the CU
    count and peak engine clock of the gfx950 agent in rocminfo (run as a
    child process), beside the published constants the roofline uses."""
    import subprocess

    import torch

    spec = {"constants": {"cus": CU, "clock_hz": CLK, "hbm_gb_s": PEAK_HBM_GBS}}
    try:
        pr = torch.cuda.get_device_properties(0)
        spec.update(name=pr.name, arch=getattr(pr, "gcnArchName", ""),
                    cus=pr.multi_processor_count, memory_bytes=pr.total_memory)
    except Exception as e:  # noqa: BLE001 - reported, not fatal
        spec["torch_error"] = str(e)
    try:
        text = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=60).stdout
        for agent in text.split("Agent ")[1:]:
            if "gfx950" not in agent:
                continue
            for line in agent.splitlines():
                key, _, val = line.strip().partition(":")
                if key == "Compute Unit":
                    spec["rocminfo_cus"] = int(val)
                elif key == "Max Clock Freq. (MHz)":
                    spec["rocminfo_max_clock_mhz"] = int(val)
            break
    except Exception as e:  # noqa: BLE001
        spec["rocminfo_error"] = str(e)
    return spec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads for the CPU baseline (default: min(16, cores))")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-slice GPU path")
    ap.add_argument("--no-lsq", action="store_true", help="skip the least-squares grad steps")
    ap.add_argument("--lsq-steps", type=int, default=200)
    ap.add_argument("--no-ragged", action="store_true", help="skip the ragged batch")
    ap.add_argument("--no-precision", action="store_true", help="skip the filter precision study")
    ap.add_argument("--no-c1", action="store_true", help="skip the 256^2 parallel FBP")
    ap.add_argument("--no-svbp", action="store_true", help="skip the single-view parallel BP")
    ap.add_argument("--no-rc", action="store_true", help="skip the ray-casting FP")
    ap.add_argument("--no-helical", action="store_true", help="skip the helical rebinning")
    ap.add_argument("--no-dd", action="store_true",
                    help="skip the distance-driven FP and its gradient")
    ap.add_argument("--no-bwd", action="store_true",
                    help="skip the differentiable FBP's backward (BP transpose)")
    ap.add_argument("--profile-bp", action="store_true",
                    help="counter-collection mode: three fan BP launches of the workload, then a "
                         "256 MiB device copy from a cold source (calibrates FETCH/WRITE_SIZE)")
    ap.add_argument("--profile", choices=("bp", "filter", "lsq", "svbp", "rc", "bpt", "ragged",
                                          "filter5", "helical", "dd"), default=None,
                    help="counter-collection mode for one kernel family: 'bp' as --profile-bp; "
                         "'filter' three fused fan-weighted ramp filters (default implementation) of the "
                         "workload; 'lsq' three FP and FP^T launches of the least-squares "
                         "workload; 'svbp' three single-view parallel BPs of its workload; 'rc' "
                         "three ray-casting FPs of its workload; 'bpt' three fan BP "
                         "transposes of the workload in each mode (scatter, then "
                         "deterministic); 'ragged' three eager runs of the ragged batch "
                         "(16x16 tiles, 32 views per round); 'filter5' three BF16x3 filters "
                         "of the 32 x 1440 x 1472 precision workload; 'helical' three 360LI "
                         "and three 180LI rebinnings of the helical workload; 'dd' three "
                         "distance-driven FPs and gradients of their workload; each followed "
                         "by the calibration copy")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this path shards by slice with no exchange, so "
              "multi-GPU throughput is N x the one-GPU figure; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import ct_recon_extensions_amd as F
    from ct_recon_extensions_amd import _lib
    from ct_recon_extensions_amd.geometry import FanBackprojection2DParam, device_tables

    if args.profile_bp or args.profile is not None:
        profile_kernels(torch, F, "bp" if args.profile_bp else args.profile)
        return

    lib = _lib.load()
    cfg = dict(CFG)
    dev = torch.device("cuda:0")
    t_gen = time.perf_counter()
    sinos = make_inputs(cfg)
    t_gen = time.perf_counter() - t_gen
    print(f"[bench] inputs generated in {t_gen:.1f}s", file=sys.stderr, flush=True)

    proj = torch.from_numpy(sinos).to(dev)
    kwf, kwb = kw_filter(cfg), kw_fan(cfg)
    filt = F.fan_filter(proj, **kwf)
    img = F.fan_backprojection_2d(filt, **kwb)

    def step():
        f = F.fan_filter(proj, **kwf)
        return F.fan_backprojection_2d(f, **kwb)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # per-kernel durations on the stream the kernels are launched on; they run
    # before the step is timed, so the GPU has also left its idle clocks
    # (after seconds of host-side input generation) by then
    ms_filter = time_events(lambda: F.fan_filter(proj, **kwf), args.steps, torch)
    ms_bp = time_events(lambda: F.fan_backprojection_2d(filt, **kwb), args.steps, torch)
    ms_step = time_events(step, args.steps, torch)
    print(f"[bench] step {ms_step:.3f} ms", file=sys.stderr, flush=True)

    B, n, na, ns = cfg["nbatch"], cfg["n"], cfg["na"], cfg["ns"]
    pv_slice = n * n * na
    pv_batch = B * pv_slice
    value = pv_batch / (ms_step * 1e-3) / 1e9
    bp_rate = pv_batch / (ms_bp * 1e-3)
    nb_group = 16
    groups = math.ceil(B / nb_group)
    slots = pv_slice * FAN_GEOM_SLOTS * groups + pv_batch * FAN_SLICE_SLOTS
    valu_achieved = slots / (ms_bp * 1e-3)
    hbm_bytes = B * 4 * (na * ns + n * n) + 16 * na + 8 * 2 * n
    hbm_achieved = hbm_bytes / (ms_bp * 1e-3) / 1e9

    # plain per-slice GPU path: one launch per slice, FP64, stream sync per slice
    plain = None
    if not args.no_plain:
        p = FanBackprojection2DParam(ns, na, cfg["ds"], 2 * math.pi / na, 0.0, 0.0, n, n,
                                     cfg["dx"], cfg["dx"], 0.0, 0.0, cfg["dso"], cfg["dsd"])
        xpos, ypos, sc = device_tables("fan", p, dev)
        out_plain = torch.empty_like(img)
        g = p.ctr()
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)

        def plain_bp():
            _lib.check(lib.ctr_plain_fan_bp(C.byref(g), 0, C.c_void_p(filt.data_ptr()),
                                            C.c_void_p(out_plain.data_ptr()), B,
                                            C.c_void_p(xpos.data_ptr()), C.c_void_p(ypos.data_ptr()),
                                            C.c_void_p(sc.data_ptr()), p.scale, st), "plain")

        plain_bp()
        torch.cuda.synchronize()
        ms_plain = time_events(plain_bp, 2, torch)
        plain_err = float((out_plain - img).double().norm() / out_plain.double().norm())
        # the whole FBP with the same structure: per slice, fan weighting, the
        # direct filter and the plain BP, the stream synchronised after each
        def plain_fbp():
            for b in range(B):
                x = F.fan_weighting(proj[b:b + 1], proj_shape=[na, ns], channel_space=cfg["ds"],
                                    dso=cfg["dso"], dsd=cfg["dsd"], tp="fan")
                torch.cuda.synchronize()
                f = F.ramp_filter(x, ns, na, cfg["ds"], cfg["dsd"], "fan", impl="direct")
                torch.cuda.synchronize()
                _lib.check(lib.ctr_plain_fan_bp(C.byref(g), 0, C.c_void_p(f.data_ptr()),
                                                C.c_void_p(out_plain[b].data_ptr()), 1,
                                                C.c_void_p(xpos.data_ptr()),
                                                C.c_void_p(ypos.data_ptr()),
                                                C.c_void_p(sc.data_ptr()), p.scale, st), "plain")

        plain_fbp()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plain_fbp()
        torch.cuda.synchronize()
        ms_plain_fbp = (time.perf_counter() - t0) * 1e3
        plain_fbp_err = float((out_plain - img).double().norm() / out_plain.double().norm())
        # the reference's own launch geometry: at most 32 workgroups of 512
        # threads, grid-stride (cuda/cuda_common.h:30-38), one launch per slice
        def plain_bp_refgrid():
            _lib.check(lib.ctr_plain_fan_bp_grid(
                C.byref(g), 0, C.c_void_p(filt.data_ptr()), C.c_void_p(out_plain.data_ptr()), B,
                C.c_void_p(xpos.data_ptr()), C.c_void_p(ypos.data_ptr()),
                C.c_void_p(sc.data_ptr()), p.scale, 32, 512, st), "plain reference grid")

        ms_plain_ref = time_events(plain_bp_refgrid, 1, torch)
        plain_ref_err = float((out_plain - img).double().norm() / out_plain.double().norm())
        plain = dict(ms_per_batch_bp=ms_plain, slices_per_s_bp=B / (ms_plain * 1e-3),
                     speedup_bp=ms_plain / ms_bp, rel_l2_fast_vs_plain=plain_err,
                     reference_grid={"grid": "min(32, ceil(npix/512)) x 512 threads, grid-stride",
                                     "ms_per_batch_bp": round(ms_plain_ref, 3),
                                     "slices_per_s_bp": round(B / (ms_plain_ref * 1e-3), 1),
                                     "speedup_bp": round(ms_plain_ref / ms_bp, 1),
                                     "rel_l2_fast_vs_plain": plain_ref_err},
                     ms_per_batch_fbp=ms_plain_fbp, slices_per_s_fbp=B / (ms_plain_fbp * 1e-3),
                     speedup_fbp_slices_per_s=ms_plain_fbp / ms_step,
                     rel_l2_fast_vs_plain_fbp=plain_fbp_err)
        print(f"[bench] plain BP {ms_plain:.2f} ms/batch, plain FBP {ms_plain_fbp:.2f} ms/batch",
              file=sys.stderr, flush=True)

    result = {
        "metric": "fan-beam FBP backprojection throughput",
        "value": round(value, 3),
        "unit": "Gpixel*view/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded analytic ellipse fan sinograms + Poisson 1e5)",
        "config": {"workload": "fan FBP: 16 x 512^2 slices, 720 views/360deg, 736 ch, "
                               "SID 595, SDD 1085.6, dx 0.7, ds 1.30; fan weighting + "
                               "Ram-Lak + bp_fan_2d",
                   "global_batch": B, "parallelism": "single GPU (slices batched per launch)"},
        "slices_per_s": round(B / (ms_step * 1e-3), 1),
        "kernels_ms": {"fan_filter": round(ms_filter, 4), "fan_bp": round(ms_bp, 4)},
        "roofline": {
            "kernel": "fan_bp_kernel",
            "bound": "valu",
            "achieved": round(valu_achieved / 1e12, 3),
            "peak": round(PEAK_VALU_SLOTS / 1e12, 3),
            "unit": "Tslot/s (FP32 VALU lane-issues)",
            "frac": round(valu_achieved / PEAK_VALU_SLOTS, 4),
            "slots_per_launch": slots,
            "traffic": profiled_traffic(),
            "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                            "scaled by a cold 256 MiB copy (" + TRAFFIC_PROFILE + ")",
            "hbm": {"achieved": round(hbm_achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(hbm_achieved / PEAK_HBM_GBS, 5),
                    "algorithmic_bytes": hbm_bytes},
            "bp_pixel_views_per_s": bp_rate,
            "unbatched_28slot_model_frac": round(bp_rate * UNBATCHED_FAN_SLOTS / PEAK_VALU_SLOTS, 4),
        },
        "plain_per_slice_gpu": plain,
    }
    cpu_threads = max(2, args.cpu_threads or min(16, os.cpu_count() or 1))
    if not args.no_bwd:
        result["fbp_backward"] = bench_fbp_backward(torch, F, proj, args.steps)
        fb = result["fbp_backward"]
        print(f"[bench] FBP forward+backward {fb['ms_fwd_bwd_step']:.3f} ms, BP^T "
              f"{fb['bpt_ms']:.3f} ms ({fb['bpt_det_ms']:.3f} deterministic)", file=sys.stderr,
              flush=True)
    if not args.no_lsq:
        result["lsq_grad_step"] = bench_lsq(torch, args.lsq_steps)
        print(f"[bench] lsq grad step {result['lsq_grad_step']['ms_per_grad_step']:.3f} ms",
              file=sys.stderr, flush=True)
        if not args.no_cpu:
            result["lsq_grad_step"]["cpu_baseline"] = cpu_lsq(cpu_threads)
            print("[bench] lsq CPU baseline done", file=sys.stderr, flush=True)
    if not args.no_ragged:
        result["ragged"] = bench_ragged(torch, args.steps)
        print(f"[bench] ragged {result['ragged']['ms_per_batch']:.3f} ms", file=sys.stderr,
              flush=True)
        if not args.no_cpu:
            result["ragged"]["cpu_baseline"] = cpu_ragged(cpu_threads)
            print("[bench] ragged CPU baseline done", file=sys.stderr, flush=True)

    if not args.no_c1:
        result["c1_parallel_fbp"] = bench_c1(torch, F, args.cpu_threads or min(16, os.cpu_count() or 1),
                                             with_cpu=not args.no_cpu)
        print(f"[bench] c1 parallel FBP {result['c1_parallel_fbp']['gpu_ms']:.3f} ms",
              file=sys.stderr, flush=True)
    if not args.no_svbp:
        result["single_view_bp"] = bench_svbp(torch, F, 20, with_cpu=not args.no_cpu)
        sv = result["single_view_bp"]
        print(f"[bench] single-view BP {sv['gpu_ms']:.3f} ms, {sv['gb_s']:.0f} GB/s",
              file=sys.stderr, flush=True)
    if not args.no_rc:
        result["raycasting_fp"] = bench_raycasting(torch, F, 5, with_cpu=not args.no_cpu)
        rc_ = result["raycasting_fp"]
        print(f"[bench] ray-casting FP {rc_['gpu_ms']:.3f} ms, {rc_['gsamples_per_s']:.0f} "
              f"Gsamples/s", file=sys.stderr, flush=True)
    if not args.no_dd:
        result["distance_driven_fp"] = bench_fp_dd(torch, F, 10, with_cpu=not args.no_cpu)
        dd = result["distance_driven_fp"]
        print(f"[bench] distance-driven FP {dd['fp_ms']:.3f} ms, gradient {dd['grad_ms']:.3f} ms",
              file=sys.stderr, flush=True)
    if not args.no_helical:
        result["helical_rebinning"] = bench_helical(torch, F, 20, with_cpu=not args.no_cpu)
        hb = result["helical_rebinning"]
        print(f"[bench] helical 360LI {hb['li360']['gpu_ms']:.3f} ms, 180LI "
              f"{hb['li180']['gpu_ms']:.3f} ms", file=sys.stderr, flush=True)
    if not args.no_precision:
        result["filter_precision"] = bench_filter_precision(torch, args.steps)
        fp_ = result["filter_precision"]
        x5, kw5 = fp_.pop("_x")
        print(f"[bench] filter bf16x3 {fp_['ms']['bf16x3']:.3f} ms, worst slice rel L2 "
              f"{fp_['worst_rel_l2']['bf16x3']:.2e}", file=sys.stderr, flush=True)
        if not args.no_cpu:
            fp_["cpu_baseline"] = cpu_filter(cpu_threads, x5, kw5)
            print("[bench] filter CPU baseline done", file=sys.stderr, flush=True)
        del x5

    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        threads = max(2, threads)
        print(f"[bench] CPU baseline on {threads} threads ...", file=sys.stderr, flush=True)
        cb = cpu_baseline(cfg, sinos, threads)
        got = img[0].double().cpu().numpy()
        e32 = float(np.linalg.norm(got - cb["ref_f32"]) / np.linalg.norm(cb["ref_f32"]))
        e64 = float(np.linalg.norm(got - cb["ref_f64"]) / np.linalg.norm(cb["ref_f64"]))
        cpu_rate = cb["pixel_views"] / cb["wall_s"] / 1e9
        result["cpu_baseline"] = {
            "value": round(cpu_rate, 5), "unit": "Gpixel*view/s", "cores": threads,
            "kind": cb["kind"],
            "sample": f"{cb['slices']} full slices of the same workload (FBP incl. weighting "
                      f"and ramp), one per thread, {cb['wall_s']:.1f}s wall",
            "host_cpu": _cpu_model(), "nproc": os.cpu_count(),
            "single_thread_value": round(cfg["n"] ** 2 * cfg["na"] / cb["single_s"] / 1e9, 5),
            "single_thread_sample": f"1 slice on 1 thread, {cb['single_s']:.1f}s",
        }
        result["vs_cpu"] = round(value / cpu_rate, 1)
        result["rel_l2_vs_reference_f32"] = e32
        result["rel_l2_vs_fp64"] = e64

    result["device"] = device_spec()
    print(json.dumps(result))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
