"""GPU parity for the workloads and operator variants the per-operator tests
leave out: the 256^2 parallel FBP stage by stage, the fan-angle transpose
against the oracle, the plain per-slice path at both launch grids, 5-D ramp
input, and the transposes' float64 gather form and deterministic scales.

Tolerances as in test_gpu_parity.py: relative L2 <= 1e-5 for float32
against the oracle (bit-identical to the reference's CPU path), <= 1e-12
for float64.
"""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import load_golden, randn, rel_l2

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-12


@pytest.fixture(scope="module")
def F():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import ct_recon_extensions_amd as F
    from ct_recon_extensions_amd import _lib
    _lib.load()  # raises if the HIP library is missing
    return F


def cuda(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def fan_kwargs(geo, angle=False):
    kw = dict(img_shape=[geo["ny"], geo["nx"]], img_space=[geo["dy"], geo["dx"]],
              proj_shape=[geo["na"], geo["ns"]], channel_space=geo["ds"],
              img_offset=[geo["offset_y"], geo["offset_x"]], channel_offset=geo["offset_s"],
              dso=geo["dso"], dsd=geo["dsd"])
    if not angle:
        kw.update(orbit_start=geo["orbit_start"], orbit=geo["orbit"])
    return kw


def par_kwargs(geo):
    return dict(img_shape=[geo["ny"], geo["nx"]], img_space=[geo["dy"], geo["dx"]],
                proj_shape=[geo["na"], geo["ns"]], channel_space=geo["ds"],
                img_offset=[geo["offset_y"], geo["offset_x"]], channel_offset=geo["offset_s"],
                orbit_start=geo["orbit_start"], orbit=geo["orbit"])


def ogeom(geo):
    return O.geom(**{k: v for k, v in geo.items() if k in dict(O.Geom._fields_)})


# --------------------------------------------- 256^2 parallel FBP (C1)
C1_SEED = 20261206


@pytest.mark.parametrize("impl", ["bf16x3", "mfma", "direct"])
def test_c1_parallel_fbp_stage_by_stage(F, impl):
    """The 256^2 ten-ellipse phantom, ray-driven FP to 360 views over 180 deg
    and 367 channels (ds = dx = 1), Ram-Lak, pixel-driven parallel BP: each
    GPU stage on the oracle's input of that stage, and the whole GPU chain,
    against the oracle chain."""
    from ct_recon_extensions_amd import phantom as P
    n, na, ns = 256, 360, 367
    img = P.rasterize(P.random_ellipses(10, C1_SEED, radius=n / 2), n, n, 1.0, 1.0)
    img = img.astype(np.float32)
    geo = dict(ns=ns, na=na, ds=1.0, orbit=math.pi / na, offset_s=0.0, orbit_start=0.0,
               nx=n, ny=n, dx=1.0, dy=1.0, offset_x=0.0, offset_y=0.0)
    g = O.geom(**geo, type=0, nrow=na)
    kw = par_kwargs(geo)
    sino_ref = O.fp(g, img, O.fp_prep(g))
    filt_ref = O.ramp(g, sino_ref, O.ramp_prep(g))
    bp_ref = O.par_bp(g, filt_ref, O.par_prep(g))

    sino = F.parallel_projection_2d(cuda(img[None]), **kw)
    assert rel_l2(host(sino)[0], sino_ref) < TOL32
    filt = F.ramp_filter(cuda(sino_ref[None]), ns, na, 1.0, type="par", impl=impl)
    assert rel_l2(host(filt)[0], filt_ref) < TOL32
    bp = F.parallel_backprojection_2d(cuda(filt_ref[None]), **kw)
    assert rel_l2(host(bp)[0], bp_ref) < TOL32
    chain = F.parallel_backprojection_2d(
        F.ramp_filter(sino, ns, na, 1.0, type="par", impl=impl), **kw)
    err = rel_l2(host(chain)[0], bp_ref)
    print(f"256^2 parallel FBP ({impl}) rel L2 vs oracle: {err:.3e}")
    assert err < TOL32


# ------------------------------------ fan-angle transpose vs the oracle (A9)
@pytest.mark.parametrize("det", [False, True])
def test_fan_angle_transpose_golden_geometry(F, det):
    geo, d = load_golden("fan_angle_small")
    g = ogeom(geo)
    t = O.fan_angle_prep(g, d["angles"])
    img = randn(31, (3, geo["ny"], geo["nx"]))
    out = host(F.fan_backprojection_2d_angle_grad(cuda(img), d["angles"],
                                                  **fan_kwargs(geo, angle=True),
                                                  deterministic=det))
    for b in range(3):
        assert rel_l2(out[b], O.fan_bpt(g, img[b], t, scaled=False)) < TOL32, b
    out64 = host(F.fan_backprojection_2d_angle_grad(cuda(img, torch.float64), d["angles"],
                                                    **fan_kwargs(geo, angle=True),
                                                    deterministic=det))
    for b in range(3):
        ref = O.fan_bpt(g, img[b].astype(np.float64), t, scaled=False)
        assert rel_l2(out64[b], ref) < TOL64, b


def test_fan_angle_transpose_ragged_problem(F):
    """The smallest problem of the seeded 64-problem ragged batch (irregular
    angles, its own detector and source distance)."""
    from ct_recon_extensions_amd import ragged as R
    params, angles = R.random_fan_problems(64, seed=20261204)
    i = min(range(64), key=lambda k: params[k].na * params[k].nx * params[k].ny)
    p, a = params[i], angles[i]
    geo = dict(ns=p.ns, na=p.na, ds=p.ds, offset_s=p.offset_s, nx=p.nx, ny=p.ny, dx=p.dx,
               dy=p.dy, offset_x=p.offset_x, offset_y=p.offset_y, dso=p.dso, dsd=p.dsd)
    g = ogeom(geo)
    t = O.fan_angle_prep(g, a)
    img = randn(32, (2, p.ny, p.nx))
    for det in (False, True):
        out = host(F.fan_backprojection_2d_angle_grad(cuda(img), a, **fan_kwargs(geo, True),
                                                      deterministic=det))
        for b in range(2):
            assert rel_l2(out[b], O.fan_bpt(g, img[b], t, scaled=False)) < TOL32, (det, b)


# ------------------------------------------------ plain per-slice path
@pytest.mark.parametrize("grid", [(0, 0), (32, 512)], ids=["full", "reference"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_plain_fan_bp_vs_oracle(F, grid, dtype):
    """ctr_plain_fan_bp_grid (one launch per slice, FP64 tables and
    accumulation) with a full grid and with the reference's <= 32 x 512
    threads, grid-stride."""
    from ct_recon_extensions_amd import _lib
    from ct_recon_extensions_amd.functional import _fan_param
    from ct_recon_extensions_amd.geometry import device_tables
    geo = dict(ns=96, na=90, ds=1.2, orbit=2 * math.pi / 90, offset_s=0.3, orbit_start=0.2,
               nx=150, ny=130, dx=0.9, dy=0.8, offset_x=0.1, offset_y=-0.2, dso=150.0,
               dsd=260.0)
    p = _fan_param([geo["ny"], geo["nx"]], [geo["dy"], geo["dx"]], [geo["na"], geo["ns"]],
                   geo["ds"], [geo["offset_y"], geo["offset_x"]], geo["offset_s"],
                   geo["orbit_start"], geo["orbit"], geo["dso"], geo["dsd"], -1)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    proj = randn(41, (3, geo["na"], geo["ns"])).astype(npdt)
    x = cuda(proj)
    out = torch.empty((3, geo["ny"], geo["nx"]), dtype=dtype, device="cuda")
    xpos, ypos, sc = device_tables("fan", p, x.device)
    lib = _lib.load()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.ctr_plain_fan_bp_grid(C.byref(p.ctr()), 0 if dtype == torch.float32 else 1,
                                         C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), 3,
                                         C.c_void_p(xpos.data_ptr()), C.c_void_p(ypos.data_ptr()),
                                         C.c_void_p(sc.data_ptr()), p.scale, grid[0], grid[1],
                                         st), "plain")
    got = host(out)
    g = ogeom(geo)
    t = O.fan_prep(g)
    for b in range(3):
        assert rel_l2(got[b], O.fan_bp(g, proj[b], t)) < (TOL32 if npdt == np.float32 else TOL64)


# ---------------------------------------------------------- 5-D ramp input
def test_ramp_filter_5d_input(F):
    """[N, D, H, ns, 1] with D*H = n_row (tensorflow/ramp_filter_ops.cu:130-143):
    every row filtered, the shape kept, and the gradient the same filter."""
    ns, d, h = 97, 3, 5
    x = randn(51, (2, d, h, ns, 1))
    g = O.geom(ns=ns, nrow=2 * d * h, ds=0.9, dsd=400.0, type=1)  # every row of the batch
    taps = O.ramp_prep(g)
    xt = cuda(x).requires_grad_(True)
    y = F.ramp_filter(xt, ns, d * h, 0.9, 400.0, "fan")
    assert tuple(y.shape) == x.shape
    got = host(y).reshape(-1, ns)
    ref = O.ramp(g, np.ascontiguousarray(x.reshape(-1, ns)), taps)
    assert rel_l2(got, ref) < TOL32
    gy = randn(52, x.shape)
    y.backward(cuda(gy))
    gref = O.ramp(g, np.ascontiguousarray(gy.reshape(-1, ns)), taps, grad=True)
    assert rel_l2(host(xt.grad).reshape(-1, ns), gref) < TOL32


# ------------------------------------------------ transposes: FP64 gather
@pytest.mark.parametrize("det", [False, True])
def test_fan_bpt_float64_gather_vs_oracle_and_repeatable(F, det):
    for name in ("fan_small", "fan_c2_coarse"):
        geo, _ = load_golden(name)
        g = ogeom(geo)
        t = O.fan_prep(g)
        img = randn(61, (5, geo["ny"], geo["nx"])).astype(np.float64)
        kw = fan_kwargs(geo)
        a = host(F.fan_backprojection_2d_grad(cuda(img), **kw, deterministic=det))
        b = host(F.fan_backprojection_2d_grad(cuda(img), **kw, deterministic=det))
        np.testing.assert_array_equal(a, b)
        for s in (0, 4):
            assert rel_l2(a[s], O.fan_bpt(g, img[s], t)) < TOL64, (name, s)


@pytest.mark.parametrize("det", [False, True])
def test_par_bpt_float64_gather_vs_oracle_and_repeatable(F, det):
    from tests.golden.make_golden import PAR_SMALL
    for geo in (dict(PAR_SMALL),
                dict(ns=367, na=360, ds=1.0, orbit=math.pi / 360, offset_s=0.0,
                     orbit_start=0.0, nx=256, ny=256, dx=1.0, dy=1.0, offset_x=0.0,
                     offset_y=0.0)):
        g = ogeom(geo)
        t = O.par_prep(g)
        img = randn(62, (5, geo["ny"], geo["nx"])).astype(np.float64)
        kw = par_kwargs(geo)
        a = host(F.parallel_backprojection_2d_grad(cuda(img), **kw, deterministic=det))
        b = host(F.parallel_backprojection_2d_grad(cuda(img), **kw, deterministic=det))
        np.testing.assert_array_equal(a, b)
        for s in (0, 4):
            assert rel_l2(a[s], O.par_bpt(g, img[s], t)) < TOL64, (geo["nx"], s)


# ------------------------------------------ transposes: float32 gather
def test_transposes_of_zero_images_are_exact_zeros(F):
    geo, _ = load_golden("fan_small")
    z = torch.zeros((3, geo["ny"], geo["nx"]), device="cuda")
    for det in (False, True):
        out = host(F.fan_backprojection_2d_grad(z, **fan_kwargs(geo), deterministic=det))
        assert np.all(out == 0) and not np.any(np.signbit(out))
    from tests.golden.make_golden import PAR_SMALL
    zp = torch.zeros((2, PAR_SMALL["ny"], PAR_SMALL["nx"]), device="cuda")
    outp = host(F.parallel_backprojection_2d_grad(zp, **par_kwargs(PAR_SMALL)))
    assert np.all(outp == 0)


def test_transpose_slices_are_independent(F):
    """Slices 1e-6 and 1e6 times one image in one batch keep float32
    accuracy each (every slice has its own accumulators), and an Inf in a
    third slice stays in that slice."""
    geo, _ = load_golden("fan_c2_coarse")
    g = ogeom(geo)
    t = O.fan_prep(g)
    img = randn(63, (geo["ny"], geo["nx"]))
    ref = O.fan_bpt(g, img, t)
    batch = np.stack([img * 1e-6, img * 1e6, img, img]).astype(np.float32)
    batch[2, 5, 7] = np.inf
    out = host(F.fan_backprojection_2d_grad(cuda(batch), **fan_kwargs(geo)))
    assert rel_l2(out[0], ref * 1e-6) < TOL32
    assert rel_l2(out[1], ref * 1e6) < TOL32
    assert not np.all(np.isfinite(out[2]))
    assert rel_l2(out[3], ref) < TOL32


@pytest.mark.parametrize("nbatch", [1, 2, 5, 16, 19])
def test_fan_bpt_gather_batches_and_full_circle(F, nbatch):
    """Every slice-group width of the gather (1, 4, 8, 16 slices per
    workgroup, partial groups) on a full-circle scan, whose views walk both
    the rows and the columns of the image, against the oracle's transpose."""
    geo = dict(ns=150, na=64, ds=1.1, orbit=2 * math.pi / 64, offset_s=0.7, orbit_start=0.15,
               nx=70, ny=45, dx=1.2, dy=1.0, offset_x=0.4, offset_y=-0.3, dso=200.0,
               dsd=350.0)
    g = ogeom(geo)
    t = O.fan_prep(g)
    img = randn(64 + nbatch, (nbatch, geo["ny"], geo["nx"]))
    out = host(F.fan_backprojection_2d_grad(cuda(img), **fan_kwargs(geo)))
    for b in range(nbatch):
        assert rel_l2(out[b], O.fan_bpt(g, img[b], t)) < TOL32, b


def test_fan_bpt_full_size_vs_oracle(F):
    """The 512^2 x 720-view x 736-channel fan geometry: one slice of the
    gather transpose against the oracle's (FP64 geometry, ~6 s on the host),
    and a 16-slice batch of power-of-two multiples of it."""
    geo = dict(ns=736, na=720, ds=1.30, orbit=2 * math.pi / 720, offset_s=0.0, orbit_start=0.0,
               nx=512, ny=512, dx=0.7, dy=0.7, offset_x=0.0, offset_y=0.0, dso=595.0,
               dsd=1085.6)
    g = ogeom(geo)
    img = randn(65, (512, 512))
    ref = O.fan_bpt(g, img.astype(np.float64), O.fan_prep(g))
    scales = 2.0 ** (np.arange(16) - 8)
    batch = (scales[:, None, None] * img[None]).astype(np.float32)
    out = host(F.fan_backprojection_2d_grad(cuda(batch), **fan_kwargs(geo)))
    for b in (0, 7, 15):
        err = rel_l2(out[b], scales[b] * ref)
        assert err < TOL32, (b, err)
    single = host(F.fan_backprojection_2d_grad(cuda(img[None]), **fan_kwargs(geo)))
    assert rel_l2(single[0], ref) < TOL32


# ------------------------------------ gather transposes: edge geometries
# Channel counts from the smallest admitted (2) to around the 63-channel
# workgroup (lane 0 owns the interval left of the workgroup's channels), a
# single view and a few views, thin and one-pixel images, a wide fan whose
# channels' level lines turn far across one workgroup, and detector offsets
# that put the image's shadow partly off the detector.
_EDGE_FAN = [
    dict(ns=3, na=5, nx=9, ny=7, ds=3.0, dso=60.0, dsd=100.0, offset_s=0.0),
    dict(ns=2, na=1, nx=13, ny=13, ds=2.0, dso=60.0, dsd=100.0, offset_s=0.3),
    dict(ns=62, na=3, nx=1, ny=40, ds=0.9, dso=80.0, dsd=140.0, offset_s=-0.4),
    dict(ns=63, na=7, nx=33, ny=1, ds=0.8, dso=80.0, dsd=140.0, offset_s=0.0),
    dict(ns=64, na=16, nx=40, ny=24, ds=1.0, dso=70.0, dsd=120.0, offset_s=5.5),
    dict(ns=127, na=9, nx=48, ny=48, ds=1.6, dso=60.0, dsd=90.0, offset_s=0.0),
    dict(ns=190, na=33, nx=31, ny=57, ds=0.6, dso=120.0, dsd=200.0, offset_s=-20.0),
]


@pytest.mark.parametrize("k", range(len(_EDGE_FAN)))
def test_fan_bpt_edge_geometries_vs_oracle(F, k):
    e = _EDGE_FAN[k]
    geo = dict(e, orbit=2 * math.pi / e["na"] * 0.93, orbit_start=0.37, dx=1.1, dy=0.9,
               offset_x=0.3, offset_y=-0.2)
    g = ogeom(geo)
    t = O.fan_prep(g)
    for nbatch in (1, 6):
        img = randn(700 + 10 * k + nbatch, (nbatch, geo["ny"], geo["nx"]))
        out = host(F.fan_backprojection_2d_grad(cuda(img), **fan_kwargs(geo)))
        for b in range(nbatch):
            ref = O.fan_bpt(g, img[b], t)
            err = rel_l2(out[b], ref) if np.any(ref) else float(np.abs(out[b]).max())
            assert err < TOL32, (k, nbatch, b, err)


_EDGE_PAR = [
    dict(ns=2, na=4, nx=6, ny=6, ds=4.0, offset_s=0.0),
    dict(ns=63, na=1, nx=20, ny=11, ds=0.5, offset_s=0.25),
    dict(ns=64, na=6, nx=1, ny=30, ds=0.7, offset_s=-3.0),
    dict(ns=126, na=13, nx=45, ny=45, ds=0.6, offset_s=0.0),
]


@pytest.mark.parametrize("k", range(len(_EDGE_PAR)))
def test_par_bpt_edge_geometries_vs_oracle(F, k):
    e = _EDGE_PAR[k]
    geo = dict(e, orbit=math.pi / e["na"] * 1.07, orbit_start=0.21, dx=1.0, dy=1.2,
               offset_x=-0.4, offset_y=0.1)
    g = ogeom(geo)
    t = O.par_prep(g)
    for nbatch in (1, 5):
        img = randn(800 + 10 * k + nbatch, (nbatch, geo["ny"], geo["nx"]))
        out = host(F.parallel_backprojection_2d_grad(cuda(img), **par_kwargs(geo)))
        for b in range(nbatch):
            ref = O.par_bpt(g, img[b], t)
            err = rel_l2(out[b], ref) if np.any(ref) else float(np.abs(out[b]).max())
            assert err < TOL32, (k, nbatch, b, err)
