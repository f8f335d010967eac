"""GPU parity: the HIP kernels against the oracle on the same seeded inputs.

Tolerances (float32 path; the north-star bar is 1e-5 relative L2 against
the reference's CPU path):
  * BP / FP / filter / weighting outputs: relative L2 <= 1e-5 vs the oracle
    (the oracle itself is bit-identical to the compiled reference);
  * float64 buffers: relative L2 <= 1e-12 (the FP64 kernels restate the
    reference's FP64 arithmetic);
  * transposes: relative L2 <= 1e-5 vs the oracle's transpose where one
    exists, and dot-product adjoint gap <= 1e-6 against the forward kernel.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import load_golden, randn, rel_l2

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-12
ADJ = 1e-6


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


def fan_kwargs(geo):
    return dict(img_shape=[geo["ny"], geo["nx"]], img_space=[geo["dy"], geo["dx"]],
                proj_shape=[geo["na"], geo["ns"]], channel_space=geo["ds"],
                img_offset=[geo["offset_y"], geo["offset_x"]], channel_offset=geo["offset_s"],
                orbit_start=geo["orbit_start"], orbit=geo["orbit"], dso=geo["dso"],
                dsd=geo["dsd"])


def par_kwargs(geo):
    kw = fan_kwargs(geo)
    kw.pop("dso")
    kw.pop("dsd")
    return kw


def adjoint_gap(A, At, x, y):
    ax = A(x)
    aty = At(y)
    lhs = float(torch.sum(ax.double() * y.double()))
    rhs = float(torch.sum(x.double() * aty.double()))
    return abs(lhs - rhs) / (float(ax.double().norm()) * float(y.double().norm()))


# ------------------------------------------------------------------ fan BP
def test_fan_bp_small_golden(F):
    geo, d = load_golden("fan_small")
    out = host(F.fan_backprojection_2d(cuda(d["proj"][None, :, :, None]), **fan_kwargs(geo)))
    assert out.shape == (1, geo["ny"], geo["nx"], 1)
    assert rel_l2(out[0, :, :, 0], d["bp_f32"]) < TOL32


def test_fan_bp_c2_coarse_golden(F):
    geo, d = load_golden("fan_c2_coarse")
    proj = randn(int(d["proj_seed"]), (geo["na"], geo["ns"]))
    out = host(F.fan_backprojection_2d(cuda(proj[None]), **fan_kwargs(geo)))
    assert rel_l2(out[0], d["bp_f32"]) < TOL32


@pytest.mark.parametrize("nbatch", [1, 3, 17])
def test_fan_bp_batches_match_per_slice_oracle(F, nbatch):
    from tests.golden.make_golden import FAN_SMALL
    geo = dict(FAN_SMALL, offset_s=0.4, fov=-1.0)
    g = O.geom(**FAN_SMALL)
    t = O.fan_prep(g)
    proj = randn(500 + nbatch, (nbatch, g.na, g.ns))
    out = host(F.fan_backprojection_2d(cuda(proj), **fan_kwargs(geo)))
    for b in range(nbatch):
        assert rel_l2(out[b], O.fan_bp(g, proj[b], t)) < TOL32, b


# The launch shape (tile, views per LDS round, slices per workgroup) does not
# change the result beyond float32 rounding: every configuration against the
# oracle, including view chunks past the LDS budget (capped) and 48- and
# 64-channel span rows.  The shape travels with the call (ctr_fan_bp_ex).
@pytest.mark.parametrize("tile,vchunk,nb", [(0, 0, 0), (0, 1, 4), (0, 6, 16), (0, 64, 8),
                                            (1, 0, 0), (2, 5, 16), (1, 3, 2)])
def test_fan_bp_launch_configs_match_oracle(F, tile, vchunk, nb):
    import ctypes as C
    from ct_recon_extensions_amd import _lib
    from ct_recon_extensions_amd.functional import _fan_param
    from ct_recon_extensions_amd.geometry import device_tables
    from tests.golden.make_golden import FAN_SMALL
    geo = dict(FAN_SMALL, fov=-1.0)
    g = O.geom(**FAN_SMALL)
    t = O.fan_prep(g)
    proj = randn(700 + vchunk, (16, g.na, g.ns))
    p = _fan_param([geo["ny"], geo["nx"]], [geo["dy"], geo["dx"]], [geo["na"], geo["ns"]],
                   geo["ds"], [geo["offset_y"], geo["offset_x"]], geo["offset_s"],
                   geo["orbit_start"], geo["orbit"], geo["dso"], geo["dsd"], -1)
    x = cuda(proj)
    out = torch.empty((16, g.ny, g.nx), device="cuda")
    xpos, ypos, sc = device_tables("fan", p, x.device)
    launch = _lib.CtrFanBpLaunch(tile, vchunk, nb)
    _lib.check(_lib.load().ctr_fan_bp_ex(
        C.byref(p.ctr()), 0, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), 16,
        C.c_void_p(xpos.data_ptr()), C.c_void_p(ypos.data_ptr()), C.c_void_p(sc.data_ptr()),
        p.scale, C.byref(launch), C.c_void_p(torch.cuda.current_stream().cuda_stream)),
        "ctr_fan_bp_ex")
    out = host(out)
    for b in range(16):
        assert rel_l2(out[b], O.fan_bp(g, proj[b], t)) < TOL32, b


def _fan_chunk_kinds(geo, tables, tw=16, th=16, chunk=64):
    """(all-inside, edge) counts of (pixel tile, 64-view chunk) pairs: a
    chunk is an edge chunk when some view puts a pixel of the tile within a
    quarter channel of the detector's ends or past them (the BP runs its
    per-pixel edge test only in those chunks)."""
    xpos, ypos, sc = tables
    g = geo
    cents = (g["ns"] - 1) / 2 + g["offset_s"]
    sb, cb = sc[0::2][:, None, None], sc[1::2][:, None, None]
    x, y = xpos[None, None, :], ypos[None, :, None]
    d = g["dso"] + x * sb - y * cb
    r = x * cb + y * sb
    s = np.arctan2(r, d) * g["dsd"] / g["ds"] + cents       # [view][iy][ix]
    inside = (s >= 0.25) & (s <= g["ns"] - 1.25)
    counts = [0, 0]
    for ty in range(0, g["ny"], th):
        for tx in range(0, g["nx"], tw):
            tile_in = inside[:, ty:ty + th, tx:tx + tw].all(axis=(1, 2))
            for v0 in range(0, g["na"], chunk):
                counts[0 if tile_in[v0:v0 + chunk].all() else 1] += 1
    return counts


# A detector that covers the centre tiles for every view and the outer tiles
# for only some, over 100 views (anchor chunks of 64 and 36): both copies of
# the BP's view loop (with and without the per-pixel edge test) run in one
# launch, on the slice-interleaved path (16 slices) and the pair path (3).
@pytest.mark.parametrize("nbatch", [16, 3])
def test_fan_bp_edge_and_inside_chunks(F, nbatch):
    geo = dict(ns=128, na=100, ds=1.0, orbit=2 * math.pi / 100, offset_s=3.3, orbit_start=0.1,
               nx=64, ny=64, dx=1.0, dy=1.0, offset_x=0.0, offset_y=0.0, dso=100.0, dsd=200.0)
    g = O.geom(**geo)
    t = O.fan_prep(g)
    n_in, n_edge = _fan_chunk_kinds(geo, t)
    assert n_in >= 4 and n_edge >= 4, (n_in, n_edge)
    proj = randn(900 + nbatch, (nbatch, g.na, g.ns))
    out = host(F.fan_backprojection_2d(cuda(proj), **fan_kwargs(geo)))
    for b in range(nbatch):
        assert rel_l2(out[b], O.fan_bp(g, proj[b], t)) < TOL32, b


def test_fan_bp_float64(F):
    geo, d = load_golden("fan_small")
    out = host(F.fan_backprojection_2d(cuda(d["proj"][None], torch.float64), **fan_kwargs(geo)))
    assert rel_l2(out[0], d["bp_f64"]) < TOL64


def test_fan_bp_ragged_sizes_and_tiny_views(F):
    """Image sizes off the tile grid, a single view, ns barely > 1."""
    for (nx, ny, na, ns) in [(37, 23, 5, 50), (1, 1, 1, 9), (70, 3, 67, 129)]:
        geo = dict(ns=ns, na=na, ds=1.1, orbit=2 * math.pi / max(na, 2), offset_s=-0.3,
                   orbit_start=0.2, nx=nx, ny=ny, dx=0.8, dy=0.9, offset_x=0.1, offset_y=0.2,
                   dso=80.0, dsd=150.0)
        g = O.geom(**geo)
        proj = randn(nx + ny, (2, na, ns))
        out = host(F.fan_backprojection_2d(cuda(proj), **fan_kwargs(geo)))
        t = O.fan_prep(g)
        for b in range(2):
            assert rel_l2(out[b], O.fan_bp(g, proj[b], t)) < TOL32, (nx, ny, na, ns, b)


def test_fan_bp_full_size_slice_vs_oracle(F):
    """One 512x512 slice of the 720-view / 736-channel fan geometry against
    the oracle (FP64, ~8 s on the host)."""
    geo = dict(ns=736, na=720, ds=1.30, orbit=2 * math.pi / 720, offset_s=0.0, orbit_start=0.0,
               nx=512, ny=512, dx=0.7, dy=0.7, offset_x=0.0, offset_y=0.0, dso=595.0,
               dsd=1085.6)
    from ct_recon_extensions_amd import phantom as P
    ell = P.random_ellipses(10, 1234, radius=0.7 * 256)
    sino = P.fan_sinogram(ell, 720, 736, 1.30, 2 * math.pi / 720, 595.0, 1085.6)
    sino = P.poisson_noise(sino, 1e5, 99).astype(np.float32)
    g = O.geom(**geo, type=1, nrow=720)
    filt = O.ramp(g, O.fanw(g, sino), O.ramp_prep(g))
    ref = O.fan_bp(g, filt, O.fan_prep(g))
    out = host(F.fan_backprojection_2d(cuda(filt[None]), **fan_kwargs(geo)))
    err = rel_l2(out[0], ref)
    print(f"fan BP 512^2 x 720 views, rel L2 vs oracle: {err:.3e}")
    assert err < TOL32
    # 16 slices (the batched kernel: slice-interleaved 48-channel spans):
    # power-of-two multiples of the same sinogram, exact in float32, so each
    # slice must be its multiple of the single-slice oracle result
    scales = 2.0 ** (np.arange(16) - 8)
    batch = (scales[:, None, None] * filt[None]).astype(np.float32)
    outb = host(F.fan_backprojection_2d(cuda(batch), **fan_kwargs(geo)))
    for b in range(16):
        assert rel_l2(outb[b], scales[b] * ref) < TOL32, b


# ------------------------------------------------------- fan BP transpose
def test_fan_bpt_small_golden(F):
    geo, d = load_golden("fan_small")
    for det in (False, True):
        out = host(F.fan_backprojection_2d_grad(cuda(d["img"][None]), **fan_kwargs(geo),
                                                deterministic=det))
        assert rel_l2(out[0], d["bpt_f32"]) < TOL32, det


def test_fan_bpt_deterministic_is_bitwise_repeatable(F):
    geo, d = load_golden("fan_c2_coarse")
    img = cuda(randn(7, (5, geo["ny"], geo["nx"])))
    a = host(F.fan_backprojection_2d_grad(img, **fan_kwargs(geo), deterministic=True))
    b = host(F.fan_backprojection_2d_grad(img, **fan_kwargs(geo), deterministic=True))
    np.testing.assert_array_equal(a, b)
    g = O.geom(**{k: v for k, v in geo.items() if k in dict(O.Geom._fields_)})
    ref = O.fan_bpt(g, np.ascontiguousarray(randn(7, (5, geo["ny"], geo["nx"]))[2]),
                    O.fan_prep(g))
    assert rel_l2(a[2], ref) < TOL32


@pytest.mark.parametrize("angle", [False, True])
def test_fan_bp_adjoint(F, angle):
    geo, _ = load_golden("fan_c2_coarse")
    kw = fan_kwargs(geo)
    x = cuda(randn(1, (4, geo["na"], geo["ns"])))
    y = cuda(randn(2, (4, geo["ny"], geo["nx"])))
    if angle:
        angles = np.sort(np.random.default_rng(5).uniform(0, 2 * math.pi, geo["na"]))
        for k in ("orbit", "orbit_start"):
            kw.pop(k)
        A = lambda p: F.fan_backprojection_2d_angle(p, angles, **kw)  # noqa: E731
        At = lambda im: F.fan_backprojection_2d_angle_grad(im, angles, **kw)  # noqa: E731
    else:
        A = lambda p: F.fan_backprojection_2d(p, **kw)  # noqa: E731
        At = lambda im: F.fan_backprojection_2d_grad(im, **kw)  # noqa: E731
    assert adjoint_gap(A, At, x, y) < ADJ


def test_fan_angle_bp_golden(F):
    geo, d = load_golden("fan_angle_small")
    kw = fan_kwargs(geo)
    for k in ("orbit", "orbit_start"):
        kw.pop(k)
    out = host(F.fan_backprojection_2d_angle(cuda(d["proj"][None]), d["angles"], **kw))
    assert rel_l2(out[0], d["bp_f32"]) < TOL32
    out64 = host(F.fan_backprojection_2d_angle(cuda(d["proj"][None], torch.float64),
                                               d["angles"], **kw))
    assert rel_l2(out64[0], d["bp_f64"]) < TOL64


# ----------------------------------------------------------- parallel BP
def test_par_bp_golden(F):
    geo, d = load_golden("par_small")
    out = host(F.parallel_backprojection_2d(cuda(d["proj"][None]), **par_kwargs(geo)))
    assert rel_l2(out[0], d["bp_f32"]) < TOL32
    out64 = host(F.parallel_backprojection_2d(cuda(d["proj"][None], torch.float64),
                                              **par_kwargs(geo)))
    assert rel_l2(out64[0], d["bp_f64"]) < TOL64


@pytest.mark.parametrize("det", [False, True])
def test_par_bp_transpose(F, det):
    from tests.golden.make_golden import PAR_SMALL
    geo = dict(PAR_SMALL)
    g = O.geom(**geo)
    img = randn(3, (3, g.ny, g.nx))
    out = host(F.parallel_backprojection_2d_grad(cuda(img), **par_kwargs(
        dict(geo, dso=0, dsd=0)), deterministic=det))
    t = O.par_prep(g)
    for b in range(3):
        assert rel_l2(out[b], O.par_bpt(g, img[b], t)) < TOL32
    kw = par_kwargs(dict(geo, dso=0, dsd=0))
    gap = adjoint_gap(lambda p: F.parallel_backprojection_2d(p, **kw),
                      lambda im: F.parallel_backprojection_2d_grad(im, **kw, deterministic=det),
                      cuda(randn(4, (3, g.na, g.ns))), cuda(img))
    assert gap < ADJ


# ------------------------------------------------------------ ray-driven FP
@pytest.mark.parametrize("name", ["fp_small", "fp_360", "fp_720"])
def test_fp_golden(F, name):
    geo, d = load_golden(name)
    kw = par_kwargs(dict(geo, dso=0, dsd=0))
    out = host(F.parallel_projection_2d(cuda(d["img"][None]), **kw))
    assert rel_l2(out[0], d["fp_f32"]) < TOL32
    out64 = host(F.parallel_projection_2d(cuda(d["img"][None], torch.float64), **kw))
    assert rel_l2(out64[0], d["fp_f64"]) < TOL64


def test_fp_transpose_golden_and_adjoint(F):
    geo, d = load_golden("fp_small")
    kw = par_kwargs(dict(geo, dso=0, dsd=0))
    out = host(F.parallel_projection_2d_grad(cuda(d["proj"][None]), **kw))
    assert rel_l2(out[0], d["grad_f32"]) < TOL32
    out64 = host(F.parallel_projection_2d_grad(cuda(d["proj"][None], torch.float64), **kw))
    assert rel_l2(out64[0], d["grad_f64"]) < TOL64
    x = cuda(randn(8, (2, geo["ny"], geo["nx"])))
    y = cuda(randn(9, (2, geo["na"], geo["ns"])))
    gap = adjoint_gap(lambda im: F.parallel_projection_2d(im, **kw),
                      lambda p: F.parallel_projection_2d_grad(p, **kw), x, y)
    assert gap < ADJ


def test_fp_non_square_image(F):
    geo = dict(ns=61, na=33, ds=1.0, orbit=math.pi / 33, offset_s=0.2, orbit_start=0.1,
               nx=30, ny=18, dx=1.0, dy=1.3, offset_x=0.3, offset_y=-0.4)
    g = O.geom(**geo)
    img = randn(10, (2, g.ny, g.nx))
    kw = par_kwargs(dict(geo, dso=0, dsd=0))
    out = host(F.parallel_projection_2d(cuda(img), **kw))
    t = O.fp_prep(g)
    for b in range(2):
        assert rel_l2(out[b], O.fp(g, img[b], t)) < TOL32


# Batches of 3 .. 8 slices take the four-slice kernel (a partial last group
# at 3, 5, 7); the second geometry is non-square with offsets, and its wide
# detector sends rays past the image on both sides.
@pytest.mark.parametrize("nbatch", [3, 4, 5, 7, 8])
@pytest.mark.parametrize("geo", [
    dict(ns=91, na=90, ds=1.0, orbit=math.pi / 90, offset_s=0.0, orbit_start=0.0, nx=64, ny=64,
         dx=1.0, dy=1.0, offset_x=0.0, offset_y=0.0),
    dict(ns=140, na=47, ds=0.8, orbit=math.pi / 47, offset_s=0.35, orbit_start=0.2, nx=50, ny=36,
         dx=1.1, dy=0.9, offset_x=-0.6, offset_y=0.45),
], ids=["square", "offset"])
def test_fp_batches_vs_oracle(F, geo, nbatch):
    g = O.geom(**geo)
    img = randn(20 + nbatch, (nbatch, g.ny, g.nx))
    kw = par_kwargs(dict(geo, dso=0, dsd=0))
    out = host(F.parallel_projection_2d(cuda(img), **kw))
    t = O.fp_prep(g)
    for b in range(nbatch):
        assert rel_l2(out[b], O.fp(g, img[b], t)) < TOL32, b


# A detector far wider than the image: of the three 200-channel workgroups
# of a view the outer ones have every ray off the image for whole row
# chunks (their staging is skipped and the rays only advance), odd image
# sizes leave a partial last row chunk, and the view counts select one and
# four views per workgroup; 1, 2 and 5 slices take the one-, two- and
# four-slice kernels.
@pytest.mark.parametrize("nbatch", [1, 2, 5])
@pytest.mark.parametrize("na", [180, 1440])
def test_fp_wide_detector_vs_oracle(F, na, nbatch):
    geo = dict(ns=600, na=na, ds=0.5, orbit=math.pi / na, offset_s=0.3, orbit_start=0.05,
               nx=41, ny=37, dx=1.0, dy=1.1, offset_x=0.4, offset_y=-0.3)
    g = O.geom(**geo)
    img = randn(60 + nbatch, (nbatch, g.ny, g.nx))
    kw = par_kwargs(dict(geo, dso=0, dsd=0))
    out = host(F.parallel_projection_2d(cuda(img), **kw))
    t = O.fp_prep(g)
    for b in range(nbatch):
        ref = O.fp(g, img[b], t)
        assert rel_l2(out[b], ref) < TOL32, b
        # channels whose rays miss the image are exactly zero
        np.testing.assert_array_equal(out[b][ref == 0], 0.0)


# ----------------------------------------------------- filter / weighting
@pytest.mark.parametrize("typ,name", [(0, "par"), (1, "fan"), (2, "flat")])
def test_ramp_golden(F, typ, name):
    geo, d = load_golden("ramp_small")
    x = d[f"in_{typ}"]
    out = host(F.ramp_filter(cuda(x[None, :, :, None]), geo["ns"], geo["nrow"], geo["ds"],
                             geo["dsd"], name))
    assert rel_l2(out[0, :, :, 0], d[f"out_f32_{typ}"]) < TOL32
    out64 = host(F.ramp_filter(cuda(x, torch.float64), geo["ns"], geo["nrow"], geo["ds"],
                               geo["dsd"], name))
    assert rel_l2(out64, d[f"out_f64_{typ}"]) < TOL64
    outg = host(F.ramp_filter_grad(cuda(x), geo["ns"], geo["nrow"], geo["ds"], geo["dsd"], name))
    assert rel_l2(outg, d[f"grad_f32_{typ}"]) < TOL32


def test_ramp_long_rows_vs_oracle(F):
    g = O.geom(ns=1472, nrow=7, ds=0.65, dsd=1085.6, type=1)
    x = randn(11, (7, 1472))
    h = O.ramp_prep(g)
    out = host(F.ramp_filter(cuda(x), 1472, 7, 0.65, 1085.6, "fan"))
    assert rel_l2(out, O.ramp(g, x, h)) < TOL32


@pytest.mark.parametrize("typ,tp", [(1, "fan"), (2, "flat")])
def test_fan_weighting_golden(F, typ, tp):
    geo, d = load_golden("fanw_small")
    x = d[f"in_{typ}"]
    kw = dict(proj_shape=[geo["nrow"], geo["ns"]], channel_space=geo["ds"],
              channel_offset=geo["offset_s"], dso=geo["dso"], dsd=geo["dsd"], tp=tp)
    out = host(F.fan_weighting(cuda(x[None]), **kw))
    assert rel_l2(out[0], d[f"out_f32_{typ}"]) < TOL32
    out = host(F.fan_weighting_grad(cuda(x[None]), **kw))
    assert rel_l2(out[0], d[f"grad_f32_{typ}"]) < TOL32


def test_fused_fan_filter_matches_composition(F):
    ns, na = 736, 40
    x = randn(12, (2, na, ns))
    kw = dict(proj_shape=[na, ns], channel_space=1.3, channel_offset=0.25, dso=595.0,
              dsd=1085.6, tp="fan")
    fused = host(F.fan_filter(cuda(x), **kw))
    g = O.geom(ns=ns, nrow=na, ds=1.3, offset_s=0.25, dso=595.0, dsd=1085.6, type=1)
    h = O.ramp_prep(g)
    for b in range(2):
        assert rel_l2(fused[b], O.ramp(g, O.fanw(g, x[b]), h)) < TOL32


# ---------------------------------------------------------------- autograd
@pytest.mark.parametrize("det", [False, True])
def test_autograd_fan_fbp_backward_is_transpose(F, det):
    """proj.grad from autograd equals the transpose op applied to the output
    gradient: bitwise in deterministic mode, to float rounding otherwise
    (LDS float atomics do not fix the summation order)."""
    geo, _ = load_golden("fan_small")
    kw = fan_kwargs(geo)
    proj = cuda(randn(20, (2, geo["na"], geo["ns"]))).requires_grad_(True)
    img = F.fan_backprojection_2d(proj, **kw, deterministic=det)
    gimg = cuda(randn(21, tuple(img.shape)))
    img.backward(gimg)
    ref = F.fan_backprojection_2d_grad(gimg, **kw, deterministic=det)
    if det:
        np.testing.assert_array_equal(host(proj.grad), host(ref))
    else:
        assert rel_l2(host(proj.grad), host(ref)) < 1e-6


# ------------------------------------------- ramp filter on the matrix cores
# 37 rows: a partial last row block.  The BF16x3 filter takes whole rows in
# LDS: the four-wave kernel up to 1024 channels (three workgroups per CU up
# to 763, two beyond), the persistent eight-wave one up to 1536; longer rows
# take the tiled kernel.
@pytest.mark.parametrize("impl", ["mfma", "bf16x3"])
@pytest.mark.parametrize("ns", [2, 3, 33, 64, 367, 736, 763, 764, 927, 1024, 1025, 1280, 1281,
                                1472, 1536, 1537])
@pytest.mark.parametrize("typ,name", [(0, "par"), (1, "fan")])
def test_ramp_matrix_cores_vs_oracle(F, ns, typ, name, impl):
    nrow = 37
    g = O.geom(ns=ns, nrow=nrow, ds=0.65, dsd=1085.6, type=typ)
    x = randn(ns + typ, (nrow, ns))
    ref = O.ramp(g, x, O.ramp_prep(g))
    out = host(F.ramp_filter(cuda(x), ns, nrow, 0.65, 1085.6, name, impl=impl))
    assert rel_l2(out, ref) < TOL32


@pytest.mark.parametrize("ns", [736, 1000, 1472])
def test_ramp_bf16x3_many_row_blocks(F, ns):
    """More 16-row blocks than workgroups, so the persistent kernels walk
    several blocks each with the next block's rows prefetched, and a partial
    last block: rows from the first, a middle and the last blocks against
    the oracle."""
    nrow = 16 * 1100 + 5
    g = O.geom(ns=ns, nrow=nrow, ds=0.65, dsd=1085.6, type=1)
    x = randn(ns + 5, (nrow, ns))
    out = host(F.ramp_filter(cuda(x), ns, nrow, 0.65, 1085.6, "fan", impl="bf16x3"))
    h = O.ramp_prep(g)
    for r0, r1 in ((0, 20), (16 * 537 - 3, 16 * 537 + 19), (nrow - 21, nrow)):
        gs = O.geom(ns=ns, nrow=r1 - r0, ds=0.65, dsd=1085.6, type=1)
        assert rel_l2(out[r0:r1], O.ramp(gs, x[r0:r1], h)) < TOL32, (r0, r1)


def test_ramp_bf16x3_precision(F):
    """BF16 hi/lo split (3 products) on the far-field taps, centre and near
    taps in FP64: relative L2 vs the FP64 filter within the float32 bar."""
    ns, nrow = 736, 64
    g = O.geom(ns=ns, nrow=nrow, ds=1.3, dsd=1085.6, type=1)
    x = randn(77, (nrow, ns)).astype(np.float64)
    ref = O.ramp(g, x, O.ramp_prep(g))
    out = host(F.ramp_filter(cuda(x, torch.float32), ns, nrow, 1.3, 1085.6, "fan", impl="bf16x3"))
    err = rel_l2(out, ref)
    print(f"bf16x3 ramp rel L2 vs FP64: {err:.3e}")
    assert err < TOL32


@pytest.mark.parametrize("impl", ["mfma", "direct", "bf16x3"])
def test_fused_fan_filter_impls(F, impl):
    ns, na = 736, 40
    x = randn(13, (2, na, ns))
    kw = dict(proj_shape=[na, ns], channel_space=1.3, channel_offset=0.25, dso=595.0,
              dsd=1085.6, tp="fan")
    fused = host(F.fan_filter(cuda(x), **kw, impl=impl))
    g = O.geom(ns=ns, nrow=na, ds=1.3, offset_s=0.25, dso=595.0, dsd=1085.6, type=1)
    h = O.ramp_prep(g)
    for b in range(2):
        assert rel_l2(fused[b], O.ramp(g, O.fanw(g, x[b]), h)) < TOL32


def test_fan_filter_backward_is_transpose(F):
    ns, na = 128, 12
    kw = dict(proj_shape=[na, ns], channel_space=1.3, channel_offset=0.1, dso=300.0,
              dsd=500.0, tp="fan")
    x = cuda(randn(14, (2, na, ns)))
    y = cuda(randn(15, (2, na, ns)))
    gap = adjoint_gap(lambda v: F.fan_filter(v, **kw),
                      lambda v: F.fan_weighting_grad(F.ramp_filter_grad(
                          v, ns, na, 1.3, 500.0, "fan"), **kw), x, y)
    assert gap < ADJ
    xr = x.clone().requires_grad_(True)
    F.fan_filter(xr, **kw).backward(y)
    ref = F.fan_weighting_grad(F.ramp_filter_grad(y, ns, na, 1.3, 500.0, "fan"), **kw)
    assert rel_l2(host(xr.grad), host(ref)) < TOL32
