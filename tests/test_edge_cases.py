"""Degenerate inputs: empty batches (the reference's per-slice loop runs zero
times, so every op returns an empty result) and the smallest geometries (a
1 x 1 image, a single view, a two-channel detector), each against the oracle
at the float32 bar."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import randn, rel_l2

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import ct_recon_extensions_amd as F
    from ct_recon_extensions_amd import _lib
    _lib.load()
    return torch, F


def _fan_kw(g):
    return dict(img_shape=[g["ny"], g["nx"]], img_space=[g["dy"], g["dx"]],
                proj_shape=[g["na"], g["ns"]], channel_space=g["ds"],
                img_offset=[g.get("offset_y", 0.0), g.get("offset_x", 0.0)],
                channel_offset=g.get("offset_s", 0.0), orbit=g["orbit"],
                orbit_start=g.get("orbit_start", 0.0), dso=g["dso"], dsd=g["dsd"])


def test_empty_batches(env):
    torch, F = env
    kw = dict(img_shape=[16, 12], img_space=[1.0, 1.0], proj_shape=[10, 21], channel_space=1.0,
              orbit=2 * math.pi / 10, dso=100.0, dsd=180.0)
    proj = torch.empty((0, 10, 21, 1), device="cuda")
    img = torch.empty((0, 16, 12, 1), device="cuda")
    assert F.fan_backprojection_2d(proj, **kw).shape == (0, 16, 12, 1)
    assert F.fan_backprojection_2d_grad(img, **kw).shape == (0, 10, 21, 1)
    pkw = dict(kw)
    pkw.pop("dso")
    pkw.pop("dsd")
    pkw["orbit"] = math.pi / 10
    assert F.parallel_backprojection_2d(proj, **pkw).shape == (0, 16, 12, 1)
    assert F.parallel_projection_2d(img, **pkw).shape == (0, 10, 21, 1)
    assert F.parallel_projection_2d_grad(proj, **pkw).shape == (0, 16, 12, 1)
    assert F.ramp_filter(proj, 21, 10, 1.0).shape == (0, 10, 21, 1)
    fkw = dict(proj_shape=[10, 21], channel_space=1.0, dso=100.0, dsd=180.0)
    assert F.fan_weighting(proj, **fkw).shape == (0, 10, 21, 1)
    assert F.fan_filter(proj, **fkw).shape == (0, 10, 21, 1)
    assert F.parallel_single_view_bp_2d(proj, **pkw).shape == (0, 16, 12, 10)
    assert F.parallel_projection_2d(img, **pkw, fov=8.0, method="raycasting").shape == \
        (0, 10, 21, 1)


@pytest.mark.parametrize("geo", [
    dict(ns=2, na=7, ds=1.0, orbit=2 * math.pi / 7, nx=5, ny=4, dx=1.0, dy=1.0, dso=50.0,
         dsd=90.0),
    dict(ns=31, na=1, ds=1.0, orbit=2 * math.pi, orbit_start=0.4, nx=9, ny=9, dx=1.0, dy=1.0,
         dso=50.0, dsd=90.0),
    dict(ns=17, na=12, ds=1.0, orbit=2 * math.pi / 12, nx=1, ny=1, dx=2.0, dy=2.0, dso=50.0,
         dsd=90.0, offset_x=0.3),
], ids=["two_channels", "one_view", "one_pixel"])
def test_fan_bp_smallest_geometries(env, geo):
    torch, F = env
    g = O.geom(**geo)
    proj = randn(3, (geo["na"], geo["ns"]))
    got = F.fan_backprojection_2d(torch.from_numpy(proj)[None].cuda(), **_fan_kw(geo))
    want = O.fan_bp(g, proj, O.fan_prep(g))
    assert rel_l2(got[0].cpu().numpy(), want) < TOL32
    img = randn(4, (geo["ny"], geo["nx"]))
    gt = F.fan_backprojection_2d_grad(torch.from_numpy(img)[None].cuda(), **_fan_kw(geo))
    wt = O.fan_bpt(g, img, O.fan_prep(g))
    assert rel_l2(gt[0].cpu().numpy(), wt) < TOL32


@pytest.mark.parametrize("geo", [
    dict(ns=2, na=5, ds=1.0, orbit=math.pi / 5, nx=3, ny=3, dx=1.0, dy=1.0),
    dict(ns=19, na=1, ds=1.0, orbit=math.pi, orbit_start=0.3, nx=8, ny=8, dx=1.0, dy=1.0),
], ids=["two_channels", "one_view"])
def test_fp_and_par_bp_smallest_geometries(env, geo):
    torch, F = env
    g = O.geom(**geo)
    kw = dict(img_shape=[geo["ny"], geo["nx"]], img_space=[geo["dy"], geo["dx"]],
              proj_shape=[geo["na"], geo["ns"]], channel_space=geo["ds"], orbit=geo["orbit"],
              orbit_start=geo.get("orbit_start", 0.0))
    img = randn(5, (geo["ny"], geo["nx"]))
    fp = F.parallel_projection_2d(torch.from_numpy(img)[None].cuda(), **kw)[0].cpu().numpy()
    assert rel_l2(fp, O.fp(g, img, O.fp_prep(g))) < TOL32
    proj = randn(6, (geo["na"], geo["ns"]))
    bp = F.parallel_backprojection_2d(torch.from_numpy(proj)[None].cuda(), **kw)[0].cpu().numpy()
    assert rel_l2(bp, O.par_bp(g, proj, O.par_prep(g))) < TOL32


@pytest.mark.parametrize("geo", [
    dict(ns=1, na=6, ds=1.0, orbit=math.pi / 6, nx=7, ny=5, dx=1.0, dy=1.0, fov=-1.0),
    dict(ns=23, na=1, ds=0.9, orbit=math.pi, orbit_start=0.7, nx=9, ny=11, dx=1.1, dy=0.8,
         fov=6.0),
    dict(ns=15, na=9, ds=1.0, orbit=math.pi / 9, nx=1, ny=1, dx=2.0, dy=2.0, fov=3.0,
         offset_s=0.2),
], ids=["one_channel", "one_view", "one_pixel"])
def test_single_view_bp_and_raycasting_smallest_geometries(env, geo):
    """Bit-identical to the oracle on the smallest geometries (a single
    channel, a single view, a 1 x 1 image; the default fov of -1 leaves only
    the rays within one unit of the centre)."""
    torch, F = env
    g = O.geom(**geo)
    kw = dict(img_shape=[geo["ny"], geo["nx"]], img_space=[geo["dy"], geo["dx"]],
              proj_shape=[geo["na"], geo["ns"]], channel_space=geo["ds"], orbit=geo["orbit"],
              orbit_start=geo.get("orbit_start", 0.0), channel_offset=geo.get("offset_s", 0.0))
    proj = randn(7, (geo["na"], geo["ns"]))
    sv = F.parallel_single_view_bp_2d(torch.from_numpy(proj).cuda(), **kw).cpu().numpy()
    np.testing.assert_array_equal(sv, O.par_svbp(g, proj, O.par_prep(g)))
    img = randn(8, (geo["ny"], geo["nx"]))
    rc = F.parallel_projection_2d(torch.from_numpy(img).cuda(), **kw, fov=geo["fov"],
                                  method="raycasting").cpu().numpy()
    np.testing.assert_array_equal(rc, O.fp_rc(g, img, O.fp_rc_prep(g)))
