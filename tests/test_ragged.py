"""Ragged persistent batches of bp_fan_2d_angle problems.

GPU: every problem of a ragged batch against the oracle (relative L2 <= 1e-5,
the float32 bar), for every tile shape; the full 64-problem workload
against the regular per-problem launch (the same tile engine, pinned to the
oracle elsewhere) and, on its smallest problems, against the oracle; graph
replays against eager runs (bitwise, with other GPU work in between); the
autotuner and its cache; the load-balance probe.  CPU: the seeded workload
generator.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2

TOL32 = 1e-5

# Small problems that together cover: non-square images, pixel and detector
# offsets, a source close to the image (exact atan mode, wide spans), each
# staged span class (64 / 128 / 256 channels), views off [0, 2 pi), unsorted
# angles, and a near-empty problem.  Every tile shape admits all of them
# (span <= 256 channels); WIDE does not admit 32x8 tiles.
SMALL = [
    dict(ns=64, na=37, ds=1.7, offset_s=0.4, nx=24, ny=20, dx=1.3, dy=1.5, offset_x=0.3,
         offset_y=-0.2, dso=60.0, dsd=110.0),
    dict(ns=96, na=50, ds=2.0, offset_s=-1.25, nx=40, ny=40, dx=3.0, dy=3.0, offset_x=0.0,
         offset_y=0.0, dso=300.0, dsd=550.0),
    dict(ns=200, na=45, ds=0.5, offset_s=0.0, nx=64, ny=48, dx=1.0, dy=1.0, offset_x=-2.5,
         offset_y=1.0, dso=120.0, dsd=240.0),
    dict(ns=300, na=29, ds=0.25, offset_s=0.7, nx=32, ny=32, dx=0.6, dy=0.6, offset_x=0.0,
         offset_y=0.0, dso=100.0, dsd=200.0),
    dict(ns=16, na=3, ds=1.0, offset_s=0.0, nx=8, ny=8, dx=0.5, dy=0.5, offset_x=0.0,
         offset_y=0.0, dso=50.0, dsd=90.0),
    dict(ns=128, na=61, ds=1.1, offset_s=0.0, nx=56, ny=72, dx=0.9, dy=0.8, offset_x=0.5,
         offset_y=0.5, dso=400.0, dsd=700.0),
]


WIDE = dict(ns=300, na=29, ds=0.25, offset_s=0.7, nx=32, ny=32, dx=1.2, dy=1.2, offset_x=0.0,
            offset_y=0.0, dso=100.0, dsd=200.0)


def _angles(i, na):
    rng = np.random.default_rng(100 + i)
    a = rng.uniform(-1.0, 2 * math.pi + 1.0, na)  # irregular, partly outside [0, 2 pi)
    return a if i % 2 else np.sort(a)


def _params(geos):
    from ct_recon_extensions_amd.geometry import FanBackprojection2DAngleParam as P
    return [P(g["ns"], g["na"], g["ds"], g["offset_s"], g["nx"], g["ny"], g["dx"], g["dy"],
              g["offset_x"], g["offset_y"], g["dso"], g["dsd"]) for g in geos]


def _oracle(g, proj, angles):
    og = O.geom(**g)
    return O.fan_bp(og, proj, O.fan_angle_prep(og, angles), scaled=False)


@pytest.fixture(scope="module")
def R():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from ct_recon_extensions_amd import _lib, ragged
    _lib.load()
    return ragged


# ------------------------------------------------------------------- CPU
def test_workload_generator_is_seeded_and_fits_the_fan():
    from ct_recon_extensions_amd.ragged import random_fan_problems
    p1, a1 = random_fan_problems(64, seed=7)
    p2, a2 = random_fan_problems(64, seed=7)
    assert p1 == p2 and all(np.array_equal(x, y) for x, y in zip(a1, a2))
    for p, a in zip(p1, a1):
        assert p.nx == p.ny and p.nx in (128, 256, 384, 512)
        assert 180 <= p.na <= 1440 and 256 <= p.ns <= 1024
        assert 0.5 <= p.dso / p.dsd <= 0.6
        # the image's circumscribed circle lies inside the fan's field of view
        half_fan = (p.ns / 2) * p.ds / p.dsd
        assert p.nx * p.dx / math.sqrt(2) <= p.dso * math.sin(half_fan)
        assert a.shape == (p.na,) and np.all(np.diff(a) >= 0)
        assert a[0] >= 0 and a[-1] < 2 * math.pi


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("tile,vchunk", [(0, 16), (0, 64), (1, 8), (2, 32)])
def test_ragged_matches_oracle(R, tile, vchunk):
    import torch
    params = _params(SMALL)
    angles = [_angles(i, g["na"]) for i, g in enumerate(SMALL)]
    rb = R.RaggedFanBackprojection2DAngle(params, angles, tile=tile, vchunk=vchunk)
    projs = [np.random.default_rng(i).standard_normal((g["na"], g["ns"])).astype(np.float32)
             for i, g in enumerate(SMALL)]
    out = rb([torch.from_numpy(p).cuda() for p in projs], graph=False)
    torch.cuda.synchronize()
    for i, g in enumerate(SMALL):
        want = _oracle(g, projs[i], angles[i])
        got = out[i].cpu().numpy()
        assert rel_l2(got, want) < TOL32, (tile, vchunk, i, rel_l2(got, want))
    rb.close()


@pytest.mark.gpu
def test_ragged_graph_replays_match_eager(R):
    import torch
    params = _params(SMALL)
    angles = [_angles(i, g["na"]) for i, g in enumerate(SMALL)]
    rb = R.RaggedFanBackprojection2DAngle(params, angles, tile=0, vchunk=16)
    gen = torch.Generator(device="cuda").manual_seed(5)
    for rep in range(4):
        rb.proj_flat.normal_(generator=gen)
        eager = [t.clone() for t in rb(graph=False)]
        rb.img_flat.fill_(float("nan"))
        side = torch.randn(1 << 20, device="cuda").sum()  # unrelated eager work
        graphed = rb(graph=True)
        torch.cuda.synchronize()
        assert torch.isfinite(side)
        for e, g in zip(eager, graphed):
            assert torch.equal(e, g), rep
    rb.close()


@pytest.mark.gpu
def test_ragged_full_workload_matches_per_problem_launches(R):
    import torch
    import ct_recon_extensions_amd as F
    params, angles = R.random_fan_problems(64, seed=20261204)
    rb = R.RaggedFanBackprojection2DAngle(params, angles, tile=0, vchunk=16)
    rb.proj_flat.normal_(generator=torch.Generator(device="cuda").manual_seed(11))
    out = [t.clone() for t in rb(graph=True)]
    torch.cuda.synchronize()
    worst = 0.0
    for p, a, x, got in zip(params, angles, rb.projs, out):
        ref = F.fan_backprojection_2d_angle(
            x[None], a, img_shape=[p.ny, p.nx], img_space=[p.dy, p.dx],
            proj_shape=[p.na, p.ns], channel_space=p.ds, dso=p.dso, dsd=p.dsd)[0]
        err = float((got.double() - ref.double()).norm() / ref.double().norm())
        worst = max(worst, err)
    assert worst < 1e-6, worst
    # the smallest problems against the oracle itself
    order = sorted(range(len(params)), key=lambda i: params[i].na * params[i].nx ** 2)[:3]
    for i in order:
        p = params[i]
        g = dict(ns=p.ns, na=p.na, ds=p.ds, offset_s=p.offset_s, nx=p.nx, ny=p.ny, dx=p.dx,
                 dy=p.dy, offset_x=p.offset_x, offset_y=p.offset_y, dso=p.dso, dsd=p.dsd)
        want = _oracle(g, rb.projs[i].cpu().numpy(), angles[i])
        assert rel_l2(out[i].cpu().numpy(), want) < TOL32, i
    rb.close()


@pytest.mark.gpu
def test_ragged_autotune_and_cache(R, tmp_path):
    params = _params(SMALL)
    angles = [_angles(i, g["na"]) for i, g in enumerate(SMALL)]
    cache = str(tmp_path / "tune.json")
    rb = R.RaggedFanBackprojection2DAngle(params, angles, cache_path=cache, tune_reps=2)
    assert rb.tile in (0, 1, 2) and rb.vchunk in R.CHUNKS
    assert len(rb.tune_table) == 12 and min(v for v in rb.tune_table.values() if v > 0) > 0
    best = min(v for v in rb.tune_table.values() if v > 0)
    assert rb.tune_table[f"{R.TILES[rb.tile]}/{rb.vchunk}"] == best
    rb2 = R.RaggedFanBackprojection2DAngle(params, angles, cache_path=cache, autotune=False)
    assert (rb2.tile, rb2.vchunk) == (rb.tile, rb.vchunk)
    rb.close()
    rb2.close()


@pytest.mark.gpu
def test_ragged_balance_and_errors(R):
    from ct_recon_extensions_amd import _lib
    params = _params(SMALL)
    angles = [_angles(i, g["na"]) for i, g in enumerate(SMALL)]
    rb = R.RaggedFanBackprojection2DAngle(params, angles, tile=1, vchunk=8)
    b = rb.balance()
    assert 0.0 < b <= 1.0
    assert rb.nitems == sum(math.ceil(g["nx"] / 8) * math.ceil(g["ny"] / 8) for g in SMALL)
    rb.close()
    with pytest.raises(_lib.CtrError):
        R.RaggedFanBackprojection2DAngle(params, angles, tile=3, vchunk=8)
    with pytest.raises(_lib.CtrError):
        R.RaggedFanBackprojection2DAngle(params, angles, tile=0, vchunk=65)


@pytest.mark.gpu
def test_ragged_rejects_tiles_with_spans_over_256_channels(R):
    import torch
    from ct_recon_extensions_amd import _lib
    params, angles = _params([WIDE]), [_angles(3, WIDE["na"])]
    with pytest.raises(_lib.CtrError, match="256 channels"):
        R.RaggedFanBackprojection2DAngle(params, angles, tile=2, vchunk=16)
    rb = R.RaggedFanBackprojection2DAngle(params, angles, tile=1, vchunk=16)
    proj = np.random.default_rng(3).standard_normal((WIDE["na"], WIDE["ns"])).astype(np.float32)
    got = rb([torch.from_numpy(proj).cuda()], graph=False)[0].cpu().numpy()
    assert rel_l2(got, _oracle(WIDE, proj, angles[0])) < TOL32
    rb.close()
