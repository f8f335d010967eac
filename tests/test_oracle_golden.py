"""Pin the C restatement (oracle/ct_oracle.c) to the reference's own outputs.

The golden vectors were written by the reference's src/*.cc compiled
unchanged (tests/golden/make_golden.py).  FP64 paths are compared bitwise;
float paths bitwise too, since the oracle keeps the reference's evaluation
order and accumulation type.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import load_golden, randn


def G(geo):
    return O.geom(**geo)


def test_fan_small_tables_and_bp():
    geo, d = load_golden("fan_small")
    g = G(geo)
    t = O.fan_prep(g)
    for a, k in zip(t, ("xpos", "ypos", "sincostbl")):
        np.testing.assert_array_equal(a, d[k])
    np.testing.assert_array_equal(O.fan_bp(g, d["proj"], t), d["bp_f32"])
    np.testing.assert_array_equal(O.fan_bp(g, d["proj"].astype(np.float64), t), d["bp_f64"])
    np.testing.assert_array_equal(O.fan_bpt(g, d["img"], t), d["bpt_f32"])
    np.testing.assert_array_equal(O.fan_bpt(g, d["img"].astype(np.float64), t), d["bpt_f64"])
    # The geometry must exercise the detector edge: some pixel·views fall off.
    assert np.count_nonzero(d["bp_f32"]) == d["bp_f32"].size


def test_fan_c2_coarse():
    geo, d = load_golden("fan_c2_coarse")
    g = G(geo)
    t = O.fan_prep(g)
    np.testing.assert_array_equal(t[2], d["sincostbl"])
    proj = randn(int(d["proj_seed"]), (g.na, g.ns))
    np.testing.assert_array_equal(O.fan_bp(g, proj, t), d["bp_f32"])
    np.testing.assert_array_equal(O.fan_bp(g, proj.astype(np.float64), t), d["bp_f64"])
    np.testing.assert_array_equal(O.fan_bpt(g, d["img"], t)[::60], d["bpt_f32_every60"])


def test_fan_angle_small():
    geo, d = load_golden("fan_angle_small")
    g = G(geo)
    t = O.fan_angle_prep(g, d["angles"])
    for a, k in zip(t, ("xpos", "ypos", "sincostbl")):
        np.testing.assert_array_equal(a, d[k])
    np.testing.assert_array_equal(O.fan_bp(g, d["proj"], t, scaled=False), d["bp_f32"])
    np.testing.assert_array_equal(
        O.fan_bp(g, d["proj"].astype(np.float64), t, scaled=False), d["bp_f64"])


def test_par_small():
    geo, d = load_golden("par_small")
    g = G(geo)
    t = O.par_prep(g)
    np.testing.assert_array_equal(t[0], d["xcos"])
    np.testing.assert_array_equal(t[1], d["ysin"])
    np.testing.assert_array_equal(O.par_bp(g, d["proj"], t), d["bp_f32"])
    np.testing.assert_array_equal(O.par_bp(g, d["proj"].astype(np.float64), t), d["bp_f64"])


def test_svbp_small():
    """Single-view parallel BP against the compiled reference's outputs."""
    geo, d = load_golden("svbp_small")
    g = G(geo)
    t = O.par_svbp_prep(g)
    np.testing.assert_array_equal(t[0], d["xcos"])
    np.testing.assert_array_equal(t[1], d["ysin"])
    np.testing.assert_array_equal(O.par_svbp(g, d["proj"], t), d["svbp_f32"])
    np.testing.assert_array_equal(O.par_svbp(g, d["proj"].astype(np.float64), t), d["svbp_f64"])


def test_rc_small():
    """Ray-casting FP against the compiled reference's outputs."""
    geo, d = load_golden("rc_small")
    g = G(geo)
    t = O.fp_rc_prep(g)
    np.testing.assert_array_equal(t[0], d["sincostbl"])
    np.testing.assert_array_equal(t[1], d["begins"])
    np.testing.assert_array_equal(t[2], d["nsteps"])
    np.testing.assert_array_equal(O.fp_rc(g, d["img"], t), d["fp_f32"])
    np.testing.assert_array_equal(O.fp_rc(g, d["img"].astype(np.float64), t), d["fp_f64"])


def test_fp_small():
    geo, d = load_golden("fp_small")
    g = G(geo)
    t = O.fp_prep(g)
    np.testing.assert_array_equal(t[0], d["sincostbl"])
    np.testing.assert_array_equal(t[1], d["beginoffset"])
    np.testing.assert_array_equal(t[2], d["usex"])
    np.testing.assert_array_equal(O.fp(g, d["img"], t), d["fp_f32"])
    np.testing.assert_array_equal(O.fp(g, d["img"].astype(np.float64), t), d["fp_f64"])
    tg = O.fp_grad_prep(g)
    np.testing.assert_array_equal(tg[0], d["weights"])
    np.testing.assert_array_equal(tg[1][: d["pos"].size], d["pos"])
    np.testing.assert_array_equal(tg[2], d["usex_grad"])
    np.testing.assert_array_equal(O.fp_grad(g, d["proj"], tg), d["grad_f32"])
    np.testing.assert_array_equal(O.fp_grad(g, d["proj"].astype(np.float64), tg), d["grad_f64"])


@pytest.mark.parametrize("name,boundary_view,expect_usex", [
    ("fp_360", 270, 0),   # accumulated angle falls just below 3pi/4 -> y-branch
    ("fp_720", 180, 1),   # accumulated angle falls just below pi/4 -> x-branch
])
def test_fp_branch_boundary(name, boundary_view, expect_usex):
    geo, d = load_golden(name)
    g = G(geo)
    t = O.fp_prep(g)
    np.testing.assert_array_equal(t[2], d["usex"])
    assert d["usex"][boundary_view] == expect_usex
    np.testing.assert_array_equal(O.fp(g, d["img"], t), d["fp_f32"])
    np.testing.assert_array_equal(O.fp(g, d["img"].astype(np.float64), t), d["fp_f64"])


@pytest.mark.parametrize("typ", [0, 1, 2])
def test_ramp_small(typ):
    geo, d = load_golden("ramp_small")
    g = O.geom(**{**geo, "type": typ})
    h = O.ramp_prep(g)
    np.testing.assert_array_equal(h, d[f"taps_{typ}"])
    x = d[f"in_{typ}"]
    np.testing.assert_array_equal(O.ramp(g, x, h), d[f"out_f32_{typ}"])
    np.testing.assert_array_equal(O.ramp(g, x.astype(np.float64), h), d[f"out_f64_{typ}"])
    np.testing.assert_array_equal(O.ramp(g, x, h, grad=True), d[f"grad_f32_{typ}"])
    np.testing.assert_array_equal(O.ramp(g, x.astype(np.float64), h, grad=True),
                                  d[f"grad_f64_{typ}"])


def test_ramp_unknown_type_fails():
    g = O.geom(ns=8, nrow=1, ds=1.0, dsd=10.0, type=7)
    with pytest.raises(RuntimeError):
        O.ramp_prep(g)


@pytest.mark.parametrize("typ", [1, 2])
def test_fanw_small(typ):
    geo, d = load_golden("fanw_small")
    g = O.geom(**{**geo, "type": typ})
    x = d[f"in_{typ}"]
    np.testing.assert_array_equal(O.fanw(g, x), d[f"out_f32_{typ}"])
    np.testing.assert_array_equal(O.fanw(g, x.astype(np.float64)), d[f"out_f64_{typ}"])
    np.testing.assert_array_equal(O.fanw(g, x, grad=True), d[f"grad_f32_{typ}"])


@pytest.mark.parametrize("name", ["fpdd_small", "fpdd_circle"])
def test_fp_dd(name):
    """Distance-driven FP and its gradient: tables and outputs bitwise.  Both
    geometries have rows whose footprint starts left of the first pixel (the
    reference's unsigned index skips them)."""
    geo, d = load_golden(name)
    g = G(geo)
    t = O.fp_dd_prep(g)
    for a, k in zip(t, ("sincostbl", "beginoffset", "usex")):
        np.testing.assert_array_equal(a, d[k])
    tg = O.fp_dd_grad_prep(g)
    np.testing.assert_array_equal(tg[0], d["weights"])
    np.testing.assert_array_equal(tg[1][:d["pos"].size], d["pos"])
    np.testing.assert_array_equal(tg[2], d["usex_grad"])
    np.testing.assert_array_equal(O.fp_dd(g, d["img"], t), d["fp_f32"])
    np.testing.assert_array_equal(O.fp_dd(g, d["img"].astype(np.float64), t), d["fp_f64"])
    np.testing.assert_array_equal(O.fp_dd_grad(g, d["proj"], tg), d["grad_f32"])
    np.testing.assert_array_equal(O.fp_dd_grad(g, d["proj"].astype(np.float64), tg),
                                  d["grad_f64"])
    bo = d["beginoffset"].reshape(g.na, g.ns, 3)
    left = np.minimum(bo[..., 0] - bo[..., 2], bo[..., 0] + bo[..., 2])
    off = bo[..., 1]
    n_rows = np.where(d["usex"][:, None] != 0, g.ny, g.nx)
    # some footprint leaves the image on the left within the scan's rows
    assert np.any((left < 0) | (left + off * (n_rows - 1) < 0))


@pytest.mark.skipif(not O.have_ref(), reason="compiled reference not built here")
def test_restatement_matches_compiled_reference_fresh_seeds():
    """Beyond the stored vectors: fresh seeds, every operator, bitwise."""
    from tests.golden.make_golden import FAN_SMALL, PAR_SMALL, FP_SMALL
    g = O.geom(**FAN_SMALL)
    t = O.fan_prep(g)
    for seed in range(3):
        p = randn(1000 + seed, (g.na, g.ns))
        np.testing.assert_array_equal(O.fan_bp(g, p, t), O.fan_bp(g, p, t, kind="ref"))
        im = randn(2000 + seed, (g.ny, g.nx))
        np.testing.assert_array_equal(O.fan_bpt(g, im, t), O.fan_bpt(g, im, t, kind="ref"))
    g = O.geom(**PAR_SMALL)
    t = O.par_prep(g)
    p = randn(3000, (g.na, g.ns))
    np.testing.assert_array_equal(O.par_bp(g, p, t), O.par_bp(g, p, t, kind="ref"))
    g = O.geom(**FP_SMALL)
    t = O.fp_prep(g)
    im = randn(4000, (g.ny, g.nx))
    np.testing.assert_array_equal(O.fp(g, im, t), O.fp(g, im, t, kind="ref"))
    from tests.golden.make_golden import FPDD_CIRCLE
    for geo in (FP_SMALL, FPDD_CIRCLE):
        g = O.geom(**geo)
        t, tg = O.fp_dd_prep(g), O.fp_dd_grad_prep(g)
        im = randn(5000, (g.ny, g.nx))
        np.testing.assert_array_equal(O.fp_dd(g, im, t), O.fp_dd(g, im, t, kind="ref"))
        p = randn(5001, (g.na, g.ns))
        np.testing.assert_array_equal(O.fp_dd_grad(g, p, tg), O.fp_dd_grad(g, p, tg, kind="ref"))
