"""Write golden vectors for the 2-D CT hot path from the compiled reference.

Run in a container that has the reference checkout:

    make -C oracle ref          # builds oracle/_ref/libctref.so from src/*.cc
    python tests/golden/make_golden.py

Every fixture is an ``.npz`` of inputs and the reference's outputs (its own
``calculate_on_cpu`` through ``oracle/ref_shim.cpp``).  Large inputs are not
stored: they are regenerated from the recorded seed with
``numpy.random.default_rng(seed).standard_normal(shape)``, the same call the
tests make.  Geometries are chosen so that every edge rule of the reference
is exercised (pixels projecting outside the detector, offsets on every axis,
the pi/4 branch boundary of the ray-driven FP under accumulated angles) while
staying inside the region where the reference is well defined.
"""
from __future__ import annotations

import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from oracle import oracle as O  # noqa: E402


def randn(seed, shape, dtype=np.float32):
    return np.random.default_rng(seed).standard_normal(shape).astype(dtype)


def gdict(g):
    return {k: getattr(g, k) for k, _ in O.Geom._fields_}


def save(name, g, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, geom=json.dumps(gdict(g)), **arrays)
    print(f"wrote {path} ({os.path.getsize(path)} B)")


# Geometries ---------------------------------------------------------------
# Small fan geometry: part of the image projects off the detector.
FAN_SMALL = dict(ns=41, na=48, ds=1.7, orbit=2 * math.pi / 48, offset_s=0.4,
                 orbit_start=0.3, nx=24, ny=20, dx=1.3, dy=1.5, offset_x=0.3,
                 offset_y=-0.2, dso=60.0, dsd=110.0)
# The 512x512 / 720-view fan geometry (SID 595, SDD 1085.6, 736 channels at
# 1.30 mm arc pitch) sampled on a coarse 64x64 grid covering the same FOV.
FAN_C2_COARSE = dict(ns=736, na=720, ds=1.30, orbit=2 * math.pi / 720, offset_s=0.0,
                     orbit_start=0.0, nx=64, ny=64, dx=0.7 * 8, dy=0.7 * 8,
                     offset_x=0.0, offset_y=0.0, dso=595.0, dsd=1085.6)
PAR_SMALL = dict(ns=41, na=36, ds=0.8, orbit=math.pi / 36, offset_s=-0.35,
                 orbit_start=0.1, nx=22, ny=26, dx=0.7, dy=0.9, offset_x=-0.4,
                 offset_y=0.25)
# Single-view parallel BP: a detector narrower than the image, so that
# samples fall past both of its ends (zeros in the output).
SVBP_SMALL = dict(ns=15, na=12, ds=0.8, orbit=math.pi / 12, offset_s=0.4,
                  orbit_start=0.2, nx=18, ny=14, dx=0.7, dy=0.9, offset_x=0.3,
                  offset_y=-0.2)
# Ray-casting FP: a FOV circle wider than the image (samples past its edges)
# and rays that miss the circle.
RC_SMALL = dict(ns=61, na=48, ds=1.1, orbit=math.pi / 48, offset_s=0.3, orbit_start=0.05,
                nx=40, ny=36, dx=1.0, dy=1.0, offset_x=0.2, offset_y=-0.3, fov=24.0)
# Ray-driven FP: nx == ny (the reference's transpose requires it); the
# detector covers the image so the reference's transpose stays defined.
FP_SMALL = dict(ns=41, na=40, ds=1.0, orbit=math.pi / 40, offset_s=0.3,
                orbit_start=0.0, nx=24, ny=24, dx=1.0, dy=1.0, offset_x=0.2,
                offset_y=-0.1)
# 360 views over 180 deg (view 270 sits on 3pi/4) and 720 views over 180 deg
# (view 180 sits on pi/4): the branch is decided on the accumulated angle.
FP_360 = dict(ns=25, na=360, ds=1.0, orbit=math.pi / 360, offset_s=0.0,
              orbit_start=0.0, nx=16, ny=16, dx=1.0, dy=1.0)
FP_720 = dict(ns=25, na=720, ds=1.0, orbit=math.pi / 720, offset_s=0.0,
              orbit_start=0.0, nx=16, ny=16, dx=1.0, dy=1.0)


def fan_fixtures():
    g = O.geom(**FAN_SMALL)
    t = O.fan_prep(g, "ref")
    proj = randn(11, (g.na, g.ns))
    img = randn(12, (g.ny, g.nx))
    save("fan_small", g, proj=proj, img=img, xpos=t[0], ypos=t[1], sincostbl=t[2],
         bp_f32=O.fan_bp(g, proj, t, kind="ref"),
         bp_f64=O.fan_bp(g, proj.astype(np.float64), t, kind="ref"),
         bpt_f32=O.fan_bpt(g, img, t, kind="ref"),
         bpt_f64=O.fan_bpt(g, img.astype(np.float64), t, kind="ref"))

    g = O.geom(**FAN_C2_COARSE)
    t = O.fan_prep(g, "ref")
    proj = randn(21, (g.na, g.ns))
    img = randn(22, (g.ny, g.nx))
    save("fan_c2_coarse", g, proj_seed=np.int64(21), img=img, sincostbl=t[2],
         bp_f32=O.fan_bp(g, proj, t, kind="ref"),
         bp_f64=O.fan_bp(g, proj.astype(np.float64), t, kind="ref"),
         bpt_f32_every60=O.fan_bpt(g, img, t, kind="ref")[::60].copy())

    # Irregular, sorted angles (sparse-view style subset plus jitter).
    g = O.geom(**{**FAN_SMALL, "na": 37})
    rng = np.random.default_rng(31)
    angles = np.sort(rng.uniform(0, 2 * math.pi, g.na))
    t = O.fan_angle_prep(g, angles, "ref")
    proj = randn(32, (g.na, g.ns))
    save("fan_angle_small", g, angles=angles, proj=proj, xpos=t[0], ypos=t[1],
         sincostbl=t[2],
         bp_f32=O.fan_bp(g, proj, t, scaled=False, kind="ref"),
         bp_f64=O.fan_bp(g, proj.astype(np.float64), t, scaled=False, kind="ref"))


def par_fixtures():
    g = O.geom(**PAR_SMALL)
    t = O.par_prep(g, "ref")
    proj = randn(41, (g.na, g.ns))
    save("par_small", g, proj=proj, xcos=t[0], ysin=t[1],
         bp_f32=O.par_bp(g, proj, t, kind="ref"),
         bp_f64=O.par_bp(g, proj.astype(np.float64), t, kind="ref"))


def svbp_fixtures():
    g = O.geom(**SVBP_SMALL)
    t = O.par_svbp_prep(g, "ref")
    proj = randn(71, (g.na, g.ns))
    save("svbp_small", g, proj=proj, xcos=t[0], ysin=t[1],
         svbp_f32=O.par_svbp(g, proj, t, kind="ref"),
         svbp_f64=O.par_svbp(g, proj.astype(np.float64), t, kind="ref"))


def rc_fixtures():
    g = O.geom(**RC_SMALL)
    t = O.fp_rc_prep(g, kind="ref")
    img = randn(81, (g.ny, g.nx))
    save("rc_small", g, img=img, sincostbl=t[0], begins=t[1], nsteps=t[2],
         fp_f32=O.fp_rc(g, img, t, kind="ref"),
         fp_f64=O.fp_rc(g, img.astype(np.float64), t, kind="ref"))


def fp_fixtures():
    g = O.geom(**FP_SMALL)
    t = O.fp_prep(g, "ref")
    tg = O.fp_grad_prep(g, "ref")
    img = randn(51, (g.ny, g.nx))
    proj = randn(52, (g.na, g.ns))
    save("fp_small", g, img=img, proj=proj, sincostbl=t[0], beginoffset=t[1], usex=t[2],
         weights=tg[0], pos=tg[1], usex_grad=tg[2],
         fp_f32=O.fp(g, img, t, kind="ref"),
         fp_f64=O.fp(g, img.astype(np.float64), t, kind="ref"),
         grad_f32=O.fp_grad(g, proj, tg, kind="ref"),
         grad_f64=O.fp_grad(g, proj.astype(np.float64), tg, kind="ref"))
    for name, geo, seed in (("fp_360", FP_360, 61), ("fp_720", FP_720, 62)):
        g = O.geom(**geo)
        t = O.fp_prep(g, "ref")
        img = randn(seed, (g.ny, g.nx))
        save(name, g, img=img, usex=t[2], fp_f32=O.fp(g, img, t, kind="ref"),
             fp_f64=O.fp(g, img.astype(np.float64), t, kind="ref"))


# Distance-driven FP: the ray-driven fixture's geometry, and a full circle
# from a negative start angle (the FP tables wrap the angle, the gradient's
# tables do not) on a square image whose projections stay on the detector
# (the reference's gradient reads past the row otherwise); both have rows
# whose footprint starts left of the first pixel.
FPDD_CIRCLE = dict(ns=70, na=72, ds=0.7, orbit=2 * math.pi / 72, offset_s=0.25,
                   orbit_start=-0.3, nx=32, ny=32, dx=1.0, dy=1.0, offset_x=0.0,
                   offset_y=0.0)


def fpdd_fixtures():
    for name, geo, seed in (("fpdd_small", FP_SMALL, 91), ("fpdd_circle", FPDD_CIRCLE, 93)):
        g = O.geom(**geo)
        t = O.fp_dd_prep(g, "ref")
        tg = O.fp_dd_grad_prep(g, "ref")
        img = randn(seed, (g.ny, g.nx))
        proj = randn(seed + 1, (g.na, g.ns))
        save(name, g, img=img, proj=proj, sincostbl=t[0], beginoffset=t[1], usex=t[2],
             weights=tg[0], pos=tg[1], usex_grad=tg[2],
             fp_f32=O.fp_dd(g, img, t, kind="ref"),
             fp_f64=O.fp_dd(g, img.astype(np.float64), t, kind="ref"),
             grad_f32=O.fp_dd_grad(g, proj, tg, kind="ref"),
             grad_f64=O.fp_dd_grad(g, proj.astype(np.float64), tg, kind="ref"))


def filter_fixtures():
    out = {}
    for typ in (0, 1, 2):
        g = O.geom(ns=33, nrow=5, ds=0.9, dsd=40.0, type=typ)
        h = O.ramp_prep(g, "ref")
        x = randn(70 + typ, (g.nrow, g.ns))
        out[f"taps_{typ}"] = h
        out[f"in_{typ}"] = x
        out[f"out_f32_{typ}"] = O.ramp(g, x, h, kind="ref")
        out[f"out_f64_{typ}"] = O.ramp(g, x.astype(np.float64), h, kind="ref")
        out[f"grad_f32_{typ}"] = O.ramp(g, x, h, grad=True, kind="ref")
        out[f"grad_f64_{typ}"] = O.ramp(g, x.astype(np.float64), h, grad=True, kind="ref")
    save("ramp_small", O.geom(ns=33, nrow=5, ds=0.9, dsd=40.0), **out)

    out = {}
    for typ in (1, 2):
        g = O.geom(ns=37, nrow=4, ds=1.3, offset_s=0.6, dso=595.0, dsd=1085.6, type=typ)
        x = randn(80 + typ, (g.nrow, g.ns))
        out[f"in_{typ}"] = x
        out[f"out_f32_{typ}"] = O.fanw(g, x, kind="ref")
        out[f"out_f64_{typ}"] = O.fanw(g, x.astype(np.float64), kind="ref")
        out[f"grad_f32_{typ}"] = O.fanw(g, x, grad=True, kind="ref")
    save("fanw_small", O.geom(ns=37, nrow=4, ds=1.3, offset_s=0.6, dso=595.0, dsd=1085.6),
         **out)


# Helical rebinning (src/helical_interpolation.cc): a pitch-1 scan whose
# windows leave the detector near +-1 rotation (the single-window branch) and
# whose channel offset puts 180LI conjugate channels off the detector; a
# pitch-1.7 scan with thick slices (several rows averaged), a row offset and
# descending slices; a pitch-3 scan with thin slices where some samples see
# no row in either window (the reference throws); a slice too close to the
# scan start (the reference throws).  The projections are regenerated from
# proj_seed (float32 draws, cast for the float64 run).
HELI_COLL = 6 * 595.0 / 1085.6  # 6 rows of 1 mm at the detector, at iso
HELI_SMALL = dict(ns=29, nt=6, na=96, ds=12.0, dt=1.0, offset_s=0.3, offset_t=0.15,
                  orbit_start=0.0, orbit=2 * math.pi / 24, couch_begin=-5.0,
                  couch_mov=HELI_COLL / 24, view_per_rot=24, nz=4, dz=0.9,
                  z0=-5.0 + 30 * HELI_COLL / 24, thickness=1.5, dso=595.0, dsd=1085.6)
HELI_PITCH = dict(ns=23, nt=5, na=80, ds=15.0, dt=1.0, offset_s=-0.7, offset_t=-0.4,
                  orbit_start=0.0, orbit=2 * math.pi / 20, couch_begin=12.0,
                  couch_mov=1.7 * 5 * 595.0 / 1085.6 / 20, view_per_rot=20, nz=5, dz=-0.8,
                  z0=12.0 + 52 * 1.7 * 5 * 595.0 / 1085.6 / 20, thickness=3.0, dso=595.0,
                  dsd=1085.6)
HELI_GAP = dict(HELI_SMALL, couch_mov=3 * HELI_COLL / 24, z0=-5.0 + 30 * 3 * HELI_COLL / 24,
                dz=2.0, thickness=0.2)
HELI_SHORT = dict(HELI_SMALL, z0=-5.0 + 20 * HELI_COLL / 24)


def heli_fixtures():
    for name, geo, seed in (("heli_small", HELI_SMALL, 101), ("heli_pitch", HELI_PITCH, 103),
                            ("heli_gap", HELI_GAP, 105), ("heli_short", HELI_SHORT, 107)):
        h = O.heli(**geo)
        proj = randn(seed, (h.na, h.nt, h.ns))
        out = {}
        for mode in (360, 180):
            for suf, x in (("f32", proj), ("f64", proj.astype(np.float64))):
                y, ok = O.helical(h, x, mode, kind="ref")
                out[f"ok_{mode}_{suf}"] = np.int64(ok)
                if ok:
                    out[f"out_{mode}_{suf}"] = y
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, geom=json.dumps(geo), proj_seed=np.int64(seed), **out)
        print(f"wrote {path} ({os.path.getsize(path)} B)",
              {k: int(v) for k, v in out.items() if k.startswith("ok")})


if __name__ == "__main__":
    if not O.have_ref():
        sys.exit("oracle/_ref/libctref.so missing: run `make -C oracle ref` first")
    fan_fixtures()
    par_fixtures()
    svbp_fixtures()
    rc_fixtures()
    fp_fixtures()
    fpdd_fixtures()
    filter_fixtures()
    heli_fixtures()
