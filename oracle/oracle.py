"""ctypes front-end of the CPU parity oracle (TEST INFRASTRUCTURE ONLY).

Two libraries share one calling convention:

* ``oracle/libct_oracle.so``: the C restatement in ``oracle/ct_oracle.c``
  (symbols ``orc_*``);
* ``oracle/_ref/libctref.so``: the reference's own ``src/*.cc`` compiled
  unchanged with ``oracle/ref_shim.cpp`` (symbols ``ref_*``), present only
  where ``make -C oracle ref`` ran against the reference checkout.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module, and only to check or to time the
CPU path.  The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libct_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libctref.so")


class Geom(C.Structure):
    """Mirror of ``orc_geom`` (one record for every operator)."""

    _fields_ = [
        ("ns", C.c_uint), ("na", C.c_uint),
        ("ds", C.c_double), ("orbit", C.c_double), ("offset_s", C.c_double),
        ("orbit_start", C.c_double),
        ("nx", C.c_uint), ("ny", C.c_uint),
        ("dx", C.c_double), ("dy", C.c_double), ("offset_x", C.c_double),
        ("offset_y", C.c_double),
        ("dso", C.c_double), ("dsd", C.c_double), ("fov", C.c_double),
        ("type", C.c_int), ("nrow", C.c_uint),
    ]


def geom(**kw) -> Geom:
    g = Geom()
    defaults = dict(ns=1, na=1, ds=1.0, orbit=0.0, offset_s=0.0, orbit_start=0.0,
                    nx=1, ny=1, dx=1.0, dy=1.0, offset_x=0.0, offset_y=0.0,
                    dso=-1.0, dsd=-1.0, fov=-1.0, type=0, nrow=1)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(g, k, v)
    return g


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


class _Lib:
    def __init__(self, path: str, prefix: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle / make -C oracle ref)")
        self.lib = C.CDLL(path)
        self.prefix = prefix
        self.path = path

    def fn(self, name: str):
        f = getattr(self.lib, self.prefix + name)
        f.restype = C.c_int
        return f

    def call(self, name: str, g: Geom, *arrays_and_ints):
        args = [C.byref(g)]
        for a in arrays_and_ints:
            if isinstance(a, np.ndarray):
                args.append(_ptr(a))
            elif a is None:
                args.append(None)
            elif isinstance(a, C.c_double):
                args.append(a)
            else:
                args.append(C.c_int(int(a)))
        ok = self.fn(name)(*args)
        if not ok:
            raise RuntimeError(f"{self.prefix}{name} returned false")


_cache: dict = {}


def lib(kind: str = "oracle") -> _Lib:
    """``kind``: "oracle" (the C restatement) or "ref" (compiled reference)."""
    if kind not in _cache:
        _cache[kind] = _Lib(ORACLE_SO, "orc_") if kind == "oracle" else _Lib(REF_SO, "ref_")
    return _cache[kind]


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def _suffix(dtype) -> str:
    return "f" if np.dtype(dtype) == np.float32 else "d"


# ---------------------------------------------------------------- fan beam
def fan_prep(g: Geom, kind="oracle"):
    xpos = np.zeros(g.nx); ypos = np.zeros(g.ny); sc = np.zeros(2 * g.na)
    lib(kind).call("fan_prep", g, xpos, ypos, sc)
    return xpos, ypos, sc


def fan_angle_prep(g: Geom, angles: np.ndarray, kind="oracle"):
    angles = np.ascontiguousarray(angles, np.float64)
    xpos = np.zeros(g.nx); ypos = np.zeros(g.ny); sc = np.zeros(2 * g.na)
    lib(kind).call("fan_angle_prep", g, angles, xpos, ypos, sc)
    return xpos, ypos, sc


def fan_bp(g: Geom, proj: np.ndarray, tables, scaled=True, kind="oracle"):
    """``scaled``: bp_fan_2d (True) or bp_fan_2d_angle (False)."""
    proj = np.ascontiguousarray(proj)
    img = np.zeros((g.ny, g.nx), proj.dtype)
    xpos, ypos, sc = tables
    s = _suffix(proj.dtype)
    if kind == "oracle":
        lib(kind).call("fan_bp_" + s, g, proj, img, xpos, ypos, sc, int(scaled))
    else:
        lib(kind).call(("fan_bp_" if scaled else "fan_angle_bp_") + s, g, proj, img,
                       xpos, ypos, sc)
    return img


def fan_bpt(g: Geom, img: np.ndarray, tables, scaled=True, kind="oracle"):
    img = np.ascontiguousarray(img)
    grad = np.zeros((g.na, g.ns), img.dtype)
    xpos, ypos, sc = tables
    s = _suffix(img.dtype)
    if kind == "oracle":
        lib(kind).call("fan_bpt_" + s, g, img, grad, xpos, ypos, sc, int(scaled))
    else:
        if not scaled:
            raise ValueError("the reference's bp_fan_2d_angle transpose is not a transpose")
        lib(kind).call("fan_bpt_" + s, g, img, grad, xpos, ypos, sc)
    return grad


# ----------------------------------------------------------- parallel beam
def par_prep(g: Geom, kind="oracle"):
    xcos = np.zeros(g.na * g.nx); ysin = np.zeros(g.na * g.ny)
    lib(kind).call("par_prep", g, xcos, ysin)
    return xcos, ysin


def par_bp(g: Geom, proj: np.ndarray, tables, kind="oracle"):
    proj = np.ascontiguousarray(proj)
    img = np.zeros((g.ny, g.nx), proj.dtype)
    xcos, ysin = tables
    lib(kind).call("par_bp_" + _suffix(proj.dtype), g, proj, img, xcos, ysin)
    return img


def par_svbp_prep(g: Geom, kind="oracle"):
    """Tables of the single-view parallel BP (src/bp_par_2d_sv.cc:18-49, the
    same as the parallel BP's; the restatement reuses orc_par_prep)."""
    xcos = np.zeros(g.na * g.nx); ysin = np.zeros(g.na * g.ny)
    lib(kind).call("par_svbp_prep" if kind == "ref" else "par_prep", g, xcos, ysin)
    return xcos, ysin


def par_svbp(g: Geom, proj: np.ndarray, tables, kind="oracle"):
    """Single-view parallel BP (src/bp_par_2d_sv.cc:54-92): img[iy, ix, ia]."""
    proj = np.ascontiguousarray(proj)
    img = np.zeros((g.ny, g.nx, g.na), proj.dtype)
    xcos, ysin = tables
    lib(kind).call("par_svbp_" + _suffix(proj.dtype), g, proj, img, xcos, ysin)
    return img


def par_bpt(g: Geom, img: np.ndarray, tables):
    img = np.ascontiguousarray(img)
    grad = np.zeros((g.na, g.ns), img.dtype)
    work = np.zeros(g.na * g.ns)
    xcos, ysin = tables
    lib("oracle").call("par_bpt_" + _suffix(img.dtype), g, img, grad, xcos, ysin, work)
    return grad


# ------------------------------------------------------- ray-driven FP
def fp_prep(g: Geom, kind="oracle"):
    sc = np.zeros(2 * g.na); bo = np.zeros(2 * g.na * g.ns); usex = np.zeros(g.na, np.int32)
    lib(kind).call("fp_prep", g, sc, bo, usex)
    return sc, bo, usex


def fp(g: Geom, img: np.ndarray, tables, kind="oracle"):
    img = np.ascontiguousarray(img)
    proj = np.zeros((g.na, g.ns), img.dtype)
    sc, bo, usex = tables
    lib(kind).call("fp_" + _suffix(img.dtype), g, img, proj, sc, bo, usex)
    return proj


def fp_rc_prep(g: Geom, step=0.125, kind="oracle"):
    """Ray-casting FP tables (src/fp_par_2d.cc:18-89): sincos [na][2],
    entry points [na][ns][2], step counts [na][ns]."""
    sc = np.zeros(2 * g.na); begins = np.zeros(2 * g.na * g.ns)
    nsteps = np.zeros(g.na * g.ns, np.int32)
    lib(kind).call("fp_rc_prep", g, C.c_double(step), sc, begins, nsteps)
    return sc, begins, nsteps


def fp_rc(g: Geom, img: np.ndarray, tables, step=0.125, kind="oracle"):
    """Ray-casting parallel FP (src/fp_par_2d.cc:91-144)."""
    img = np.ascontiguousarray(img)
    proj = np.zeros((g.na, g.ns), img.dtype)
    sc, begins, nsteps = tables
    lib(kind).call("fp_rc_" + _suffix(img.dtype), g, C.c_double(step), img, proj, sc, begins,
                   nsteps)
    return proj


def fp_grad_prep(g: Geom, kind="oracle"):
    w = np.zeros(g.na); pos = np.zeros(2 * g.na * max(g.nx, g.ny)); usex = np.zeros(g.na, np.int32)
    lib(kind).call("fp_grad_prep", g, w, pos, usex)
    return w, pos, usex


def fp_grad(g: Geom, proj: np.ndarray, tables, kind="oracle"):
    proj = np.ascontiguousarray(proj)
    img = np.zeros((g.ny, g.nx), proj.dtype)
    w, pos, usex = tables
    lib(kind).call("fp_grad_" + _suffix(proj.dtype), g, proj, img, w, pos, usex)
    return img


# -------------------------------------------------- distance-driven FP
def fp_dd_prep(g: Geom, kind="oracle"):
    """Distance-driven FP tables (src/fp_par_2d.cc:396-457): sincos [na][2],
    (begin, offset, half-width) [na][ns][3], branch [na]."""
    sc = np.zeros(2 * g.na); bo = np.zeros(3 * g.na * g.ns); usex = np.zeros(g.na, np.int32)
    lib(kind).call("fp_dd_prep", g, sc, bo, usex)
    return sc, bo, usex


def fp_dd(g: Geom, img: np.ndarray, tables, kind="oracle"):
    """Distance-driven parallel FP (src/fp_par_2d.cc:459-543)."""
    img = np.ascontiguousarray(img)
    proj = np.zeros((g.na, g.ns), img.dtype)
    sc, bo, usex = tables
    lib(kind).call("fp_dd_" + _suffix(img.dtype), g, img, proj, sc, bo, usex)
    return proj


def fp_dd_grad_prep(g: Geom, kind="oracle"):
    """Its gradient's tables (src/fp_par_2d.cc:548-599); the restatement's are
    the ray-driven gradient's, the same arithmetic."""
    w = np.zeros(g.na); pos = np.zeros(2 * g.na * max(g.nx, g.ny))
    usex = np.zeros(g.na, np.int32)
    lib(kind).call("fp_dd_grad_prep" if kind == "ref" else "fp_grad_prep", g, w, pos, usex)
    return w, pos, usex


def fp_dd_grad(g: Geom, proj: np.ndarray, tables, kind="oracle"):
    """The distance-driven FP's gradient as the reference defines it
    (src/fp_par_2d.cc:601-655)."""
    proj = np.ascontiguousarray(proj)
    img = np.zeros((g.ny, g.nx), proj.dtype)
    w, pos, usex = tables
    lib(kind).call("fp_dd_grad_" + _suffix(proj.dtype), g, proj, img, w, pos, usex)
    return img


# ------------------------------------------------- ramp filter / weighting
def ramp_prep(g: Geom, kind="oracle"):
    h = np.zeros(2 * g.ns + 1)
    lib(kind).call("ramp_prep", g, h)
    return h


def ramp(g: Geom, x: np.ndarray, h: np.ndarray, grad=False, kind="oracle"):
    x = np.ascontiguousarray(x)
    out = np.zeros_like(x)
    lib(kind).call(("ramp_grad_" if grad else "ramp_") + _suffix(x.dtype), g, x, h, out)
    return out


def fanw(g: Geom, x: np.ndarray, grad=False, kind="oracle"):
    x = np.ascontiguousarray(x)
    out = np.zeros_like(x)
    lib(kind).call("fanw_" + _suffix(x.dtype), g, x, out, int(grad))
    return out


# ------------------------------------------------------- helical rebinning
class Heli(C.Structure):
    """Mirror of ``orc_heli``: the reference's HelicalInterpolationParam
    (include/helical_interpolation.h:17-63) in its field order."""

    _fields_ = [
        ("ns", C.c_uint), ("nt", C.c_uint), ("na", C.c_uint),
        ("ds", C.c_double), ("dt", C.c_double), ("offset_s", C.c_double),
        ("offset_t", C.c_double), ("orbit_start", C.c_double), ("orbit", C.c_double),
        ("couch_begin", C.c_double), ("couch_mov", C.c_double),
        ("view_per_rot", C.c_uint), ("nz", C.c_uint),
        ("dz", C.c_double), ("z0", C.c_double), ("thickness", C.c_double),
        ("dso", C.c_double), ("dsd", C.c_double),
    ]


def heli(**kw) -> Heli:
    h = Heli()
    for k, _ in Heli._fields_:
        setattr(h, k, kw[k])
    return h


def helical(h: Heli, proj: np.ndarray, mode: int, kind="oracle"):
    """360LI (mode 360, src/helical_interpolation.cc:81-160 via :336-348) or
    180LI (mode 180, :162-269 via :353-365) of proj[na][nt][ns] into
    out[nz][view_per_rot][ns].  Returns (out, code): 1, or 0 for a slice
    short of views and -1 for a sample with no detector row in either window
    (the reference throws for both), -2 (oracle only) for a 180LI conjugate
    view interpolated from outside the scan (the reference reads outside its
    input).  The compiled reference returns 0 for either throw."""
    proj = np.ascontiguousarray(proj)
    assert proj.shape == (h.na, h.nt, h.ns)
    out = np.zeros((h.nz, h.view_per_rot, h.ns), proj.dtype)
    f = lib(kind).lib[f"{lib(kind).prefix}heli{mode}_{_suffix(proj.dtype)}"]
    f.restype = C.c_int
    return out, int(f(C.byref(h), _ptr(proj), _ptr(out)))
