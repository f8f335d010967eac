"""Geometry records and their FP64 tables.

The dataclasses mirror the reference's Param structs field for field and in
constructor order (include/bp_fan_2d.h:20-59, include/bp_fan_2d_angle.h:20-55,
include/bp_par_2d.h:20-55, include/fp_par_2d.h:20-55, include/filter.h:21-36,
include/fan_weighting.h:21-39, include/helical_interpolation.h:17-63).
Tables are built on the host by the library's C ABI (the reference's
accumulation order, FP64) and cached on the device per geometry.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import math
from collections import OrderedDict
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass(frozen=True)
class FanBackprojection2DParam:
    ns: int
    na: int
    ds: float
    orbit: float
    offset_s: float
    orbit_start: float
    nx: int
    ny: int
    dx: float
    dy: float
    offset_x: float
    offset_y: float
    dso: float
    dsd: float
    fov: float = -1.0

    def ctr(self) -> _lib.CtrGeom:
        return _geom(self)

    @property
    def scale(self) -> float:
        """|orbit| / round(|na*orbit|/pi) (src/bp_fan_2d.cc:58,83)."""
        return abs(self.orbit) / round(abs(self.na * self.orbit) / math.pi)


@dataclass(frozen=True)
class FanBackprojection2DAngleParam:
    ns: int
    na: int
    ds: float
    offset_s: float
    nx: int
    ny: int
    dx: float
    dy: float
    offset_x: float
    offset_y: float
    dso: float
    dsd: float
    fov: float = -1.0

    def ctr(self) -> _lib.CtrGeom:
        return _geom(self)


@dataclass(frozen=True)
class ParallelBackprojection2DParam:
    ns: int
    na: int
    ds: float
    orbit: float
    offset_s: float
    orbit_start: float
    nx: int
    ny: int
    dx: float
    dy: float
    offset_x: float
    offset_y: float
    fov: float = -1.0

    def ctr(self) -> _lib.CtrGeom:
        return _geom(self)


@dataclass(frozen=True)
class ParallelProjection2DParam(ParallelBackprojection2DParam):
    pass


@dataclass(frozen=True)
class FilterParam:
    ns: int
    nrow: int
    ds: float
    dsd: float = -1.0
    type: int = 0  # 0 par, 1 fan (equiangular), 2 flat
    window: int = 0  # not implemented by the reference either (include/filter.h:28)

    def ctr(self) -> _lib.CtrGeom:
        return _geom(self)


@dataclass(frozen=True)
class FanWeightingParam:
    ns: int
    nrow: int
    ds: float
    offset: float = 0.0
    dso: float = -1.0
    dsd: float = -1.0
    type: int = 1  # 1 fan, 2 flat

    def ctr(self) -> _lib.CtrGeom:
        g = _geom(self)
        g.offset_s = self.offset
        return g


@dataclass(frozen=True)
class HelicalInterpolationParam:
    """include/helical_interpolation.h:17-63, in its constructor order."""
    ns: int
    nt: int
    na: int
    ds: float
    dt: float
    offset_s: float
    offset_t: float
    orbit_start: float
    orbit: float
    couch_begin: float
    couch_mov: float
    view_per_rot: int
    nz: int
    dz: float
    z0: float
    thickness: float
    dso: float
    dsd: float

    def ctr(self) -> _lib.CtrHelicalGeom:
        g = _lib.CtrHelicalGeom()
        for name, _ in _lib.CtrHelicalGeom._fields_:
            setattr(g, name, getattr(self, name))
        return g


def _geom(p) -> _lib.CtrGeom:
    g = _lib.CtrGeom()
    for name, _ in _lib.CtrGeom._fields_:
        if hasattr(p, name):
            setattr(g, name, getattr(p, name))
    if not hasattr(p, "fov"):
        g.fov = -1.0
    return g


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------ host tables
def fan_tables(p: FanBackprojection2DParam):
    xpos, ypos, sc = np.zeros(p.nx), np.zeros(p.ny), np.zeros(2 * p.na)
    _lib.call("ctr_fan_tables", C.byref(p.ctr()), _ptr(xpos), _ptr(ypos), _ptr(sc))
    return xpos, ypos, sc


def fan_angle_tables(p: FanBackprojection2DAngleParam, angles):
    angles = np.ascontiguousarray(angles, np.float64)
    if angles.shape != (p.na,):
        raise ValueError(f"angles must have shape ({p.na},), got {angles.shape}")
    xpos, ypos, sc = np.zeros(p.nx), np.zeros(p.ny), np.zeros(2 * p.na)
    _lib.call("ctr_fan_angle_tables", C.byref(p.ctr()), _ptr(angles), _ptr(xpos), _ptr(ypos),
              _ptr(sc))
    return xpos, ypos, sc


def par_tables(p: ParallelBackprojection2DParam):
    xcos, ysin = np.zeros(p.na * p.nx), np.zeros(p.na * p.ny)
    _lib.call("ctr_par_tables", C.byref(p.ctr()), _ptr(xcos), _ptr(ysin))
    return xcos, ysin


def fp_tables(p: ParallelProjection2DParam):
    sc, bo = np.zeros(2 * p.na), np.zeros(2 * p.na * p.ns)
    usex = np.zeros(p.na, np.int32)
    _lib.call("ctr_fp_tables", C.byref(p.ctr()), _ptr(sc), _ptr(bo), _ptr(usex))
    return sc, bo, usex


RC_STEP = 0.125  # the reference's ray-casting step (include/fp_par_2d.h:160,186)


def fp_dd_tables(p: ParallelProjection2DParam):
    sc, bo = np.zeros(2 * p.na), np.zeros(3 * p.na * p.ns)
    usex = np.zeros(p.na, np.int32)
    _lib.call("ctr_fp_dd_tables", C.byref(p.ctr()), _ptr(sc), _ptr(bo), _ptr(usex))
    return sc, bo, usex


def fp_rc_tables(p: ParallelProjection2DParam, step: float = RC_STEP):
    sc, begins = np.zeros(2 * p.na), np.zeros(2 * p.na * p.ns)
    nsteps = np.zeros(p.na * p.ns, np.int32)
    _lib.call("ctr_fp_rc_tables", C.byref(p.ctr()), float(step), _ptr(sc), _ptr(begins),
              _ptr(nsteps))
    return sc, begins, nsteps


def fp_grad_tables(p: ParallelProjection2DParam, stride: int | None = None):
    stride = max(p.nx, p.ny) if stride is None else stride
    w, pos = np.zeros(p.na), np.zeros(2 * p.na * stride)
    usex = np.zeros(p.na, np.int32)
    _lib.call("ctr_fp_grad_tables", C.byref(p.ctr()), _ptr(w), _ptr(pos), _ptr(usex), stride)
    return w, pos, usex


def helical_tables(p: HelicalInterpolationParam, mode: int):
    """Per channel cos(gamma), (pi - 2 gamma)/orbit and cs - s/ds, then per
    slice (z, centre view, window begin, first output view); raises where the
    reference reports "View not enough for this slice!"."""
    t = np.zeros(3 * p.ns + 4 * p.nz)
    _lib.call("ctr_helical_tables", C.byref(p.ctr()), int(mode), _ptr(t))
    return t


def ramp_taps(p: FilterParam):
    h = np.zeros(2 * p.ns + 1)
    _lib.call("ctr_ramp_taps", C.byref(p.ctr()), _ptr(h))
    return h


# ---------------------------------------------------------- device cache
_CACHE: "OrderedDict" = OrderedDict()
_CACHE_MAX = 64


def device_tables(kind: str, p, device, angles=None):
    """FP64 tables of ``kind`` for geometry ``p`` as torch tensors on ``device``."""
    import torch

    extra = b""
    if angles is not None:
        extra = hashlib.sha1(np.ascontiguousarray(angles, np.float64).tobytes()).digest()
    # the parameter structs are frozen dataclasses: hashable by their fields
    key = (kind, type(p).__name__, p, str(device), extra)
    hit = _CACHE.get(key)
    if hit is not None:
        _CACHE.move_to_end(key)
        # Kernels about to read these tables may run on a stream other than
        # the one that allocated them: tell the caching allocator, so that
        # once the entry is evicted its memory is not handed out again
        # before that stream's work is done.
        st = torch.cuda.current_stream(hit[0].device)
        for t in hit:
            t.record_stream(st)
        return hit
    if kind == "fan":
        host = fan_tables(p)
    elif kind == "fan_angle":
        host = fan_angle_tables(p, angles)
    elif kind == "par":
        host = par_tables(p)
    elif kind == "fp":
        host = fp_tables(p)
    elif kind == "fp_rc":
        host = fp_rc_tables(p)
    elif kind == "fp_dd":
        host = fp_dd_tables(p)
    elif kind == "fp_grad":
        host = fp_grad_tables(p)
    elif kind == "ramp":
        host = (ramp_taps(p),)
    elif kind in ("helical_360", "helical_180"):
        host = (helical_tables(p, int(kind[-3:])),)
    else:
        raise ValueError(kind)
    dev = tuple(torch.from_numpy(a).to(device) for a in host)
    _CACHE[key] = dev
    while len(_CACHE) > _CACHE_MAX:
        _CACHE.popitem(last=False)
    return dev
