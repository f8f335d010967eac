"""numpy front-end on the reference's numpy C ABI (include/numpy_ctext.h).

Mirrors numpy_ext/np_ctrecon.py:9-99 of the reference: the same ctypes
bindings of lib/libnumpy_ctext.so and the same ``FanFBP2DAngle`` class
(fan weighting -> equiangular ramp filter -> bp_fan_2d_angle, times orbit),
with two corrections: return types are declared (``restype``; the reference
sets a misspelt ``rettype`` and so reads every bool as an int), and the
caller's sinogram is not overwritten (the reference weights it in place).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import numpy.ctypeslib as npct

# argument kinds of the eleven entry points: contiguous float64 vectors and
# matrices, and the C scalars of include/numpy_ctext.h
_VEC = npct.ndpointer(np.float64, ndim=1, flags="C_CONTIGUOUS")
_MAT = npct.ndpointer(np.float64, ndim=2, flags="C_CONTIGUOUS")
_U, _I, _D, _B = C.c_uint, C.c_int, C.c_double, C.c_bool

_LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
_lib = None


def _load():
    global _lib
    if _lib is None:
        lib = npct.load_library("libnumpy_ctext.so", _LIBDIR)
        sigs = {
            "set_device": (_B, [_I]),
            "clear": (_B, []),
            "fan_weighting_create": (_I, [_U, _U, _D, _D, _D, _D, _I]),
            "fan_weighting_run": (_B, [_I, _MAT, _MAT]),
            "fan_weighting_destroy": (_B, [_I]),
            "ramp_filter_create": (_I, [_U, _U, _D, _D, _I]),
            "ramp_filter_run": (_B, [_I, _MAT, _MAT]),
            "ramp_filter_destroy": (_B, [_I]),
            "fan_bp_2d_angle_create": (_I, [_VEC, _U, _U, _D, _D, _U, _U] + [_D] * 7),
            "fan_bp_2d_angle_run": (_B, [_I, _MAT, _MAT]),
            "fan_bp_2d_angle_destroy": (_B, [_I]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class FanFBP2DAngle:
    """Fan-beam FBP at arbitrary source angles (numpy_ext/np_ctrecon.py:66-99).

    ``param`` keys: device, ns, na, ds, offset_s, nx, ny, dx, dy, offset_x,
    offset_y, dso, dsd, fov, orbit, angles.
    """

    def __init__(self, param: dict):
        lib = _load()
        self._param = param
        itype = 1
        if not lib.set_device(param["device"]):
            raise RuntimeError("set_device failed (this library computes on a GPU; device < 0 "
                               "selects GPU 0)")
        self._angles = np.ascontiguousarray(param["angles"], dtype=np.double)
        self._fw_handle = lib.fan_weighting_create(param["ns"], param["na"], param["ds"],
                                                   param["offset_s"], param["dso"],
                                                   param["dsd"], itype)
        self._flt_handle = lib.ramp_filter_create(param["ns"], param["na"], param["ds"],
                                                  param["dsd"], itype)
        self._bp_handle = lib.fan_bp_2d_angle_create(
            self._angles, param["ns"], param["na"], param["ds"], param["offset_s"], param["nx"],
            param["ny"], param["dx"], param["dy"], param["offset_x"], param["offset_y"],
            param["dso"], param["dsd"], param["fov"])
        if not (self._fw_handle and self._flt_handle and self._bp_handle):
            raise RuntimeError("operator creation failed (check the geometry)")

    def __del__(self):
        lib = _lib
        if lib is not None:
            for h, f in ((getattr(self, "_fw_handle", 0), lib.fan_weighting_destroy),
                         (getattr(self, "_flt_handle", 0), lib.ramp_filter_destroy),
                         (getattr(self, "_bp_handle", 0), lib.fan_bp_2d_angle_destroy)):
                if h:
                    f(h)

    def run(self, proj):
        lib = _lib
        proj = np.array(proj, dtype=np.double, order="C", copy=True)
        na, ns = proj.shape
        if (na, ns) != (self._param["na"], self._param["ns"]):
            raise RuntimeError(f"sinogram is {na} x {ns}; this operator takes "
                               f"{self._param['na']} x {self._param['ns']} (views x channels)")
        weighted = np.empty_like(proj)
        filtered = np.empty_like(proj)
        recon = np.zeros([self._param["ny"], self._param["nx"]], dtype=np.double)
        for stage, ok in (("fan weighting", lambda: lib.fan_weighting_run(self._fw_handle, proj,
                                                                          weighted)),
                          ("ramp filter", lambda: lib.ramp_filter_run(self._flt_handle, weighted,
                                                                      filtered)),
                          ("backprojection", lambda: lib.fan_bp_2d_angle_run(
                              self._bp_handle, filtered, recon))):
            if not ok():
                raise RuntimeError(f"FanFBP2DAngle: the {stage} stage failed on the GPU")
        return recon * self._param["orbit"]
