"""ctypes binding of lib/libctrecon_hip.so (the C ABI of include/ctr_api.h).

There is no fallback: if the library is missing or cannot be loaded, every
operator raises.  Build it with ``python -m ct_recon_extensions_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libctrecon_hip.so")

CTR_F32, CTR_F64 = 0, 1


class CtrGeom(C.Structure):
    """Mirror of ``ctr_geom`` (include/ctr_api.h)."""

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


class CtrRaggedProblem(C.Structure):
    """Mirror of ``ctr_ragged_problem`` (include/ctr_api.h)."""

    _fields_ = [
        ("g", CtrGeom),
        ("xpos", C.c_void_p), ("ypos", C.c_void_p), ("sincostbl", C.c_void_p),
        ("proj", C.c_void_p), ("img", C.c_void_p),
        ("scale", C.c_double),
    ]


class CtrFanBpLaunch(C.Structure):
    """Mirror of ``ctr_fan_bp_launch`` (include/ctr_api.h)."""

    _fields_ = [("tile", C.c_int), ("vchunk", C.c_int), ("nb", C.c_int)]


class CtrHelicalGeom(C.Structure):
    """Mirror of ``ctr_helical_geom`` (include/ctr_api.h)."""

    _fields_ = [
        ("ns", C.c_uint), ("nt", C.c_uint), ("na", C.c_uint),
        ("ds", C.c_double), ("dt", C.c_double), ("offset_s", C.c_double),
        ("offset_t", C.c_double), ("orbit_start", C.c_double), ("orbit", C.c_double),
        ("couch_begin", C.c_double), ("couch_mov", C.c_double),
        ("view_per_rot", C.c_uint), ("nz", C.c_uint),
        ("dz", C.c_double), ("z0", C.c_double), ("thickness", C.c_double),
        ("dso", C.c_double), ("dsd", C.c_double),
    ]


class CtrError(RuntimeError):
    pass


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_G = C.POINTER(CtrGeom)
_H = C.POINTER(CtrHelicalGeom)

_SIGS = {
    "ctr_last_error": (C.c_char_p, []),
    "ctr_version": (C.c_int, []),
    "ctr_device_info": (C.c_int, [_I, _I, _I, C.c_char_p, C.c_int]),
    "ctr_fan_tables": (C.c_int, [_G, _P, _P, _P]),
    "ctr_fan_angle_tables": (C.c_int, [_G, _P, _P, _P, _P]),
    "ctr_par_tables": (C.c_int, [_G, _P, _P]),
    "ctr_fp_tables": (C.c_int, [_G, _P, _P, _P]),
    "ctr_fp_rc_tables": (C.c_int, [_G, C.c_double, _P, _P, _P]),
    "ctr_fp_grad_tables": (C.c_int, [_G, _P, _P, _P, C.c_uint]),
    "ctr_fp_dd_tables": (C.c_int, [_G, _P, _P, _P]),
    "ctr_ramp_taps": (C.c_int, [_G, _P]),
    "ctr_fan_bp": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, _P, C.c_double, _P]),
    "ctr_fan_bp_ex": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, _P, C.c_double,
                                C.POINTER(CtrFanBpLaunch), _P]),
    "ctr_fan_bp_t": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, _P, C.c_double, C.c_int, _P]),
    "ctr_plain_fan_bp": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, _P, C.c_double, _P]),
    "ctr_plain_fan_bp_grid": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, _P, C.c_double,
                                        C.c_int, C.c_int, _P]),
    "ctr_par_bp": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, _P]),
    "ctr_par_bp_t": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, C.c_int, _P]),
    "ctr_par_sv_bp": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, _P]),
    "ctr_fp": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, _P, _P]),
    "ctr_fp_rc": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, C.c_double, _P, _P, _P, _P]),
    "ctr_fp_t": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, _P, C.c_uint, _P]),
    "ctr_fp_t_noalloc": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, _P, C.c_uint, _P]),
    "ctr_fp_dd": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, _P, _P]),
    "ctr_fp_dd_grad": (C.c_int, [_G, C.c_int, _P, _P, C.c_int, _P, _P, _P, C.c_uint, _P]),
    "ctr_ramp": (C.c_int, [_G, C.c_int, _P, _P, C.c_uint, _P, C.c_int, _P]),
    "ctr_fan_weighting": (C.c_int, [_G, C.c_int, _P, _P, C.c_uint, _P]),
    "ctr_ramp_toeplitz": (C.c_int, [_G, C.c_int, _P, _P, C.c_uint, _P, C.c_int, C.c_int, _P]),
    "ctr_helical_tables": (C.c_int, [_H, C.c_int, _P]),
    "ctr_helical": (C.c_int, [_H, C.c_int, C.c_int, _P, _P, C.c_int, _P, _P, _P]),
    "ctr_ragged_fan_bp_plan": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(_P)]),
    "ctr_ragged_fan_bp_run": (C.c_int, [_P, C.c_int, _P]),
    "ctr_ragged_fan_bp_info": (C.c_int, [_P, _I, _I, _I, _I, _D]),
    "ctr_ragged_fan_bp_balance": (C.c_int, [_P, _P, _D]),
    "ctr_ragged_fan_bp_destroy": (C.c_int, [_P]),
    "ctr_ragged_fan_bp_autotune": (C.c_int, [_P, C.c_int, C.c_int, _P, _I, _I,
                                             C.POINTER(C.c_float)]),
}

# Entry points that only exist once the corresponding sources are built.
_OPTIONAL: dict = {}

_lock = threading.Lock()
_lib = None


def load():
    """Load the HIP library (once).  Raises if it is absent."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise CtrError(
                    f"{LIB_PATH} is missing: build it with "
                    "`python -m ct_recon_extensions_amd.build` (there is no CPU fallback)")
            lib = C.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                f = getattr(lib, name)
                f.restype = res
                f.argtypes = args
            for name, (res, args) in _OPTIONAL.items():
                if hasattr(lib, name):
                    f = getattr(lib, name)
                    f.restype = res
                    f.argtypes = args
            _lib = lib
    return _lib


def has(name: str) -> bool:
    return hasattr(load(), name)


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().ctr_last_error().decode(errors="replace")
        raise CtrError(f"{what or 'ctr call'} failed (status {rc}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def last_error() -> str:
    return load().ctr_last_error().decode(errors="replace")
