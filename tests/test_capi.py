"""The C ABI boundary on the CPU (no GPU calls): the library loads, exports
every entry point its headers declare, builds the reference's FP64 tables
bit-for-bit, and reports errors as status codes."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import ROOT

LIB = os.path.join(ROOT, "ct_recon_extensions_amd", "lib", "libctrecon_hip.so")
NPLIB = os.path.join(ROOT, "ct_recon_extensions_amd", "lib", "libnumpy_ctext.so")


def declared(header, prefix_re):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(prefix_re, text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} not built: run python -m ct_recon_extensions_amd.build")
    from ct_recon_extensions_amd import _lib
    return _lib.load()


def test_exports_every_declared_ctr_symbol(lib):
    names = declared("ctr_api.h", r"\b(ctr_[a-z0-9_]+)\s*\(")
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(NPLIB), reason="numpy C ABI not built")
def test_numpy_library_exports_the_reference_symbols():
    nl = C.CDLL(NPLIB)
    names = declared("numpy_ctext.h", r"\b([a-z_0-9]+)\s*\([^;]*\)\s*;")
    expected = {"set_device", "clear", "fan_weighting_create", "fan_weighting_run",
                "fan_weighting_destroy", "ramp_filter_create", "ramp_filter_run",
                "ramp_filter_destroy", "fan_bp_2d_angle_create", "fan_bp_2d_angle_run",
                "fan_bp_2d_angle_destroy"}
    assert expected <= set(names)
    for n in expected:
        assert hasattr(nl, n), n


def _geom(**kw):
    from ct_recon_extensions_amd import _lib
    g = _lib.CtrGeom()
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def test_tables_match_oracle_bitwise(lib):
    from ct_recon_extensions_amd import geometry as G
    from tests.golden.make_golden import FAN_SMALL, FP_SMALL, PAR_SMALL
    f = FAN_SMALL
    p = G.FanBackprojection2DParam(f["ns"], f["na"], f["ds"], f["orbit"], f["offset_s"],
                                   f["orbit_start"], f["nx"], f["ny"], f["dx"], f["dy"],
                                   f["offset_x"], f["offset_y"], f["dso"], f["dsd"])
    for a, b in zip(G.fan_tables(p), O.fan_prep(O.geom(**f))):
        np.testing.assert_array_equal(a, b)
    angles = np.sort(np.random.default_rng(3).uniform(0, 6, f["na"]))
    pa = G.FanBackprojection2DAngleParam(f["ns"], f["na"], f["ds"], f["offset_s"], f["nx"],
                                         f["ny"], f["dx"], f["dy"], f["offset_x"],
                                         f["offset_y"], f["dso"], f["dsd"])
    for a, b in zip(G.fan_angle_tables(pa, angles), O.fan_angle_prep(O.geom(**f), angles)):
        np.testing.assert_array_equal(a, b)
    q = PAR_SMALL
    pp = G.ParallelBackprojection2DParam(q["ns"], q["na"], q["ds"], q["orbit"], q["offset_s"],
                                         q["orbit_start"], q["nx"], q["ny"], q["dx"], q["dy"],
                                         q["offset_x"], q["offset_y"])
    for a, b in zip(G.par_tables(pp), O.par_prep(O.geom(**q))):
        np.testing.assert_array_equal(a, b)
    r = FP_SMALL
    pf = G.ParallelProjection2DParam(r["ns"], r["na"], r["ds"], r["orbit"], r["offset_s"],
                                     r["orbit_start"], r["nx"], r["ny"], r["dx"], r["dy"],
                                     r["offset_x"], r["offset_y"])
    for a, b in zip(G.fp_tables(pf), O.fp_prep(O.geom(**r))):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(G.fp_grad_tables(pf, stride=r["nx"]), O.fp_grad_prep(O.geom(**r))):
        np.testing.assert_array_equal(a, b)
    for typ in (0, 1, 2):
        fp_ = G.FilterParam(33, 5, 0.9, 40.0, typ)
        np.testing.assert_array_equal(G.ramp_taps(fp_),
                                      O.ramp_prep(O.geom(ns=33, nrow=5, ds=0.9, dsd=40.0, type=typ)))


def test_fp_boundary_views_follow_accumulated_angles(lib):
    """360 views over 180 deg: view 270 -> y-branch; 720 views: view 180 -> x-branch."""
    from ct_recon_extensions_amd import geometry as G
    for na, view, expect in ((360, 270, 0), (720, 180, 1)):
        p = G.ParallelProjection2DParam(25, na, 1.0, math.pi / na, 0.0, 0.0, 16, 16, 1.0, 1.0,
                                        0.0, 0.0)
        assert G.fp_tables(p)[2][view] == expect


def test_error_paths_return_codes(lib):
    from ct_recon_extensions_amd import _lib
    g = _geom(ns=8, nrow=1, ds=1.0, dsd=10.0, type=7)
    h = np.zeros(17)
    rc = lib.ctr_ramp_taps(C.byref(g), h.ctypes.data_as(C.c_void_p))
    assert rc == 1 and "unknown filter type" in _lib.last_error()
    rc = lib.ctr_fan_tables(None, None, None, None)
    assert rc == 1 and "null" in _lib.last_error()
    g = _geom(ns=8, na=4, nx=4, ny=6, ds=1.0, dx=1.0, dy=1.0)
    w, pos, usex = np.zeros(4), np.zeros(64), np.zeros(4, np.int32)
    rc = lib.ctr_fp_grad_tables(C.byref(g), w.ctypes.data_as(C.c_void_p),
                                pos.ctypes.data_as(C.c_void_p), usex.ctypes.data_as(C.c_void_p), 4)
    assert rc == 1 and "pos_stride" in _lib.last_error()
    # operators validate before touching the device
    rc = lib.ctr_fan_bp(C.byref(g), 5, C.c_void_p(1), C.c_void_p(1), 1, C.c_void_p(1),
                        C.c_void_p(1), C.c_void_p(1), 1.0, None)
    assert rc == 1 and "dtype" in _lib.last_error()
    rc = lib.ctr_par_bp(C.byref(g), 0, None, C.c_void_p(1), 1, C.c_void_p(1), C.c_void_p(1), None)
    assert rc == 1
    rc = lib.ctr_fp(C.byref(g), 0, C.c_void_p(1), C.c_void_p(1), 0, C.c_void_p(1), C.c_void_p(1),
                    C.c_void_p(1), None)
    assert rc == 1 and "nbatch" in _lib.last_error()
    rc = lib.ctr_ramp(C.byref(g), 0, C.c_void_p(1), C.c_void_p(1), 4, C.c_void_p(1), 9, None)
    assert rc == 1 and "weight_type" in _lib.last_error()


def test_fan_bp_launch_shape_is_validated_per_call_on_each_thread(lib):
    """The fan BP's launch shape travels with each call (ctr_fan_bp_ex): no
    process-wide setter exists, and two host threads passing different bad
    shapes at once each get their own status and message (thread-local),
    before any device work."""
    import threading

    from ct_recon_extensions_amd import _lib
    assert not hasattr(lib, "ctr_fan_bp_config")
    g = _geom(ns=8, na=4, nx=4, ny=4, ds=1.0, dx=1.0, dy=1.0, dso=50.0, dsd=90.0)
    cases = {"tile": _lib.CtrFanBpLaunch(3, 0, 0), "vchunk": _lib.CtrFanBpLaunch(0, 65, 0),
             "nb": _lib.CtrFanBpLaunch(0, 0, 3)}
    errors = []

    def worker(word, launch):
        for _ in range(300):
            rc = lib.ctr_fan_bp_ex(C.byref(g), 0, C.c_void_p(1), C.c_void_p(1), 1,
                                   C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), 1.0,
                                   C.byref(launch), None)
            msg = _lib.last_error()
            if rc != 1 or word not in msg:
                errors.append((word, rc, msg))
                return

    threads = [threading.Thread(target=worker, args=kv) for kv in cases.items()]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    # plain path: the reference grid's caps are checked the same way
    rc = lib.ctr_plain_fan_bp_grid(C.byref(g), 0, C.c_void_p(1), C.c_void_p(1), 1, C.c_void_p(1),
                                   C.c_void_p(1), C.c_void_p(1), 1.0, 32, 0, None)
    assert rc == 1 and "max_threads" in _lib.last_error()


def test_python_ops_refuse_cpu_tensors(lib):
    import torch
    import ct_recon_extensions_amd as F
    with pytest.raises(ValueError, match="GPU tensor"):
        F.fan_backprojection_2d(torch.zeros(1, 4, 8), [4, 4], [1, 1], [4, 8], 1.0)


def test_tables_equal_oracle_at_full_size():
    """The host tables at the workloads' sizes equal the oracle's (built by gcc
    like the reference, whose -O2/-O3 turns sin and cos of one angle into
    sincos; view 464 of 720 over 180 deg is an angle where sincos and cos
    give cosines one ulp apart)."""
    import ct_recon_extensions_amd.geometry as G
    n, na, ns = 512, 720, 729
    q = dict(ns=ns, na=na, ds=1.0, orbit=math.pi / na, nx=n, ny=n, dx=1.0, dy=1.0)
    g = O.geom(**q)
    pp = G.ParallelBackprojection2DParam(ns, na, 1.0, math.pi / na, 0.0, 0.0, n, n, 1.0, 1.0,
                                         0.0, 0.0)
    for a, b in zip(G.par_tables(pp), O.par_prep(g)):
        np.testing.assert_array_equal(a, b)
    pf = G.ParallelProjection2DParam(ns, na, 1.0, math.pi / na, 0.0, 0.0, n, n, 1.0, 1.0,
                                     0.0, 0.0)
    for a, b in zip(G.fp_tables(pf), O.fp_prep(g)):
        np.testing.assert_array_equal(a, b)
    gf = O.geom(ns=736, na=720, ds=1.30, orbit=2 * math.pi / 720, nx=512, ny=512, dx=0.7,
                dy=0.7, dso=595.0, dsd=1085.6)
    pfan = G.FanBackprojection2DParam(736, 720, 1.30, 2 * math.pi / 720, 0.0, 0.0, 512, 512,
                                      0.7, 0.7, 0.0, 0.0, 595.0, 1085.6)
    for a, b in zip(G.fan_tables(pfan), O.fan_prep(gf)):
        np.testing.assert_array_equal(a, b)
