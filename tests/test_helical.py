"""Helical rebinning, 360LI and 180LI (src/helical_interpolation.cc:81-269).

CPU: the oracle against the compiled reference's golden vectors (bitwise),
the reference's two errors, and closed forms.  GPU: the HIP kernel through
the torch op and the C ABI, bit-identical to the oracle in float32 and
float64, on the golden scans, batches, random geometries (every branch:
one window off the detector, the 180LI conjugate channel off the detector,
thick and zero-thickness slices) and a clinical-size scan.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import GOLDEN, randn

GOLDEN_CASES = ("heli_small", "heli_pitch", "heli_gap", "heli_short")


def _golden(name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    geo = json.loads(str(d.pop("geom")))
    h = O.heli(**geo)
    proj = randn(int(d.pop("proj_seed")), (h.na, h.nt, h.ns))
    return geo, h, proj, d


# ------------------------------------------------------------------- CPU
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_matches_reference_golden(name):
    _, h, proj, d = _golden(name)
    for mode in (360, 180):
        for suf, x in (("f32", proj), ("f64", proj.astype(np.float64))):
            out, code = O.helical(h, x, mode)
            assert (code == 1) == bool(d[f"ok_{mode}_{suf}"]), (name, mode, suf, code)
            if code == 1:
                np.testing.assert_array_equal(out, d[f"out_{mode}_{suf}"])


def test_oracle_error_codes():
    """heli_gap: samples with no detector row in either window (-1, the
    reference's "No row data available!"); heli_short: a slice too close to
    the scan start for 360LI (0, "View not enough for this slice!") that
    180LI, needing half the views, still covers."""
    _, h, proj, _ = _golden("heli_gap")
    for mode in (360, 180):
        out, code = O.helical(h, proj, mode)
        assert code == -1 and np.isnan(out).any() and not np.isnan(out).all()
    _, h, proj, _ = _golden("heli_short")
    assert O.helical(h, proj, 360)[1] == 0
    assert O.helical(h, proj, 180)[1] == 1


def _dense_scan(vpr=32, nt=16, ns=21, pitch=0.3, nz=6):
    """Pitch 0.3 on 16 rows: every window of both views stays on the
    detector, so every sample interpolates between two views."""
    coll = nt * 500.0 / 1000.0
    m = pitch * coll / vpr
    return O.heli(ns=ns, nt=nt, na=4 * vpr, ds=20.0, dt=1.0, offset_s=0.2, offset_t=0.3,
                  orbit_start=0.0, orbit=2 * math.pi / vpr, couch_begin=3.0, couch_mov=m,
                  view_per_rot=vpr, nz=nz, dz=0.37, z0=3.0 + 1.6 * vpr * m, thickness=1.0,
                  dso=500.0, dsd=1000.0)


@pytest.mark.parametrize("mode", [360, 180])
def test_oracle_view_linear_scan_recovers_slice_position(mode):
    """A scan whose every sample equals its view index v: both window means
    are the (interpolated) view indices c1, c2, and the z interpolation
    (z1 c2 - z2 c1)/(z1 - z2) with z_i = zc - c_i m returns zc/m, the
    fractional view at which the source passes the slice -- for every view
    and channel.  A wrong weight or sign breaks this.  (No channel offset:
    every 180LI mirrored channel stays on the detector.)"""
    h = _dense_scan()
    h.offset_s = 0.0
    v = np.arange(h.na, dtype=np.float64)[:, None, None]
    proj = np.broadcast_to(v, (h.na, h.nt, h.ns)).copy()
    out, code = O.helical(h, proj, mode)
    assert code == 1
    z = h.z0 + h.dz * np.arange(h.nz)
    want = (z - h.couch_begin) / h.couch_mov
    np.testing.assert_allclose(out, np.broadcast_to(want[:, None, None], out.shape),
                               rtol=0, atol=1e-9)


def test_oracle_z_invariant_scan_is_reproduced_by_360li():
    """A scan that does not change along z or with the view (proj = f(s))
    comes back unchanged from 360LI: window means and z interpolation of
    equal values."""
    h = _dense_scan()
    f = np.random.default_rng(5).standard_normal(h.ns)
    proj = np.broadcast_to(f, (h.na, h.nt, h.ns)).copy()
    out, code = O.helical(h, proj, 360)
    assert code == 1
    np.testing.assert_allclose(out, np.broadcast_to(f, out.shape), rtol=0, atol=1e-12)


def test_oracle_180li_conjugate_channel_off_detector_keeps_the_view():
    """180LI reads the conjugate ray at the mirrored channel 2 cs - is; with
    cs = 12.5 on 21 channels the mirror of is < 5 is off the detector and the
    reference keeps the output view's own value there.  On the view-linear
    scan those channels return a whole view index congruent to the output
    view iv (mod view_per_rot) within half a rotation of the slice; every
    other channel returns the slice position zc/m."""
    h = _dense_scan(ns=21)
    h.offset_s = 2.5
    v = np.arange(h.na, dtype=np.float64)[:, None, None]
    proj = np.broadcast_to(v, (h.na, h.nt, h.ns)).copy()
    out, code = O.helical(h, proj, 180)
    assert code == 1
    z = h.z0 + h.dz * np.arange(h.nz)
    vz = ((z - h.couch_begin) / h.couch_mov)[:, None]
    off = out[:, :, :5]
    iv = np.arange(h.view_per_rot)[None, :, None]
    assert np.array_equal(off, np.round(off))
    assert np.all((off - iv) % h.view_per_rot == 0)
    assert np.all(np.abs(off - vz[:, :, None]) <= h.view_per_rot / 2 + 0.5)
    rest = out[:, :, 5:]
    np.testing.assert_allclose(rest, np.broadcast_to(vz[:, :, None], rest.shape), rtol=0,
                               atol=1e-9)


# ------------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def F():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import ct_recon_extensions_amd as F
    from ct_recon_extensions_amd import _lib
    _lib.load()
    return F


def _param(F, h):
    return F.HelicalInterpolationParam(**{k: getattr(h, k) for k, _ in O.Heli._fields_})


def _run(F, h, x, mode, check=True):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    out = F.helical_interpolation(t, _param(F, h), mode, check=check)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["heli_small", "heli_pitch"])
def test_gpu_golden_bitwise(F, name):
    _, h, proj, d = _golden(name)
    for mode in (360, 180):
        np.testing.assert_array_equal(_run(F, h, proj, mode), d[f"out_{mode}_f32"])
        np.testing.assert_array_equal(_run(F, h, proj.astype(np.float64), mode),
                                      d[f"out_{mode}_f64"])


@pytest.mark.gpu
def test_gpu_reference_errors(F):
    """The reference's two exceptions: a short slice raises before any work;
    a sample without rows raises after the launch with check, and is NaN
    exactly where the oracle's is without it."""
    _, h, proj, _ = _golden("heli_short")
    with pytest.raises(ValueError, match="View not enough"):
        _run(F, h, proj, 360)
    _, h, proj, _ = _golden("heli_gap")
    for mode in (360, 180):
        with pytest.raises(RuntimeError, match="No row data"):
            _run(F, h, proj, mode)
        got = _run(F, h, proj, mode, check=False)
        want, code = O.helical(h, proj, mode)
        assert code == -1
        np.testing.assert_array_equal(got, want)


@pytest.mark.gpu
def test_gpu_batches_are_independent_scans(F):
    _, h, _, _ = _golden("heli_pitch")
    x = randn(61, (3, h.na, h.nt, h.ns))
    for mode in (360, 180):
        got = _run(F, h, x, mode)
        for b in range(3):
            np.testing.assert_array_equal(got[b], O.helical(h, x[b], mode)[0])


def _random_geometry(rng):
    vpr = int(rng.integers(8, 70))
    nt = int(rng.integers(1, 12))
    ns = int(rng.integers(2, 150))
    dso = float(rng.uniform(300, 600))
    dsd = dso * float(rng.uniform(1.5, 2.2))
    m = nt * dso / dsd / vpr * float(rng.uniform(0.3, 2.0))
    na = int(rng.integers(3 * vpr, 5 * vpr))
    cb = float(rng.uniform(-20, 20))
    nz = int(rng.integers(1, 6))
    dz = float(rng.uniform(-0.5, 0.5))
    return O.heli(ns=ns, nt=nt, na=na, ds=float(rng.uniform(0.5, 3)), dt=1.0,
                  offset_s=float(rng.uniform(-3, 3)), offset_t=float(rng.uniform(-1, 1)),
                  orbit_start=0.0, orbit=2 * math.pi / vpr, couch_begin=cb, couch_mov=m,
                  view_per_rot=vpr, nz=nz, dz=dz,
                  z0=cb + m * float(rng.uniform(1.2 * vpr, 1.8 * vpr)),
                  thickness=float(rng.choice([0.0, rng.uniform(0.1, 3), rng.uniform(3, 20)])),
                  dso=dso, dsd=dsd)


@pytest.mark.gpu
def test_gpu_random_geometries_vs_oracle(F):
    rng = np.random.default_rng(2024)
    seen = set()
    for trial in range(40):
        h = _random_geometry(rng)
        mode = (360, 180)[trial % 2]
        dt = (np.float32, np.float64)[(trial // 2) % 2]
        x = rng.standard_normal((h.na, h.nt, h.ns)).astype(dt)
        want, code = O.helical(h, x, mode)
        seen.add(code)
        if code == 0:
            with pytest.raises(ValueError):
                _run(F, h, x, mode)
            continue
        got = _run(F, h, x, mode, check=False)
        np.testing.assert_array_equal(got, want, err_msg=f"trial {trial} mode {mode}")
    assert 1 in seen


def clinical_scan(nz, na, dz=1.0):
    """736 channels of 1 mm and 32 rows at dso 595 / dsd 1085.6 mm (17.5 mm
    collimation at the isocentre), 1152 views per rotation, pitch 1, 1 mm
    slices starting just over one rotation into the scan."""
    m = 32 * 595.0 / 1085.6 / 1152
    return dict(ns=736, nt=32, na=na, ds=1.0, dt=1.0, offset_s=1.25, offset_t=0.0,
                orbit_start=0.0, orbit=2 * math.pi / 1152, couch_begin=0.0, couch_mov=m,
                view_per_rot=1152, nz=nz, dz=dz, z0=1160 * m, thickness=1.0, dso=595.0,
                dsd=1085.6)


@pytest.mark.gpu
def test_gpu_clinical_size_vs_oracle(F):
    """736 channels x 32 rows, 1152 views per rotation over 3.5 rotations,
    pitch 1, 16 slices of 1 mm: the bench workload's geometry with fewer
    slices, bit-identical to the oracle in float32 for both modes."""
    h = O.heli(**clinical_scan(nz=16, na=3456))
    x = np.random.default_rng(71).standard_normal((h.na, h.nt, h.ns), dtype=np.float32)
    for mode in (360, 180):
        want, code = O.helical(h, x, mode)
        assert code == 1
        np.testing.assert_array_equal(_run(F, h, x, mode), want)


@pytest.mark.gpu
def test_c_abi_direct(F):
    """ctr_helical_tables + ctr_helical through ctypes, on a caller-owned
    stream: the torch op's result, and argument errors as codes."""
    import torch
    from ct_recon_extensions_amd import _lib
    lib = _lib.load()
    _, h, proj, d = _golden("heli_small")
    g = _lib.CtrHelicalGeom(*[getattr(h, k) for k, _ in O.Heli._fields_])
    tab = np.zeros(3 * h.ns + 4 * h.nz)
    assert lib.ctr_helical_tables(C.byref(g), 180, tab.ctypes.data_as(C.c_void_p)) == 0
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        x = torch.from_numpy(proj).cuda()
        t = torch.from_numpy(tab).cuda()
        out = torch.empty((h.nz, h.view_per_rot, h.ns), dtype=torch.float32, device="cuda")
        st = torch.zeros(1, dtype=torch.int32, device="cuda")
        rc = lib.ctr_helical(C.byref(g), 180, _lib.CTR_F32, C.c_void_p(x.data_ptr()),
                             C.c_void_p(out.data_ptr()), 1, C.c_void_p(t.data_ptr()),
                             C.c_void_p(st.data_ptr()), C.c_void_p(s.cuda_stream))
    assert rc == 0
    s.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), d["out_180_f32"])
    assert int(st.item()) == 0
    vp = C.c_void_p(x.data_ptr())
    CTR_ERR_ARG = 1
    assert lib.ctr_helical(C.byref(g), 90, 0, vp, vp, 1, vp, None, None) == CTR_ERR_ARG
    assert lib.ctr_helical(C.byref(g), 360, 0, vp, vp, 0, vp, None, None) == CTR_ERR_ARG
    assert lib.ctr_helical(C.byref(g), 360, 0, None, vp, 1, vp, None, None) == CTR_ERR_ARG
