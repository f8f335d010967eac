"""Synthetic-input generators: the torch fan sinogram equals the numpy one
(FP64), and the analytic parallel projection of a disk is its chord."""
import math

import numpy as np

from ct_recon_extensions_amd import phantom as P


def test_torch_fan_sinogram_matches_numpy():
    ell = P.shepp_logan(250.0)
    kw = dict(na=90, ns=101, ds=1.3, orbit=2 * math.pi / 90, dso=595.0, dsd=1085.6,
              orbit_start=0.2, offset_s=0.3)
    a = P.fan_sinogram(ell, **kw)
    b = P.fan_sinogram_torch(ell, device="cpu", **kw).numpy()
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))


def test_disk_projection_is_chord_length():
    ell = np.array([[3.0, -2.0, 40.0, 40.0, 0.7, 1.5]])
    theta = np.linspace(0, math.pi, 7)[:, None]
    t = np.linspace(-60, 60, 31)[None, :]
    p = P.parallel_projection(ell, theta, t)
    tau = t - 3.0 * np.cos(theta) + 2.0 * np.sin(theta)
    want = 1.5 * 2 * np.sqrt(np.maximum(40.0 ** 2 - tau ** 2, 0))
    np.testing.assert_allclose(p, want, atol=1e-9)
