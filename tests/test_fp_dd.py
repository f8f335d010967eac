"""Distance-driven parallel FP (method "disdriven") and the gradient the
reference defines for it, against the compiled reference's golden vectors and
the oracle: bit-identical in float32 and float64 (the kernels keep the
reference's operation order and float / double mix)."""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import load_golden, randn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def F():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import ct_recon_extensions_amd as F
    from ct_recon_extensions_amd import _lib
    _lib.load()
    return F


def cuda(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t.to(dtype) if dtype is not None else t).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def kw(geo):
    return dict(img_shape=[geo["ny"], geo["nx"]], img_space=[geo["dy"], geo["dx"]],
                proj_shape=[geo["na"], geo["ns"]], channel_space=geo["ds"],
                img_offset=[geo["offset_y"], geo["offset_x"]], channel_offset=geo["offset_s"],
                orbit_start=geo["orbit_start"], orbit=geo["orbit"], method="disdriven")


@pytest.mark.parametrize("name", ["fpdd_small", "fpdd_circle"])
def test_fp_dd_golden_bitwise(F, name):
    geo, d = load_golden(name)
    out = host(F.parallel_projection_2d(cuda(d["img"][None]), **kw(geo)))
    np.testing.assert_array_equal(out[0], d["fp_f32"])
    out64 = host(F.parallel_projection_2d(cuda(d["img"][None], torch.float64), **kw(geo)))
    np.testing.assert_array_equal(out64[0], d["fp_f64"])
    g = host(F.parallel_projection_2d_grad(cuda(d["proj"][None]), **kw(geo)))
    np.testing.assert_array_equal(g[0], d["grad_f32"])
    g64 = host(F.parallel_projection_2d_grad(cuda(d["proj"][None], torch.float64), **kw(geo)))
    np.testing.assert_array_equal(g64[0], d["grad_f64"])


@pytest.mark.parametrize("nbatch", [1, 2, 3, 6])
def test_fp_dd_batches_vs_oracle_bitwise(F, nbatch):
    """Batches through every slices-per-thread width (1, 2, 4 and a partial
    group), a non-square image with offsets for the FP, and the 720-view
    512^2 geometry's angles on a smaller image."""
    geos = [dict(ns=61, na=33, ds=1.0, orbit=math.pi / 33, offset_s=0.2, orbit_start=0.1,
                 nx=30, ny=18, dx=1.0, dy=1.3, offset_x=0.3, offset_y=-0.4),
            dict(ns=181, na=720, ds=1.0, orbit=math.pi / 720, offset_s=0.0, orbit_start=0.0,
                 nx=128, ny=128, dx=1.0, dy=1.0, offset_x=0.0, offset_y=0.0)]
    for geo in geos:
        g = O.geom(**geo)
        t = O.fp_dd_prep(g)
        img = randn(300 + nbatch, (nbatch, g.ny, g.nx))
        out = host(F.parallel_projection_2d(cuda(img), **kw(geo)))
        for b in range(nbatch):
            np.testing.assert_array_equal(out[b], O.fp_dd(g, img[b], t))
        if geo["nx"] == geo["ny"]:
            tg = O.fp_dd_grad_prep(g)
            proj = randn(400 + nbatch, (nbatch, g.na, g.ns))
            gout = host(F.parallel_projection_2d_grad(cuda(proj), **kw(geo)))
            for b in range(nbatch):
                np.testing.assert_array_equal(gout[b], O.fp_dd_grad(g, proj[b], tg))


def test_fp_dd_autograd_uses_the_reference_gradient(F):
    """Autograd's backward of the distance-driven FP is the reference's
    distance-driven gradient op on the output gradient (bitwise), and the
    pair's inner products match the oracle pair's: <FP x, y> and <x, G y>
    differ by the same amount on the GPU as on the host (the reference's
    gradient is an approximate transpose)."""
    geo, _ = load_golden("fpdd_circle")
    g = O.geom(**geo)
    x = randn(501, (1, g.ny, g.nx)).astype(np.float64)
    y = randn(502, (1, g.na, g.ns)).astype(np.float64)
    xt = cuda(x).requires_grad_(True)
    fx = F.parallel_projection_2d(xt, **kw(geo))
    fx.backward(cuda(y))
    gy = F.parallel_projection_2d_grad(cuda(y), **kw(geo))
    np.testing.assert_array_equal(host(xt.grad), host(gy))
    lhs, rhs = float((host(fx) * y).sum()), float((x * host(gy)).sum())
    fo = O.fp_dd(g, x[0], O.fp_dd_prep(g))
    go = O.fp_dd_grad(g, y[0], O.fp_dd_grad_prep(g))
    lo, ro = float((fo * y[0]).sum()), float((x[0] * go).sum())
    assert abs(lhs - lo) <= 1e-12 * abs(lo) and abs(rhs - ro) <= 1e-12 * abs(ro)
