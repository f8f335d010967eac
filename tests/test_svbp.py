"""Single-view parallel BP (src/bp_par_2d_sv.cc:54-92) on the GPU against the
compiled reference's outputs and the oracle: bit-identical in float32 and
float64 (the kernel uses the reference's FP64 arithmetic without
contraction)."""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import load_golden, randn

pytestmark = pytest.mark.gpu


def kwargs(geo):
    return dict(img_shape=[geo["ny"], geo["nx"]], img_space=[geo["dy"], geo["dx"]],
                proj_shape=[geo["na"], geo["ns"]], channel_space=geo["ds"],
                img_offset=[geo["offset_y"], geo["offset_x"]], channel_offset=geo["offset_s"],
                orbit_start=geo["orbit_start"], orbit=geo["orbit"])


@pytest.fixture(scope="module")
def F():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import ct_recon_extensions_amd as F
    return F


def test_svbp_golden_bit_exact(F):
    geo, d = load_golden("svbp_small")
    out = F.parallel_single_view_bp_2d(torch.from_numpy(d["proj"]).cuda(), **kwargs(geo))
    assert out.shape == (geo["ny"], geo["nx"], geo["na"])
    np.testing.assert_array_equal(out.cpu().numpy(), d["svbp_f32"])
    out64 = F.parallel_single_view_bp_2d(torch.from_numpy(d["proj"]).double().cuda()[None],
                                         **kwargs(geo))
    assert out64.shape == (1, geo["ny"], geo["nx"], geo["na"])
    np.testing.assert_array_equal(out64[0].cpu().numpy(), d["svbp_f64"])


# 100 views (chunks of 32 and a partial one), 97 x 61 pixels (partial last
# pixel tile), a detector narrower than the image, three slices.
GEO = dict(ns=90, na=100, ds=1.1, orbit=math.pi / 100, offset_s=0.35, orbit_start=0.15,
           nx=97, ny=61, dx=1.3, dy=1.7, offset_x=-0.4, offset_y=0.6)


# na = 100: a full 64-view chunk and a partial one; na = 37: one partial
# chunk, and rows that are not 16-byte multiples (element-wise stores instead
# of 16-byte ones).  The 97 x 61 image leaves partial 8 x 8 pixel blocks.
@pytest.mark.parametrize("na", [100, 37])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_svbp_batch_bit_exact_vs_oracle(F, dtype, na):
    geo = dict(GEO, na=na, orbit=math.pi / na)
    g = O.geom(**geo)
    t = O.par_prep(g)
    proj = randn(77, (3, g.na, g.ns), dtype)
    out = F.parallel_single_view_bp_2d(torch.from_numpy(proj).cuda()[..., None], **kwargs(geo))
    assert out.shape == (3, g.ny, g.nx, g.na)
    out = out.cpu().numpy()
    for b in range(3):
        ref = O.par_svbp(g, proj[b], t)
        assert (ref == 0).sum() > 300 and (ref != 0).sum() > 1000
        np.testing.assert_array_equal(out[b], ref, err_msg=f"slice {b}")


def test_svbp_views_sum_to_parallel_bp(F):
    proj = torch.from_numpy(randn(78, (2, GEO["na"], GEO["ns"]), np.float64)).cuda()
    sv = F.parallel_single_view_bp_2d(proj, **kwargs(GEO))
    bp = F.parallel_backprojection_2d(proj, **kwargs(GEO))
    err = (GEO["orbit"] * sv.sum(-1) - bp).abs().max() / bp.abs().max()
    assert float(err) < 1e-12


def test_svbp_rejects_other_methods(F):
    with pytest.raises(ValueError):
        F.parallel_single_view_bp_2d(torch.zeros(GEO["na"], GEO["ns"], device="cuda"),
                                     **kwargs(GEO), method="raydriven")
