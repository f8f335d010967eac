"""The reference's PyTorch operator pair (pytorch/bp_par_2d_ops.cc):
bppar2d_prep_forward / bppar2d_forward, same names and argument orders."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import load_golden, randn, rel_l2
from tests.golden.make_golden import PAR_SMALL

ARGS = ("ns", "na", "ds", "orbit", "offset_s", "orbit_start", "nx", "ny", "dx", "dy",
        "offset_x", "offset_y")


def positional(geo):
    return [geo[k] for k in ARGS] + [-1.0]


def test_prep_tables_match_oracle():
    import ct_recon_extensions_amd as F
    b1, b2, b3 = F.bppar2d_prep_forward("pixdriven", *positional(PAR_SMALL))
    xcos, ysin = O.par_prep(O.geom(**PAR_SMALL))
    assert b1.shape == (PAR_SMALL["nx"], PAR_SMALL["na"]) and b1.dtype == torch.float64
    assert b2.shape == (PAR_SMALL["ny"], PAR_SMALL["na"]) and b2.dtype == torch.float64
    assert b3.dtype == torch.int32 and b1.device.type == "cpu"
    np.testing.assert_array_equal(b1.numpy().reshape(-1), xcos)
    np.testing.assert_array_equal(b2.numpy().reshape(-1), ysin)


def test_unknown_method_raises_reference_error():
    import ct_recon_extensions_amd as F
    with pytest.raises(RuntimeError, match="Only pixdriven is available now"):
        F.bppar2d_prep_forward("raydriven", *positional(PAR_SMALL))
    with pytest.raises(RuntimeError, match="Only pixdriven is available now"):
        F.bppar2d_forward(torch.zeros(1), None, None, None, "raydriven",
                          *positional(PAR_SMALL))


@pytest.mark.gpu
def test_forward_matches_golden_and_functional_op():
    import ct_recon_extensions_amd as F
    geo, d = load_golden("par_small")
    bufs = F.bppar2d_prep_forward("pixdriven", *positional(geo))
    proj = torch.from_numpy(d["proj"]).float().cuda()[None]
    out = F.bppar2d_forward(proj, *bufs, "pixdriven", *positional(geo))
    assert out.shape == (1, geo["ny"], geo["nx"])
    assert rel_l2(out[0].double().cpu().numpy(), d["bp_f32"]) < 1e-5
    kw = dict(img_shape=[geo["ny"], geo["nx"]], img_space=[geo["dy"], geo["dx"]],
              proj_shape=[geo["na"], geo["ns"]], channel_space=geo["ds"],
              img_offset=[geo["offset_y"], geo["offset_x"]], channel_offset=geo["offset_s"],
              orbit_start=geo["orbit_start"], orbit=geo["orbit"])
    ref = F.parallel_backprojection_2d(proj, **kw)
    assert torch.equal(out, ref)
    # float64, tables already on the device, [na, ns] input
    dev_bufs = F.bppar2d_prep_forward("pixdriven", *positional(geo), device="cuda")
    out64 = F.bppar2d_forward(torch.from_numpy(d["proj"]).double().cuda(), *dev_bufs,
                              "pixdriven", *positional(geo))
    assert out64.shape == (geo["ny"], geo["nx"]) and out64.dtype == torch.float64
    assert rel_l2(out64.cpu().numpy(), d["bp_f64"]) < 1e-12


@pytest.mark.gpu
def test_forward_gradient_is_transpose():
    import ct_recon_extensions_amd as F
    geo = PAR_SMALL
    bufs = F.bppar2d_prep_forward("pixdriven", *positional(geo), device="cuda")
    proj = torch.from_numpy(randn(31, (2, geo["na"], geo["ns"]))).cuda().requires_grad_()
    g = torch.from_numpy(randn(32, (2, geo["ny"], geo["nx"]))).cuda()
    out = F.bppar2d_forward(proj, *bufs, "pixdriven", *positional(geo))
    (out * g).sum().backward()
    # <BP x, g> = <x, BP^T g>
    lhs = float((out.detach().double() * g.double()).sum())
    rhs = float((proj.detach().double() * proj.grad.double()).sum())
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
