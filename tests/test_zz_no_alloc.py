"""ctr_fp_t_noalloc (the FP transpose without device allocation) against
ctr_fp_t through the C ABI: two calls in a row, the second on a smaller
problem, each bitwise equal to the allocating entry point (same kernels,
same summation order) and within float32 rounding of the oracle; a request
larger than the static scratch is refused before any device work."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import randn, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from ct_recon_extensions_amd import _lib
    return _lib.load()


def _run(lib, name, p, proj, tables):
    from ct_recon_extensions_amd import _lib
    n = proj.shape[0]
    out = torch.full((n, p.ny, p.nx), float("nan"), device="cuda")
    w, pos, usex = tables
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(lib, name)(C.byref(p.ctr()), 0, C.c_void_p(proj.data_ptr()),
                            C.c_void_p(out.data_ptr()), n, C.c_void_p(w.data_ptr()),
                            C.c_void_p(pos.data_ptr()), C.c_void_p(usex.data_ptr()),
                            max(p.nx, p.ny), stream)
    _lib.check(rc, name)
    torch.cuda.synchronize()
    return out


def test_fp_t_noalloc_matches_fp_t_on_two_calls(lib):
    from ct_recon_extensions_amd.geometry import ParallelProjection2DParam, device_tables
    # the second call is smaller: it reuses the same scratch after the first
    for nx, ny, na, ns, nb, seed in ((96, 80, 90, 137, 8, 1), (40, 40, 33, 61, 3, 2)):
        p = ParallelProjection2DParam(ns, na, 1.0, math.pi / na, 0.25, 0.1, nx, ny, 1.0, 1.0,
                                      0.3, -0.2)
        tables = device_tables("fp_grad", p, torch.device("cuda"))
        proj_h = randn(seed, (nb, na, ns))
        proj = torch.from_numpy(proj_h).cuda()
        ref = _run(lib, "ctr_fp_t", p, proj, tables)
        got = _run(lib, "ctr_fp_t_noalloc", p, proj, tables)
        assert torch.equal(got, ref), (nx, na, nb)
        g = O.geom(ns=ns, na=na, ds=1.0, orbit=math.pi / na, offset_s=0.25, orbit_start=0.1,
                   nx=nx, ny=ny, dx=1.0, dy=1.0, offset_x=0.3, offset_y=-0.2)
        if nx == ny:  # the oracle's table layout is the reference's (nx == ny)
            t = O.fp_grad_prep(g)
            for b in range(nb):
                assert rel_l2(got[b].cpu().numpy(), O.fp_grad(g, proj_h[b], t)) < 1e-5, b


def test_fp_t_noalloc_refuses_more_than_its_scratch(lib):
    from ct_recon_extensions_amd import _lib
    g = _lib.CtrGeom()
    g.ns, g.na, g.ds, g.orbit = 2900, 2880, 1.0, math.pi / 2880
    g.nx, g.ny, g.dx, g.dy = 2048, 2048, 1.0, 1.0
    one = C.c_void_p(1)
    rc = lib.ctr_fp_t_noalloc(C.byref(g), 0, one, one, 8, one, one, one, 2048, None)
    assert rc == 2 and "scratch" in _lib.last_error()
