"""The reference's PyTorch operator pair for the parallel-beam BP
(pytorch/bp_par_2d_ops.cc:17-72), same names and argument orders:

* ``bppar2d_prep_forward(method, ns, na, ds, orbit, offset_s, orbit_start,
  nx, ny, dx, dy, offset_x, offset_y, fov)`` -> ``[buffer1, buffer2,
  buffer3]``: the FP64 tables of ParallelBackprojection2DPixDrivenPrep
  (src/bp_par_2d.cc:18-49), ``xcos[ix][ia]`` and ``ysin[iy][ia]``, and the
  int buffer the prep's signature carries (unused by the pixel-driven BP);
* ``bppar2d_forward(input, buffer1, buffer2, buffer3, method, ns, ..., fov)``
  -> the backprojected image (src/bp_par_2d.cc:51-90) from those tables.

The reference file is a stub that does not compile (``prep`` and
``buffer`` are undeclared, the buffer shapes count bytes, the forward
returns nothing); these functions implement what it declares.  Unknown
methods raise RuntimeError with the reference's message.  ``input`` is a
CUDA tensor [N, na, ns, 1], [N, na, ns] or [na, ns], float32 or float64; the
output has the matching image shape.  The forward is differentiable (its
backward is the exact transpose kernel).
"""
import torch

from .functional import _ParBP, _to_batch
from .geometry import ParallelBackprojection2DParam, par_tables

_METHOD_ERROR = "Unrecognised backprojection method. Only pixdriven is available now."


def _param(ns, na, ds, orbit, offset_s, orbit_start, nx, ny, dx, dy, offset_x, offset_y, fov):
    return ParallelBackprojection2DParam(int(ns), int(na), float(ds), float(orbit),
                                         float(offset_s), float(orbit_start), int(nx), int(ny),
                                         float(dx), float(dy), float(offset_x), float(offset_y),
                                         float(fov))


def bppar2d_prep_forward(method, ns, na, ds, orbit, offset_s, orbit_start, nx, ny, dx, dy,
                         offset_x, offset_y, fov, device="cpu"):
    """Tables of the pixel-driven parallel BP (pytorch/bp_par_2d_ops.cc:17-47).
    The reference fills them on the host (calculate_on_cpu); ``device``
    chooses where the returned tensors live."""
    if method != "pixdriven":
        raise RuntimeError(_METHOD_ERROR)
    p = _param(ns, na, ds, orbit, offset_s, orbit_start, nx, ny, dx, dy, offset_x, offset_y, fov)
    xcos, ysin = par_tables(p)
    buffer1 = torch.from_numpy(xcos.reshape(p.nx, p.na)).to(device)
    buffer2 = torch.from_numpy(ysin.reshape(p.ny, p.na)).to(device)
    buffer3 = torch.zeros(1, dtype=torch.int32, device=device)
    return [buffer1, buffer2, buffer3]


def _table(buf, n, na, name, device):
    if not isinstance(buf, torch.Tensor) or buf.numel() != n * na:
        raise ValueError(f"bppar2d_forward: {name} must hold {n} x {na} FP64 table entries")
    return buf.to(device=device, dtype=torch.float64).contiguous().reshape(-1)


def bppar2d_forward(input, buffer1, buffer2, buffer3, method, ns, na, ds, orbit, offset_s,
                    orbit_start, nx, ny, dx, dy, offset_x, offset_y, fov):
    """Pixel-driven parallel BP with caller-held tables
    (pytorch/bp_par_2d_ops.cc:50-72)."""
    if method != "pixdriven":
        raise RuntimeError(_METHOD_ERROR)
    p = _param(ns, na, ds, orbit, offset_s, orbit_start, nx, ny, dx, dy, offset_x, offset_y, fov)
    x, restore = _to_batch(input, p.na, p.ns, "input")
    tables = (_table(buffer1, p.nx, p.na, "buffer1", x.device),
              _table(buffer2, p.ny, p.na, "buffer2", x.device))
    return restore(_ParBP.apply(x, p, tables, False))
