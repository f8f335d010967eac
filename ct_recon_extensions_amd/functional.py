"""PyTorch operators of the 2-D CT hot path, on the HIP library.

Names, argument order and keyword defaults follow the reference's Python
front-end python/tf_ctrecon.py:112-299 (TensorFlow there, torch here); the
tensors use the same NHWC layout, [N, H, W, 1] (``[N, H, W]`` and ``[H, W]``
are accepted too), with H = views and W = channels for sinograms and
H = ny, W = nx for images.  Attribute mapping follows the reference's TF ops
(e.g. tensorflow/bp_fan_2d_ops.cu:156-159): img_shape = [ny, nx],
img_space = [dy, dx], img_offset = [offset_y, offset_x],
proj_shape = [na, ns].

Every operator is a torch.autograd.Function whose backward is the
hand-written transpose kernel of its forward.  All work is enqueued on
torch's current stream of the tensor's device; float32 and float64 tensors
are supported.  There is no CPU path: tensors must live on the GPU.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Sequence

import torch

from . import _lib
from .geometry import (RC_STEP, FanBackprojection2DAngleParam, FanBackprojection2DParam,
                       FanWeightingParam, FilterParam, HelicalInterpolationParam,
                       ParallelBackprojection2DParam, ParallelProjection2DParam, device_tables)

__all__ = [
    "ramp_filter", "ramp_filter_grad", "parallel_backprojection_2d",
    "parallel_backprojection_2d_grad", "parallel_projection_2d", "parallel_projection_2d_grad",
    "fan_backprojection_2d", "fan_backprojection_2d_grad", "fan_backprojection_2d_angle",
    "fan_backprojection_2d_angle_grad", "fan_weighting", "fan_weighting_grad",
    "helical_interpolation",
]

_FILTER_TYPES = {"par": 0, "fan": 1, "flat": 2}
_WEIGHT_TYPES = {"fan": 1, "flat": 2}


# ---------------------------------------------------------------- helpers
def _dtype_code(t: torch.Tensor) -> int:
    if t.dtype == torch.float32:
        return _lib.CTR_F32
    if t.dtype == torch.float64:
        return _lib.CTR_F64
    raise TypeError(f"expected float32 or float64, got {t.dtype}")


def _stream(t: torch.Tensor):
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _ptr(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


def _require_gpu(t: torch.Tensor, name: str):
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor (this library has no CPU path)")


def _call_batched(name: str, n: int, *args) -> None:
    """Call a batched C ABI operator; an empty batch is a no-op (the
    reference's per-slice loop runs zero times)."""
    if n > 0:
        _lib.call(name, *args)


def _to_batch(x: torch.Tensor, h: int, w: int, name: str):
    """View x as contiguous [N, h, w]; return it and a function restoring the
    caller's rank for an output of spatial shape (h2, w2)."""
    _require_gpu(x, name)
    if x.dim() == 4:
        if x.shape[3] != 1:
            raise ValueError(f"{name}: channel dimension must be 1 (NHWC), got {tuple(x.shape)}")
        n = x.shape[0]
        restore = lambda y: y.unsqueeze(-1)  # noqa: E731
    elif x.dim() == 3:
        n = x.shape[0]
        restore = lambda y: y  # noqa: E731
    elif x.dim() == 2:
        n = 1
        restore = lambda y: y[0]  # noqa: E731
    else:
        raise ValueError(f"{name}: expected [N,H,W,1], [N,H,W] or [H,W], got {tuple(x.shape)}")
    if x.shape[-3 if x.dim() == 4 else -2] != h or x.shape[-2 if x.dim() == 4 else -1] != w:
        raise ValueError(f"{name}: expected spatial shape ({h}, {w}), got {tuple(x.shape)}")
    return x.reshape(n, h, w).contiguous(), restore


# ------------------------------------------------------ fan-beam BP kernels
def _fan_bp(proj3, p, tables, scale):
    n = proj3.shape[0]
    out = torch.empty((n, p.ny, p.nx), dtype=proj3.dtype, device=proj3.device)
    g = p.ctr()
    xpos, ypos, sc = tables
    _call_batched("ctr_fan_bp", n, C.byref(g), _dtype_code(proj3), _ptr(proj3), _ptr(out), n,
                  _ptr(xpos), _ptr(ypos), _ptr(sc), scale, _stream(proj3))
    return out


def _fan_bpt(img3, p, tables, scale, deterministic):
    n = img3.shape[0]
    out = torch.empty((n, p.na, p.ns), dtype=img3.dtype, device=img3.device)
    g = p.ctr()
    xpos, ypos, sc = tables
    _call_batched("ctr_fan_bp_t", n, C.byref(g), _dtype_code(img3), _ptr(img3), _ptr(out), n,
                  _ptr(xpos), _ptr(ypos), _ptr(sc), scale, int(bool(deterministic)),
                  _stream(img3))
    return out


class _FanBP(torch.autograd.Function):
    @staticmethod
    def forward(ctx, proj3, p, tables, scale, deterministic):
        ctx.p, ctx.tables, ctx.scale, ctx.det = p, tables, scale, deterministic
        return _fan_bp(proj3, p, tables, scale)

    @staticmethod
    def backward(ctx, g):
        return _FanBPT.apply(g.contiguous(), ctx.p, ctx.tables, ctx.scale, ctx.det), None, None, None, None


class _FanBPT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img3, p, tables, scale, deterministic):
        ctx.p, ctx.tables, ctx.scale, ctx.det = p, tables, scale, deterministic
        return _fan_bpt(img3, p, tables, scale, deterministic)

    @staticmethod
    def backward(ctx, g):
        return _FanBP.apply(g.contiguous(), ctx.p, ctx.tables, ctx.scale, ctx.det), None, None, None, None


def _fan_param(img_shape, img_space, proj_shape, channel_space, img_offset, channel_offset,
               orbit_start, orbit, dso, dsd, fov):
    return FanBackprojection2DParam(
        int(proj_shape[1]), int(proj_shape[0]), float(channel_space), float(orbit),
        float(channel_offset), float(orbit_start), int(img_shape[1]), int(img_shape[0]),
        float(img_space[1]), float(img_space[0]), float(img_offset[1]), float(img_offset[0]),
        float(dso), float(dsd), float(fov))


def _check_method(method, allowed, op):
    if method not in allowed:
        raise ValueError(f"{op}: unrecognised method {method!r}; available: {', '.join(allowed)}")


def fan_backprojection_2d(proj, img_shape, img_space, proj_shape, channel_space,
                          img_offset=(0, 0), channel_offset=0, orbit_start=0,
                          orbit=math.pi / 180, dso=1000, dsd=1500, fov=-1, method="pixdriven",
                          deterministic=False):
    """Equiangular fan-beam pixel-driven BP (python/tf_ctrecon.py:221-244;
    src/bp_fan_2d.cc:46-91).  proj [N, na, ns, 1] -> img [N, ny, nx, 1]."""
    _check_method(method, ("pixdriven",), "fan_backprojection_2d")
    p = _fan_param(img_shape, img_space, proj_shape, channel_space, img_offset, channel_offset,
                   orbit_start, orbit, dso, dsd, fov)
    x, restore = _to_batch(proj, p.na, p.ns, "proj")
    tables = device_tables("fan", p, x.device)
    return restore(_FanBP.apply(x, p, tables, p.scale, deterministic))


def fan_backprojection_2d_grad(grad_in, img_shape, img_space, proj_shape, channel_space,
                               img_offset=(0, 0), channel_offset=0, orbit_start=0,
                               orbit=math.pi / 180, dso=1000, dsd=1500, fov=-1,
                               method="pixdriven", deterministic=False):
    """Transpose of fan_backprojection_2d (python/tf_ctrecon.py:246-269;
    src/bp_fan_2d.cc:124-170).  img [N, ny, nx, 1] -> proj [N, na, ns, 1]."""
    _check_method(method, ("pixdriven",), "fan_backprojection_2d_grad")
    p = _fan_param(img_shape, img_space, proj_shape, channel_space, img_offset, channel_offset,
                   orbit_start, orbit, dso, dsd, fov)
    x, restore = _to_batch(grad_in, p.ny, p.nx, "grad_in")
    tables = device_tables("fan", p, x.device)
    return restore(_FanBPT.apply(x, p, tables, p.scale, deterministic))


def _fan_angle_param(angles, img_shape, img_space, proj_shape, channel_space, img_offset,
                     channel_offset, dso, dsd, fov):
    p = FanBackprojection2DAngleParam(
        int(proj_shape[1]), int(proj_shape[0]), float(channel_space), float(channel_offset),
        int(img_shape[1]), int(img_shape[0]), float(img_space[1]), float(img_space[0]),
        float(img_offset[1]), float(img_offset[0]), float(dso), float(dsd), float(fov))
    ang = torch.as_tensor(angles, dtype=torch.float64).cpu().numpy()
    return p, ang


def fan_backprojection_2d_angle(proj, angles, img_shape, img_space, proj_shape, channel_space,
                                img_offset=(0, 0), channel_offset=0, dso=1000, dsd=1500, fov=-1,
                                deterministic=False):
    """Fan-beam BP at caller-given source angles, unscaled
    (src/bp_fan_2d_angle.cc:47-91; numpy_ext/np_ctrecon.py:66-99 multiplies
    the result by orbit itself)."""
    p, ang = _fan_angle_param(angles, img_shape, img_space, proj_shape, channel_space,
                              img_offset, channel_offset, dso, dsd, fov)
    x, restore = _to_batch(proj, p.na, p.ns, "proj")
    tables = device_tables("fan_angle", p, x.device, angles=ang)
    return restore(_FanBP.apply(x, p, tables, 1.0, deterministic))


def fan_backprojection_2d_angle_grad(grad_in, angles, img_shape, img_space, proj_shape,
                                     channel_space, img_offset=(0, 0), channel_offset=0,
                                     dso=1000, dsd=1500, fov=-1, deterministic=False):
    """Exact transpose of fan_backprojection_2d_angle (the reference's own,
    src/bp_fan_2d_angle.cc:124-165, divides by the weights and is not one)."""
    p, ang = _fan_angle_param(angles, img_shape, img_space, proj_shape, channel_space,
                              img_offset, channel_offset, dso, dsd, fov)
    x, restore = _to_batch(grad_in, p.ny, p.nx, "grad_in")
    tables = device_tables("fan_angle", p, x.device, angles=ang)
    return restore(_FanBPT.apply(x, p, tables, 1.0, deterministic))


# -------------------------------------------------- parallel-beam kernels
def _par_param(img_shape, img_space, proj_shape, channel_space, img_offset, channel_offset,
               orbit_start, orbit, fov, cls=ParallelBackprojection2DParam):
    return cls(int(proj_shape[1]), int(proj_shape[0]), float(channel_space), float(orbit),
               float(channel_offset), float(orbit_start), int(img_shape[1]), int(img_shape[0]),
               float(img_space[1]), float(img_space[0]), float(img_offset[1]),
               float(img_offset[0]), float(fov))


class _ParBP(torch.autograd.Function):
    @staticmethod
    def forward(ctx, proj3, p, tables, deterministic):
        ctx.p, ctx.tables, ctx.det = p, tables, deterministic
        n = proj3.shape[0]
        out = torch.empty((n, p.ny, p.nx), dtype=proj3.dtype, device=proj3.device)
        _call_batched("ctr_par_bp", n, C.byref(p.ctr()), _dtype_code(proj3), _ptr(proj3),
                      _ptr(out), n, _ptr(tables[0]), _ptr(tables[1]), _stream(proj3))
        return out

    @staticmethod
    def backward(ctx, g):
        return _ParBPT.apply(g.contiguous(), ctx.p, ctx.tables, ctx.det), None, None, None


class _ParBPT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img3, p, tables, deterministic):
        ctx.p, ctx.tables, ctx.det = p, tables, deterministic
        n = img3.shape[0]
        out = torch.empty((n, p.na, p.ns), dtype=img3.dtype, device=img3.device)
        _call_batched("ctr_par_bp_t", n, C.byref(p.ctr()), _dtype_code(img3), _ptr(img3),
                      _ptr(out), n, _ptr(tables[0]), _ptr(tables[1]), int(bool(deterministic)),
                      _stream(img3))
        return out

    @staticmethod
    def backward(ctx, g):
        return _ParBP.apply(g.contiguous(), ctx.p, ctx.tables, ctx.det), None, None, None


def parallel_backprojection_2d(proj, img_shape, img_space, proj_shape, channel_space,
                               img_offset=(0, 0), channel_offset=0, orbit_start=0,
                               orbit=math.pi / 180, fov=-1, method="pixdriven",
                               deterministic=False):
    """Parallel-beam pixel-driven BP (python/tf_ctrecon.py:149-170;
    src/bp_par_2d.cc:51-90).  proj [N, na, ns, 1] -> img [N, ny, nx, 1]."""
    _check_method(method, ("pixdriven",), "parallel_backprojection_2d")
    p = _par_param(img_shape, img_space, proj_shape, channel_space, img_offset, channel_offset,
                   orbit_start, orbit, fov)
    x, restore = _to_batch(proj, p.na, p.ns, "proj")
    tables = device_tables("par", p, x.device)
    return restore(_ParBP.apply(x, p, tables, deterministic))


def parallel_backprojection_2d_grad(grad_in, img_shape, img_space, proj_shape, channel_space,
                                    img_offset=(0, 0), channel_offset=0, orbit_start=0,
                                    orbit=math.pi / 180, fov=-1, method="pixdriven",
                                    deterministic=False):
    """Exact transpose of parallel_backprojection_2d (the reference's,
    src/bp_par_2d.cc:147-199, reads uninitialised pointers)."""
    _check_method(method, ("pixdriven",), "parallel_backprojection_2d_grad")
    p = _par_param(img_shape, img_space, proj_shape, channel_space, img_offset, channel_offset,
                   orbit_start, orbit, fov)
    x, restore = _to_batch(grad_in, p.ny, p.nx, "grad_in")
    tables = device_tables("par", p, x.device)
    return restore(_ParBPT.apply(x, p, tables, deterministic))


def parallel_single_view_bp_2d(proj, img_shape, img_space, proj_shape, channel_space,
                               img_offset=(0, 0), channel_offset=0, orbit_start=0,
                               orbit=math.pi / 180, fov=-1, method="pixdriven"):
    """Single-view parallel BP (python/tf_ctrecon.py:198-219;
    src/bp_par_2d_sv.cc:54-92): every pixel keeps each view's interpolated
    value, no sum and no scale.  proj [N, na, ns, 1] or [N, na, ns] -> img
    [N, ny, nx, na]; [na, ns] -> [ny, nx, na].  Bit-identical to the
    reference's CPU path.  The reference defines no gradient for it, and
    neither does this op (the output does not require grad)."""
    _check_method(method, ("pixdriven",), "parallel_single_view_bp_2d")
    p = _par_param(img_shape, img_space, proj_shape, channel_space, img_offset, channel_offset,
                   orbit_start, orbit, fov)
    x, _ = _to_batch(proj, p.na, p.ns, "proj")
    tables = device_tables("par", p, x.device)
    n = x.shape[0]
    out = torch.empty((n, p.ny, p.nx, p.na), dtype=x.dtype, device=x.device)
    _call_batched("ctr_par_sv_bp", n, C.byref(p.ctr()), _dtype_code(x), _ptr(x), _ptr(out), n,
                  _ptr(tables[0]), _ptr(tables[1]), _stream(x))
    return out[0] if proj.dim() == 2 else out


class _FP(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img3, p, t_fp, t_grad):
        ctx.p, ctx.t_fp, ctx.t_grad = p, t_fp, t_grad
        n = img3.shape[0]
        out = torch.empty((n, p.na, p.ns), dtype=img3.dtype, device=img3.device)
        sc, bo, usex = t_fp
        _call_batched("ctr_fp", n, C.byref(p.ctr()), _dtype_code(img3), _ptr(img3),
                      _ptr(out), n, _ptr(sc), _ptr(bo), _ptr(usex), _stream(img3))
        return out

    @staticmethod
    def backward(ctx, g):
        return _FPT.apply(g.contiguous(), ctx.p, ctx.t_fp, ctx.t_grad), None, None, None


class _FPT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, proj3, p, t_fp, t_grad):
        ctx.p, ctx.t_fp, ctx.t_grad = p, t_fp, t_grad
        n = proj3.shape[0]
        out = torch.empty((n, p.ny, p.nx), dtype=proj3.dtype, device=proj3.device)
        w, pos, usex = t_grad
        _call_batched("ctr_fp_t", n, C.byref(p.ctr()), _dtype_code(proj3), _ptr(proj3),
                      _ptr(out), n, _ptr(w), _ptr(pos), _ptr(usex), max(p.nx, p.ny),
                      _stream(proj3))
        return out

    @staticmethod
    def backward(ctx, g):
        return _FP.apply(g.contiguous(), ctx.p, ctx.t_fp, ctx.t_grad), None, None, None


class _FPDD(torch.autograd.Function):
    """Distance-driven FP; backward: the reference's distance-driven gradient
    (its TF op registers that one for method "disdriven")."""

    @staticmethod
    def forward(ctx, img3, p, t_fp, t_grad):
        ctx.p, ctx.t_fp, ctx.t_grad = p, t_fp, t_grad
        n = img3.shape[0]
        out = torch.empty((n, p.na, p.ns), dtype=img3.dtype, device=img3.device)
        sc, bo, usex = t_fp
        _call_batched("ctr_fp_dd", n, C.byref(p.ctr()), _dtype_code(img3), _ptr(img3),
                      _ptr(out), n, _ptr(sc), _ptr(bo), _ptr(usex), _stream(img3))
        return out

    @staticmethod
    def backward(ctx, g):
        return _FPDDGrad.apply(g.contiguous(), ctx.p, ctx.t_fp, ctx.t_grad), None, None, None


class _FPDDGrad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, proj3, p, t_fp, t_grad):
        ctx.p, ctx.t_fp, ctx.t_grad = p, t_fp, t_grad
        n = proj3.shape[0]
        out = torch.empty((n, p.ny, p.nx), dtype=proj3.dtype, device=proj3.device)
        w, pos, usex = t_grad
        _call_batched("ctr_fp_dd_grad", n, C.byref(p.ctr()), _dtype_code(proj3), _ptr(proj3),
                      _ptr(out), n, _ptr(w), _ptr(pos), _ptr(usex), max(p.nx, p.ny),
                      _stream(proj3))
        return out

    @staticmethod
    def backward(ctx, g):
        # the reference defines no gradient of its gradient op; the
        # distance-driven FP is the pair's other half
        return _FPDD.apply(g.contiguous(), ctx.p, ctx.t_fp, ctx.t_grad), None, None, None


def parallel_projection_2d(img, img_shape, img_space, proj_shape, channel_space,
                           img_offset=(0, 0), channel_offset=0, orbit_start=0,
                           orbit=math.pi / 180, fov=-1, method="raydriven"):
    """Parallel FP (python/tf_ctrecon.py:173-195).  img [N, ny, nx, 1] ->
    proj [N, na, ns, 1].  method "raydriven" (src/fp_par_2d.cc:210-280;
    backward: the gather-form transpose, src/fp_par_2d.cc:332-391),
    "raycasting" (src/fp_par_2d.cc:18-144: rays from their entry points on
    the FOV circle of radius fov, step 0.125, bilinear samples; bit-identical
    to the reference; like the reference it has no gradient) or "disdriven"
    (src/fp_par_2d.cc:396-543: channel footprints averaging the image per
    row or column; backward: the reference's distance-driven gradient,
    src/fp_par_2d.cc:548-655, which approximates the transpose; both
    bit-identical to the reference)."""
    _check_method(method, ("raydriven", "raycasting", "disdriven"), "parallel_projection_2d")
    p = _par_param(img_shape, img_space, proj_shape, channel_space, img_offset, channel_offset,
                   orbit_start, orbit, fov, cls=ParallelProjection2DParam)
    x, restore = _to_batch(img, p.ny, p.nx, "img")
    if method == "raycasting":
        sc, begins, nsteps = device_tables("fp_rc", p, x.device)
        n = x.shape[0]
        out = torch.empty((n, p.na, p.ns), dtype=x.dtype, device=x.device)
        _call_batched("ctr_fp_rc", n, C.byref(p.ctr()), _dtype_code(x), _ptr(x), _ptr(out), n,
                      RC_STEP, _ptr(sc), _ptr(begins), _ptr(nsteps), _stream(x))
        return restore(out)
    if method == "disdriven":
        return restore(_FPDD.apply(x, p, device_tables("fp_dd", p, x.device),
                                   device_tables("fp_grad", p, x.device)))
    return restore(_FP.apply(x, p, device_tables("fp", p, x.device),
                             device_tables("fp_grad", p, x.device)))


def parallel_projection_2d_grad(grad_in, img_shape, img_space, proj_shape, channel_space,
                                img_offset=(0, 0), channel_offset=0, orbit_start=0,
                                orbit=math.pi / 180, fov=-1, method="raydriven"):
    """Transpose of parallel_projection_2d (tensorflow/fp_par_2d_ops.cu:245-374)."""
    if method == "raycasting":  # tensorflow/fp_par_2d_ops.cu:287-291
        raise RuntimeError("Gradient is not yet implemented for raycasting forward projector.")
    _check_method(method, ("raydriven", "disdriven"), "parallel_projection_2d_grad")
    p = _par_param(img_shape, img_space, proj_shape, channel_space, img_offset, channel_offset,
                   orbit_start, orbit, fov, cls=ParallelProjection2DParam)
    x, restore = _to_batch(grad_in, p.na, p.ns, "grad_in")
    if method == "disdriven":
        return restore(_FPDDGrad.apply(x, p, device_tables("fp_dd", p, x.device),
                                       device_tables("fp_grad", p, x.device)))
    return restore(_FPT.apply(x, p, device_tables("fp", p, x.device),
                              device_tables("fp_grad", p, x.device)))


# ------------------------------------------------- ramp filter / weighting
class _Ramp(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x2, p, taps, impl):
        ctx.p, ctx.taps, ctx.impl = p, taps, impl
        return _ramp_launch(x2, p, taps, None, impl)

    @staticmethod
    def backward(ctx, g):
        # the taps are symmetric: the transpose is the same filter
        return _Ramp.apply(g.contiguous(), ctx.p, ctx.taps, ctx.impl), None, None, None


class _FanFilter(torch.autograd.Function):
    """Fan weighting then ramp filter in one kernel; backward: filter then
    weighting (the transpose of the composition)."""

    @staticmethod
    def forward(ctx, x2, p, taps, wp, impl):
        ctx.p, ctx.taps, ctx.wp, ctx.impl = p, taps, wp, impl
        return _ramp_launch(x2, p, taps, wp, impl)

    @staticmethod
    def backward(ctx, g):
        y = _ramp_launch(g.contiguous(), ctx.p, ctx.taps, None, ctx.impl)
        return _FanW.apply(y, ctx.wp), None, None, None, None


# Ramp-filter implementation for float32 tensors:
#   "mfma"   Toeplitz GEMM on the matrix cores, FP32 MFMA (default)
#   "bf16x3" Toeplitz GEMM on BF16 MFMA with hi/lo split operands
#   "direct" LDS-staged direct convolution on the vector ALUs
# float64 tensors always use the direct FP64 kernel.
RAMP_IMPLS = ("mfma", "bf16x3", "direct")
# BF16x3 MFMA is the fastest of the three at both the 736- and the
# 1472-channel shapes and as accurate as the FP32 MFMA filter (both carry the
# centre and near taps in FP64; see csrc/toeplitz.hip).
default_ramp_impl = "bf16x3"


def _ramp_launch(x2, p, taps, wp, impl=None):
    impl = impl or default_ramp_impl
    if impl not in RAMP_IMPLS:
        raise ValueError(f"ramp filter impl must be one of {RAMP_IMPLS}")
    out = torch.empty_like(x2)
    g = p.ctr()
    wtype = 0
    if wp is not None:
        g.dso, g.dsd, g.offset_s, wtype = wp.dso, wp.dsd, wp.offset, wp.type
    rows = x2.shape[0]
    if x2.dtype == torch.float32 and impl != "direct":
        _call_batched("ctr_ramp_toeplitz", rows, C.byref(g), _lib.CTR_F32, _ptr(x2), _ptr(out),
                      rows, _ptr(taps[0]), wtype, int(impl == "bf16x3"), _stream(x2))
    else:
        _call_batched("ctr_ramp", rows, C.byref(g), _dtype_code(x2), _ptr(x2), _ptr(out), rows,
                      _ptr(taps[0]), wtype, _stream(x2))
    return out


def _rows(x: torch.Tensor, ns: int, name: str):
    _require_gpu(x, name)
    if x.dim() in (4, 5) and x.shape[-1] == 1:
        if x.shape[-2] != ns:
            raise ValueError(f"{name}: width must equal n_channel={ns}, got {tuple(x.shape)}")
    elif x.shape[-1] != ns:
        raise ValueError(f"{name}: last dimension must equal n_channel={ns}")
    return x.reshape(-1, ns).contiguous()


def ramp_filter(proj, n_channel, n_row, channel_space, dsd=-1, type="par", window="None",
                impl=None):
    """Ramp filter along the channels of every row (python/tf_ctrecon.py:112-127;
    src/filter.cc:67-91).  4-D [N, nrow, ns, 1] or 5-D [N, D, H, ns, 1]
    (D*H = nrow, tensorflow/ramp_filter_ops.cu:130-143)."""
    if type not in _FILTER_TYPES:
        raise ValueError("ramp_filter: type must be 'par', 'fan' or 'flat'")
    p = FilterParam(int(n_channel), int(n_row), float(channel_space), float(dsd),
                    _FILTER_TYPES[type], 0)
    x2 = _rows(proj, p.ns, "proj")
    if x2.shape[0] % p.nrow:
        raise ValueError(f"ramp_filter: {x2.shape[0]} rows is not a multiple of n_row={p.nrow}")
    taps = device_tables("ramp", p, x2.device)
    return _Ramp.apply(x2, p, taps, impl).reshape(proj.shape)


def ramp_filter_grad(grad_in, n_channel, n_row, channel_space, dsd=-1, type="par",
                     window="None", impl=None):
    """Transpose of ramp_filter (python/tf_ctrecon.py:129-146; src/filter.cc:96-124);
    the taps are symmetric, so it is the same filter."""
    return ramp_filter(grad_in, n_channel, n_row, channel_space, dsd, type, window, impl)


class _FanW(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x2, p):
        ctx.p = p
        out = torch.empty_like(x2)
        _call_batched("ctr_fan_weighting", x2.shape[0], C.byref(p.ctr()), _dtype_code(x2),
                      _ptr(x2), _ptr(out), x2.shape[0], _stream(x2))
        return out

    @staticmethod
    def backward(ctx, g):
        return _FanW.apply(g.contiguous(), ctx.p), None


def _fanw_param(proj_shape, channel_space, channel_offset, dso, dsd, tp):
    if tp not in _WEIGHT_TYPES:
        raise ValueError("fan_weighting: tp must be 'fan' or 'flat'")
    return FanWeightingParam(int(proj_shape[1]), int(proj_shape[0]), float(channel_space),
                             float(channel_offset), float(dso), float(dsd), _WEIGHT_TYPES[tp])


def fan_weighting(proj, proj_shape, channel_space, channel_offset=0, dso=1000, dsd=1500,
                  tp="fan"):
    """Cosine pre-weighting dso*|cos(gamma)|/dsd per channel
    (python/tf_ctrecon.py:271-284; src/fan_weighting.cc:19-35)."""
    p = _fanw_param(proj_shape, channel_space, channel_offset, dso, dsd, tp)
    x2 = _rows(proj, p.ns, "proj")
    return _FanW.apply(x2, p).reshape(proj.shape)


def fan_weighting_grad(grad_in, proj_shape, channel_space, channel_offset=0, dso=1000,
                       dsd=1500, tp="fan"):
    """Transpose of fan_weighting: the same diagonal (src/fan_weighting.cc:39-56)."""
    return fan_weighting(grad_in, proj_shape, channel_space, channel_offset, dso, dsd, tp)


def fan_filter(proj, proj_shape, channel_space, channel_offset=0, dso=1000, dsd=1500,
               tp="fan", impl=None):
    """fan_weighting followed by the matching ramp filter (numpy_ext/np_ctrecon.py:91-93),
    fused into one kernel: the weighting is applied while the filter stages
    its rows.  Differentiable (backward: filter, then weighting)."""
    wp = _fanw_param(proj_shape, channel_space, channel_offset, dso, dsd, tp)
    p = FilterParam(wp.ns, wp.nrow, wp.ds, float(dsd), wp.type, 0)
    x2 = _rows(proj, p.ns, "proj")
    taps = device_tables("ramp", p, x2.device)
    return _FanFilter.apply(x2, p, taps, wp, impl).reshape(proj.shape)


# ------------------------------------------------------- helical rebinning
def helical_interpolation(proj, param: HelicalInterpolationParam, mode: int = 360,
                          check: bool = True):
    """HelicalInterpolation360 (mode 360) or HelicalInterpolation180 (mode
    180) of include/helical_interpolation.h:86-118 (src/helical_interpolation.cc:
    81-160, 162-269): a multi-row helical scan ``proj`` of shape
    [na, nt, ns] or [B, na, nt, ns] (views, detector rows, channels) to nz
    fan-beam sinograms of one rotation each, [nz, view_per_rot, ns] or
    [B, nz, view_per_rot, ns], bit-identical to the reference per element.

    A slice without enough views raises ValueError before any work, as the
    reference throws "View not enough for this slice!".  Where a sample sees
    no detector row in either window the reference throws "No row data
    available!": with ``check`` the call synchronises and raises
    RuntimeError; without it the sample is NaN and nothing synchronises.
    Not differentiable: the reference defines no gradient for it."""
    if mode not in (360, 180):
        raise ValueError("mode must be 360 or 180")
    _require_gpu(proj, "proj")
    p = param
    shape = (p.na, p.nt, p.ns)
    if tuple(proj.shape[-3:]) != shape or proj.dim() not in (3, 4):
        raise ValueError(f"proj must be [na, nt, ns] = {list(shape)} or [B, na, nt, ns], "
                         f"got {list(proj.shape)}")
    x = proj.contiguous()
    nb = 1 if x.dim() == 3 else x.shape[0]
    try:
        (tables,) = device_tables(f"helical_{mode}", p, x.device)
    except _lib.CtrError as e:
        raise ValueError(str(e)) from None
    out = torch.empty(x.shape[:-3] + (p.nz, p.view_per_rot, p.ns), dtype=x.dtype,
                      device=x.device)
    if nb == 0:
        return out
    status = torch.zeros(1, dtype=torch.int32, device=x.device) if check else None
    _lib.call("ctr_helical", C.byref(p.ctr()), int(mode), _dtype_code(x), _ptr(x), _ptr(out),
              int(nb), _ptr(tables), None if status is None else _ptr(status), _stream(x))
    if check:
        code = int(status.item())
        if code & 1:
            raise RuntimeError("helical_interpolation: No row data available! (a sample sees "
                               "no detector row in either window)")
        if code & 2:
            raise RuntimeError("helical_interpolation: a 180LI conjugate view lies outside "
                               "the scan")
    return out
