"""Gradient-based least-squares reconstruction on the differentiable
projector: min_x 0.5 ||FP(x) - y||^2 with FP the ray-driven parallel
projector (src/fp_par_2d.cc:210-280) and its gradient FP^T(FP x - y) taken
through autograd (the backward is the hand-written gather transpose).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch

from . import functional as F


@dataclass(frozen=True)
class ParallelScan:
    """Parallel-beam scan: n x n image of pixel dx, na views over 180 deg,
    ns channels of pitch ds, centred."""
    n: int = 512
    na: int = 720
    ns: int = 729
    dx: float = 1.0
    ds: float = 1.0

    def kwargs(self):
        return dict(img_shape=[self.n, self.n], img_space=[self.dx, self.dx],
                    proj_shape=[self.na, self.ns], channel_space=self.ds,
                    orbit=math.pi / self.na)


def projector(scan: ParallelScan):
    kw = scan.kwargs()

    def fp(x):
        return F.parallel_projection_2d(x, **kw)

    def fpt(y):
        return F.parallel_projection_2d_grad(y, **kw)

    return fp, fpt


def operator_norm_sq(scan: ParallelScan, device, iters: int = 20, seed: int = 0) -> float:
    """||FP||^2 by power iteration on FP^T FP (one slice)."""
    fp, fpt = projector(scan)
    g = torch.Generator(device=device).manual_seed(seed)
    x = torch.randn((1, scan.n, scan.n), generator=g, device=device)
    lam = 0.0
    for _ in range(iters):
        x = x / x.norm()
        z = fpt(fp(x))
        lam = float((x * z).sum())
        x = z
    return lam


def loss_and_grad(x: torch.Tensor, y: torch.Tensor, fp):
    """0.5 ||fp(x) - y||^2 and its gradient through autograd."""
    xr = x.detach().requires_grad_(True)
    r = fp(xr) - y
    loss = 0.5 * (r * r).sum()
    (g,) = torch.autograd.grad(loss, xr)
    return loss.detach(), g


def gradient_descent(x0: torch.Tensor, y: torch.Tensor, scan: ParallelScan, steps: int,
                     alpha: float):
    """steps of x <- x - alpha * FP^T (FP x - y); returns x and the losses
    (a tensor on the device, read back once at the end)."""
    fp, _ = projector(scan)
    x = x0.clone()
    losses = torch.empty(steps, dtype=torch.float64, device=x.device)
    for k in range(steps):
        loss, g = loss_and_grad(x, y, fp)
        losses[k] = loss
        x = x - alpha * g
    return x, losses
