"""Seeded ellipse phantoms and their analytic sinograms (synthetic inputs).

Conventions are the reference's (src/bp_par_2d.cc:22-45, src/bp_fan_2d.cc:22-42,
src/bp_fan_2d.cc:66-69):

* pixel (ix, iy) sits at x = (ix - centx)*dx, y = (centy - iy)*dy with
  centx = (nx-1)/2 + offset_x, centy = (ny-1)/2 + offset_y (row 0 is +y);
* the parallel ray of view theta at detector coordinate t is
  x*cos(theta) + y*sin(theta) = t, with t = (is - cents)*ds;
* the fan source of view beta sits at (-dso*sin(beta), dso*cos(beta)); the
  ray at fan angle gamma is the parallel ray theta = beta + gamma,
  t = dso*sin(gamma); on an equiangular detector gamma = (is - cents)*ds/dsd.

An ellipse is (x0, y0, a, b, phi, rho): centre, semi-axes, rotation, density.
Its parallel projection is 2*rho*a*b*sqrt(m2 - tau^2)/m2 where
m2 = a^2 cos^2(theta-phi) + b^2 sin^2(theta-phi) and
tau = t - x0 cos(theta) - y0 sin(theta) (zero where tau^2 > m2).
"""
from __future__ import annotations

import math

import numpy as np


def random_ellipses(n: int, seed: int, radius: float = 1.0) -> np.ndarray:
    """n ellipses in a disk of the given radius, drawn from PCG64(seed):
    centres U[-0.8,0.8]^2, semi-axes U[0.05,0.5]^2, angle U[0,pi),
    density U[-0.5,1], all lengths scaled by ``radius``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    e = np.empty((n, 6))
    e[:, 0:2] = rng.uniform(-0.8, 0.8, (n, 2)) * radius
    e[:, 2:4] = rng.uniform(0.05, 0.5, (n, 2)) * radius
    e[:, 4] = rng.uniform(0.0, math.pi, n)
    e[:, 5] = rng.uniform(-0.5, 1.0, n)
    return e


def shepp_logan(radius: float = 1.0) -> np.ndarray:
    """The modified Shepp-Logan head (Toft's contrasts), scaled to ``radius``."""
    e = np.array([
        [0.0, 0.0, 0.69, 0.92, 0.0, 1.0],
        [0.0, -0.0184, 0.6624, 0.874, 0.0, -0.8],
        [0.22, 0.0, 0.11, 0.31, -18.0, -0.2],
        [-0.22, 0.0, 0.16, 0.41, 18.0, -0.2],
        [0.0, 0.35, 0.21, 0.25, 0.0, 0.1],
        [0.0, 0.1, 0.046, 0.046, 0.0, 0.1],
        [0.0, -0.1, 0.046, 0.046, 0.0, 0.1],
        [-0.08, -0.605, 0.046, 0.023, 0.0, 0.1],
        [0.0, -0.606, 0.023, 0.023, 0.0, 0.1],
        [0.06, -0.605, 0.023, 0.046, 0.0, 0.1],
    ])
    e[:, 4] = np.deg2rad(e[:, 4])
    e[:, 0:4] *= radius
    return e


def pixel_grid(nx, ny, dx, dy, offset_x=0.0, offset_y=0.0):
    centx = (nx - 1) / 2 + offset_x
    centy = (ny - 1) / 2 + offset_y
    x = (np.arange(nx) - centx) * dx
    y = (centy - np.arange(ny)) * dy
    return x, y


def rasterize(ell: np.ndarray, nx, ny, dx, dy, offset_x=0.0, offset_y=0.0,
              supersample: int = 1) -> np.ndarray:
    """Image [ny, nx] of the ellipse densities, point-sampled at pixel centres
    (or averaged over supersample^2 sub-points)."""
    x, y = pixel_grid(nx, ny, dx, dy, offset_x, offset_y)
    img = np.zeros((ny, nx))
    offs = (np.arange(supersample) + 0.5) / supersample - 0.5
    for ox in offs:
        for oy in offs:
            X, Y = np.meshgrid(x + ox * dx, y + oy * dy)
            for x0, y0, a, b, phi, rho in ell:
                c, s = math.cos(phi), math.sin(phi)
                u = (X - x0) * c + (Y - y0) * s
                v = -(X - x0) * s + (Y - y0) * c
                img += rho * ((u / a) ** 2 + (v / b) ** 2 <= 1.0)
    return img / supersample ** 2


def parallel_projection(ell: np.ndarray, theta: np.ndarray, t: np.ndarray) -> np.ndarray:
    """Analytic line integrals of the ellipse set along the rays (theta, t),
    broadcast element-wise (FP64)."""
    th = np.asarray(theta, np.float64)
    tt = np.asarray(t, np.float64)
    out = np.zeros(np.broadcast_shapes(th.shape, tt.shape))
    for x0, y0, a, b, phi, rho in ell:
        m2 = a * a * np.cos(th - phi) ** 2 + b * b * np.sin(th - phi) ** 2
        tau = tt - x0 * np.cos(th) - y0 * np.sin(th)
        q = m2 - tau * tau
        out += np.where(q > 0, 2.0 * rho * a * b * np.sqrt(np.maximum(q, 0.0)) / m2, 0.0)
    return out


def parallel_sinogram(ell, na, ns, ds, orbit, orbit_start=0.0, offset_s=0.0):
    theta = orbit_start + orbit * np.arange(na)
    t = (np.arange(ns) - ((ns - 1) / 2 + offset_s)) * ds
    return parallel_projection(ell, theta[:, None], t[None, :])


def fan_sinogram(ell, na, ns, ds, orbit, dso, dsd, orbit_start=0.0, offset_s=0.0,
                 angles=None):
    """Equiangular fan sinogram [na, ns]: p_par(beta + gamma, dso*sin(gamma))."""
    beta = (orbit_start + orbit * np.arange(na)) if angles is None else np.asarray(angles)
    gamma = (np.arange(ns) - ((ns - 1) / 2 + offset_s)) * ds / dsd
    theta = beta[:, None] + gamma[None, :]
    return parallel_projection(ell, theta, dso * np.sin(gamma)[None, :])


def poisson_noise(p: np.ndarray, i0: float, seed: int, peak: float = 4.0) -> np.ndarray:
    """Poisson counts at i0 incident photons on a sinogram scaled so that its
    largest line integral is ``peak``; returns -ln(max(counts,1)/i0) rescaled
    back to the input's units."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mu = peak / max(float(np.max(np.abs(p))), 1e-30)
    counts = rng.poisson(i0 * np.exp(-mu * p))
    return -np.log(np.maximum(counts, 1) / i0) / mu


def fan_sinogram_torch(ell, na, ns, ds, orbit, dso, dsd, device, orbit_start=0.0,
                       offset_s=0.0, dtype=None):
    """fan_sinogram evaluated with torch on ``device`` (FP64 by default), for
    workloads too large to generate on the host; same formula and
    conventions."""
    import torch
    dtype = dtype or torch.float64
    beta = orbit_start + orbit * torch.arange(na, dtype=dtype, device=device)
    gamma = (torch.arange(ns, dtype=dtype, device=device) - ((ns - 1) / 2 + offset_s)) * ds / dsd
    th = beta[:, None] + gamma[None, :]
    tt = (dso * torch.sin(gamma))[None, :]
    c, s = torch.cos(th), torch.sin(th)
    out = torch.zeros((na, ns), dtype=dtype, device=device)
    for x0, y0, a, b, phi, rho in np.asarray(ell, np.float64):
        cp, sp = math.cos(phi), math.sin(phi)
        ct = c * cp + s * sp  # cos(theta - phi)
        st = s * cp - c * sp  # sin(theta - phi)
        m2 = a * a * ct * ct + b * b * st * st
        tau = tt - x0 * c - y0 * s
        q = m2 - tau * tau
        out += torch.where(q > 0, 2.0 * rho * a * b * torch.sqrt(q.clamp_min(0.0)) / m2,
                           torch.zeros((), dtype=dtype, device=device))
    return out


def poisson_noise_torch(p, i0: float, generator, peak: float = 4.0):
    """poisson_noise with torch's Poisson sampler (``generator`` seeds it)."""
    import torch
    mu = peak / max(float(p.abs().max()), 1e-30)
    counts = torch.poisson(i0 * torch.exp(-mu * p), generator=generator)
    return -torch.log(counts.clamp_min(1.0) / i0) / mu


def shepp_logan_fan_batch(nbatch: int, seed: int, device, na: int = 1440, ns: int = 1472,
                          ds: float = 0.65, dso: float = 595.0, dsd: float = 1085.6,
                          fov_radius: float = 1024 * 0.35 / 2):
    """Seeded batch of Shepp-Logan-type fan sinograms, float32 on ``device``:
    per slice the head (scaled to fov_radius) is rotated by U[0, 2 pi) and its
    densities scaled by U[0.9, 1.1]; equiangular fan over 360 deg; Poisson
    noise at I0 = 10^U[4, 6] photons.  Returns (sinograms [nbatch, na, ns],
    the fan_filter keywords of the geometry, the I0 of every slice)."""
    import torch
    rng = np.random.Generator(np.random.PCG64(seed))
    gen = torch.Generator(device=device).manual_seed(seed)
    out = torch.empty((nbatch, na, ns), dtype=torch.float32, device=device)
    i0s = []
    for b in range(nbatch):
        ell = shepp_logan(fov_radius)
        rot = rng.uniform(0, 2 * math.pi)
        c, s = math.cos(rot), math.sin(rot)
        ell[:, 0], ell[:, 1] = c * ell[:, 0] - s * ell[:, 1], s * ell[:, 0] + c * ell[:, 1]
        ell[:, 4] += rot
        ell[:, 5] *= rng.uniform(0.9, 1.1)
        i0 = float(10 ** rng.uniform(4, 6))
        i0s.append(i0)
        p = fan_sinogram_torch(ell, na, ns, ds, 2 * math.pi / na, dso, dsd, device)
        out[b] = poisson_noise_torch(p, i0, gen).float()
    kw = dict(proj_shape=[na, ns], channel_space=ds, dso=dso, dsd=dsd, tp="fan")
    return out, kw, i0s
