"""Ragged batches of fan-beam backprojections with arbitrary view angles.

``RaggedFanBackprojection2DAngle`` runs many ``bp_fan_2d_angle`` problems
(src/bp_fan_2d_angle.cc:47-91) -- each with its own image size, view count
and angles, detector and source distance -- in one persistent kernel
launch, replayed from a HIP graph.  The sinograms and images of the batch
live in two packed device buffers; ``projs`` and ``imgs`` are per-problem
views into them.

The pixel tile and the views per LDS round are chosen by the library's
autotuner (``ctr_ragged_fan_bp_autotune``: every tile {16x16, 8x8, 32x8} x
chunk {8, 16, 32, 64}, timed on the device) and remembered in a JSON cache
keyed by the batch's geometry and the device name.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import math
import os
from typing import Sequence

import numpy as np

from . import _lib
from .geometry import FanBackprojection2DAngleParam, fan_angle_tables

TILES = {0: "16x16", 1: "8x8", 2: "32x8"}
CHUNKS = (8, 16, 32, 64)


def _cache_path() -> str:
    p = os.environ.get("CTR_AUTOTUNE_CACHE")
    if p:
        return p
    base = os.environ.get("XDG_CACHE_HOME", os.path.join(os.path.expanduser("~"), ".cache"))
    return os.path.join(base, "ct_recon_extensions_amd", "ragged_autotune.json")


def _signature(params, angles, device_name: str) -> str:
    h = hashlib.sha1(device_name.encode())
    for p, a in zip(params, angles):
        h.update(repr(tuple(round(float(v), 9) if isinstance(v, float) else v
                            for v in (p.ns, p.na, p.ds, p.offset_s, p.nx, p.ny, p.dx, p.dy,
                                      p.offset_x, p.offset_y, p.dso, p.dsd))).encode())
        h.update(np.ascontiguousarray(a, np.float64).tobytes())
    return h.hexdigest()


def _load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(path: str, key: str, entry: dict) -> None:
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = _load_cache(path)
        data[key] = entry
        tmp = path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only costs a re-tune next time


class RaggedFanBackprojection2DAngle:
    """One plan for a fixed ragged batch.

    params: FanBackprojection2DAngleParam per problem; angles: the na source
    angles of each problem (radians); scale: factor applied to every image
    (1 is the reference's bp_fan_2d_angle, which does not scale).
    tile / vchunk: fix the configuration (tile 0 16x16, 1 8x8, 2 32x8);
    left as None they come from the autotune cache, or from a fresh
    autotune run when ``autotune`` is true (else 16x16 and 16).
    """

    def __init__(self, params: Sequence[FanBackprojection2DAngleParam], angles, device="cuda",
                 tile: int | None = None, vchunk: int | None = None, autotune: bool = True,
                 scale: float = 1.0, cache_path: str | None = None, tune_reps: int = 3):
        import torch

        if len(params) == 0 or len(params) != len(angles):
            raise ValueError("need one angle array per problem and at least one problem")
        self._lib = _lib.load()
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("the ragged backprojection runs on a GPU only")
        self.params = list(params)
        self.angles = [np.ascontiguousarray(a, np.float64) for a in angles]
        sizes_p = [p.na * p.ns for p in self.params]
        sizes_i = [p.nx * p.ny for p in self.params]
        self.proj_flat = torch.zeros(sum(sizes_p), dtype=torch.float32, device=self.device)
        self.img_flat = torch.zeros(sum(sizes_i), dtype=torch.float32, device=self.device)
        self.projs, self.imgs, self._tables = [], [], []
        op, oi = 0, 0
        probs = (_lib.CtrRaggedProblem * len(self.params))()
        for i, (p, a) in enumerate(zip(self.params, self.angles)):
            tab = tuple(torch.from_numpy(t).to(self.device) for t in fan_angle_tables(p, a))
            self._tables.append(tab)
            pr = self.proj_flat[op:op + sizes_p[i]].view(p.na, p.ns)
            im = self.img_flat[oi:oi + sizes_i[i]].view(p.ny, p.nx)
            self.projs.append(pr)
            self.imgs.append(im)
            op += sizes_p[i]
            oi += sizes_i[i]
            q = probs[i]
            q.g = p.ctr()
            q.xpos, q.ypos, q.sincostbl = (t.data_ptr() for t in tab)
            q.proj, q.img = pr.data_ptr(), im.data_ptr()
            q.scale = float(scale)
        self._probs = probs
        self.pixel_views = int(sum(p.na * p.nx * p.ny for p in self.params))
        self.tune_table = None

        if tile is None or vchunk is None:
            path = cache_path or _cache_path()
            key = _signature(self.params, self.angles, torch.cuda.get_device_name(self.device))
            hit = _load_cache(path).get(key)
            if hit is not None:
                tile, vchunk, self.tune_table = hit["tile"], hit["vchunk"], hit.get("ms")
            elif autotune:
                tile, vchunk, self.tune_table = self._autotune(tune_reps)
                _store_cache(path, key, {"tile": tile, "vchunk": vchunk, "ms": self.tune_table})
            else:
                tile, vchunk = 0, 16
        self.tile, self.vchunk = int(tile), int(vchunk)
        plan = C.c_void_p()
        _lib.check(self._lib.ctr_ragged_fan_bp_plan(C.byref(self._probs), len(self.params),
                                                    self.tile, self.vchunk, C.byref(plan)),
                   "ctr_ragged_fan_bp_plan")
        self._plan = plan
        ni, gr, bl, lds = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        pv = C.c_double()
        _lib.check(self._lib.ctr_ragged_fan_bp_info(plan, C.byref(ni), C.byref(gr), C.byref(bl),
                                                    C.byref(lds), C.byref(pv)), "info")
        self.nitems, self.grid, self.block, self.lds_bytes = ni.value, gr.value, bl.value, lds.value

    def _stream(self):
        import torch
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _autotune(self, reps: int):
        bt, bv = C.c_int(), C.c_int()
        table = (C.c_float * 12)()
        _lib.check(self._lib.ctr_ragged_fan_bp_autotune(C.byref(self._probs), len(self.params),
                                                        reps, self._stream(), C.byref(bt),
                                                        C.byref(bv), table),
                   "ctr_ragged_fan_bp_autotune")
        ms = {f"{TILES[t]}/{CHUNKS[c]}": round(float(table[4 * t + c]), 5)
              for t in range(3) for c in range(4)}
        return bt.value, bv.value, ms

    def __call__(self, projs=None, graph: bool = True):
        """Backproject the batch.  projs (optional): one [na, ns] tensor per
        problem, copied into the packed buffer first.  Returns the image
        views of the packed output, overwritten by the next run."""
        if projs is not None:
            if len(projs) != len(self.projs):
                raise ValueError(f"expected {len(self.projs)} sinograms, got {len(projs)}")
            for dst, src in zip(self.projs, projs):
                dst.copy_(src.reshape(dst.shape))
        _lib.check(self._lib.ctr_ragged_fan_bp_run(self._plan, 1 if graph else 0, self._stream()),
                   "ctr_ragged_fan_bp_run")
        return self.imgs

    def balance(self) -> float:
        """One run that measures the fraction of workgroup time spent busy."""
        b = C.c_double()
        _lib.check(self._lib.ctr_ragged_fan_bp_balance(self._plan, self._stream(), C.byref(b)),
                   "ctr_ragged_fan_bp_balance")
        return b.value

    def close(self):
        plan = getattr(self, "_plan", None)
        if plan:
            self._lib.ctr_ragged_fan_bp_destroy(plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def random_fan_problems(count: int = 64, seed: int = 20261204, sizes=(128, 256, 384, 512),
                        views=(180, 1440), channels=(256, 1024), dsd: float = 1085.6,
                        ratio=(0.5, 0.6), half_fan: float = 0.44):
    """A seeded ragged batch of equiangular fan problems with irregular angles.

    Per problem: image n x n with n drawn from ``sizes``; na ~ U{views};
    angles = a jittered regular grid over [0, 2 pi) (each angle moved by
    U(-0.45, 0.45) of the mean step, from a random start), sorted; ns ~
    U{channels} at the arc pitch that spans +-half_fan; dso = U[ratio] * dsd;
    pixel size such that the image's circumscribed circle fills 98 % of the
    fan's field of view dso sin(half_fan).
    Returns (params, angles).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    params, angles = [], []
    for _ in range(count):
        n = int(rng.choice(sizes))
        na = int(rng.integers(views[0], views[1] + 1))
        ns = int(rng.integers(channels[0], channels[1] + 1))
        dso = float(rng.uniform(*ratio)) * dsd
        ds = 2.0 * half_fan * dsd / ns
        radius = 0.98 * dso * math.sin(half_fan)
        dx = 2.0 * radius / (n * math.sqrt(2.0))
        step = 2 * math.pi / na
        a = rng.uniform(0, 2 * math.pi) + step * (np.arange(na) + rng.uniform(-0.45, 0.45, na))
        a = np.sort(np.mod(a, 2 * math.pi))
        params.append(FanBackprojection2DAngleParam(ns, na, ds, 0.0, n, n, dx, dx, 0.0, 0.0,
                                                    dso, dsd))
        angles.append(a)
    return params, angles
