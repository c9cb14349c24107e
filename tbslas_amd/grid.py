"""Leaves with regular sample grids and tricubic evaluation (the large
cubic-interpolation variant of the semi-Lagrangian step) over libtbx.

A leaf holds an N^3 grid of samples at corner + 2^-depth (a, b, c) / (N - 1)
(x fastest, samples [L][dof][N^3]); values between samples come from the
reference's fast_interp (tree_functor.h:89-153) in the leaf's unit
coordinates.  Leaves are given as for Tree (Morton order, anchor + depth).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BC_FREESPACE, WRITE_BACK_WRAP, _GridDesc, check, lib, tfac
from .tree import _check_dep, _init_device, _mode, _ptr, _require_dev, _stream


class Grid:
    """Device handle of one rank's grid-leaf tree (tbx_grid)."""

    def __init__(self, anchor, depth, n: int, dof: int, samples=None):
        self.anchor = np.ascontiguousarray(anchor, dtype=np.uint32).reshape(-1, 3)
        self.depth = np.ascontiguousarray(depth, dtype=np.uint8).reshape(-1)
        self.n, self.dof = int(n), int(dof)
        self.L = len(self.depth)
        self.P = self.n ** 3
        cf = None
        if samples is not None:
            cf = np.ascontiguousarray(samples, dtype=np.float64)
            if cf.size != self.L * self.dof * self.P:
                raise ValueError("samples must hold L*dof*n^3 values")
        _init_device()
        d = _GridDesc(self.L, _ptr(self.anchor), _ptr(self.depth), self.n, self.dof, _ptr(cf))
        h = C.c_void_p()
        check(lib().tbx_grid_create(C.byref(d), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().tbx_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_samples(self, samples):
        import torch

        if isinstance(samples, torch.Tensor):
            _require_dev(samples, torch.float64, "samples")
            if samples.numel() != self.L * self.dof * self.P:
                raise ValueError("sample tensor has the wrong size")
            check(lib().tbx_grid_set_samples(self._h, _ptr(samples), 1, _stream()))
        else:
            cf = np.ascontiguousarray(samples, dtype=np.float64)
            if cf.size != self.L * self.dof * self.P:
                raise ValueError("sample array has the wrong size")
            check(lib().tbx_grid_set_samples(self._h, _ptr(cf), 0, _stream()))

    def get_samples(self) -> np.ndarray:
        out = np.empty((self.L, self.dof, self.P))
        check(lib().tbx_grid_get_samples(self._h, _ptr(out), 0, _stream()))
        return out

    def eval(self, pts, bc: int = BC_FREESPACE, mode: str = "fp64", write_back: bool = False):
        import torch

        _require_dev(pts, torch.float64, "pts", (3,))
        out = torch.empty((pts.shape[0], self.dof), dtype=torch.float64, device=pts.device)
        flags = _mode(mode) | (WRITE_BACK_WRAP if write_back else 0)
        check(lib().tbx_grid_eval(self._h, _ptr(pts), pts.shape[0], bc, flags, _ptr(out),
                                  _stream()))
        return out


def grid_points(anchor, depth, n: int) -> np.ndarray:
    """Sample points [L*n^3][3] of the leaves (host, setup)."""
    anchor = np.asarray(anchor, dtype=np.float64).reshape(-1, 3)
    s = 2.0 ** np.asarray(depth, dtype=np.float64)
    node = np.arange(n) / (n - 1)
    a, b, c = np.meshgrid(node, node, node, indexing="ij")
    loc = np.stack([c.reshape(-1), b.reshape(-1), a.reshape(-1)], axis=1)  # x fastest
    return ((anchor / s[:, None])[:, None, :] + loc[None, :, :] / s[:, None, None]).reshape(-1, 3)


def sample_function(fn, anchor, depth, n: int, dof: int, batch: int = 4096) -> np.ndarray:
    """Samples [L, dof, n^3] of fn (pts [m,3] -> [m,dof]) on every leaf."""
    L = len(depth)
    out = np.empty((L, dof, n ** 3))
    for a in range(0, L, batch):
        b = min(L, a + batch)
        v = np.asarray(fn(grid_points(anchor[a:b], depth[a:b], n)), dtype=np.float64)
        out[a:b] = v.reshape(b - a, n ** 3, dof).transpose(0, 2, 1)
    return out


def grid_semilag_insitu(vel: Grid, conc: Grid, timestep: int, dt: float, num_rk_step: int = 1,
                        bc: int = BC_FREESPACE, mode: str = "fp64", g=None, dep=None):
    """One fused in-situ step of grid leaves: samples <- values at the
    departure points of the sample points."""
    _check_dep(dep, conc.L * conc.P)
    cb = tfac(g)
    check(lib().tbx_grid_semilag_insitu(vel.handle, cb, None, conc.handle, timestep, dt,
                                        num_rk_step, bc, _mode(mode), _ptr(dep), _stream()))
