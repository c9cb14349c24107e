"""CheckChebOutput (applications/include/utils.txx:118-297): the error
metric of the reference's convergence tests.

For a tree of degree q the field is sampled at the n^3 first-kind
Chebyshev nodes of every leaf, n = q/3 + 1 (utils.txx:149-153); the
volume-weighted mean error (weights s^3 / n^3, s the leaf edge) is removed
(utils.txx:158-187), and the absolute / relative L2 and L-inf errors of the
remainder are reported (utils.txx:240-293).  The evaluation of the tree is
a caller-supplied function, so the same metric serves device trees and
the host oracle.
"""
from __future__ import annotations

import numpy as np


def first_kind_nodes(n: int) -> np.ndarray:
    """n first-kind Chebyshev nodes on [0, 1] (the reference's cheb_nodes)."""
    return -np.cos((np.arange(n) + 0.5) * np.pi / n) * 0.5 + 0.5


def error_points(anchor, depth, q: int):
    """The metric's sample points [L * n^3, 3] and per-point volume weights
    s^3 / n^3."""
    n = q // 3 + 1
    x = first_kind_nodes(n)
    a, b, c = np.meshgrid(x, x, x, indexing="ij")
    loc = np.stack([c.reshape(-1), b.reshape(-1), a.reshape(-1)], axis=1)
    s = 0.5 ** np.asarray(depth, dtype=np.float64)
    corner = np.asarray(anchor, dtype=np.float64).reshape(-1, 3) * s[:, None]
    pts = (corner[:, None, :] + s[:, None, None] * loc[None, :, :]).reshape(-1, 3)
    w = np.repeat(s ** 3 / n ** 3, n ** 3)
    return pts, w


def check_cheb_output(anchor, depth, q: int, evaluate, fn):
    """(al2, rl2, ali, rli) of the tree against fn.  evaluate(pts [m,3]) and
    fn(pts) return [m, dof] arrays."""
    pts, w = error_points(anchor, depth, q)
    val = np.asarray(evaluate(pts), dtype=np.float64).reshape(len(pts), -1)
    ref = np.asarray(fn(pts), dtype=np.float64).reshape(len(pts), -1)
    err = val - ref
    avg = (err * w[:, None]).sum(0)           # glb_err_avg, utils.txx:158-187
    err = err - avg[None, :]
    l2 = (ref * ref * w[:, None]).sum()
    l2_err = (err * err * w[:, None]).sum()
    mx, mx_err = np.abs(ref).max(), np.abs(err).max()
    return (float(np.sqrt(l2_err)), float(np.sqrt(l2_err / l2)), float(mx_err),
            float(mx_err / mx))


def check_cheb_output_device(tree, fn, mode: str = "fp64"):
    """The same metric computed on the device (tbx_check_cheb_output) for a
    device tree against fn (torch: points [m,3] -> [m,dof] or [m])."""
    pts = tree.error_points()
    ref = fn(pts).reshape(pts.shape[0], tree.dof).contiguous()
    return tree.check_cheb_output(ref, mode)
