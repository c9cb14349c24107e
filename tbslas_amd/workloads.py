"""The advection workloads the benchmark measures and the parity tests check,
built in one place so both run exactly the same inputs.

adaptive_q14   the adaptive Chebyshev octree at degree 14: a 2:1-balanced
               octree seeded at uniform depth 5 and refined (up to depth 7)
               where the top-shell tail of a Gaussian blob (sigma 0.04,
               centre (0.7, 0.5, 0.5)) exceeds 1e-13 -- 36,324 leaves,
               122,593,500 grid points -- with the Taylor-Green velocity
               (sin 2pi x cos 2pi y cos 2pi z, -cos 2pi x sin 2pi y cos 2pi z,
               0) sampled on the same tree and scaled by cos(pi t);
               periodic, dt = 1e-3, one RK2 substep, 100 steps.
uniform_q8     the uniform depth-4 tree (4,096 leaves) at degree 8, the
               rigid rotation v = 2 pi (0.5 - y, x - 0.5, 0), a Gaussian blob
               (sigma 0.06 at (0.7, 0.5, 0.5)); free space, dt = 2 pi / 200,
               50 steps.
batch_problems the many-small-problems batch: depth uniform in {3,4,5,6},
               degree in {6, 8, 10}, rigid rotations about random axes.
"""
from __future__ import annotations

import math

import numpy as np

from . import fields
from .tree import adaptive_leaves, fit_function, uniform_leaves

# adaptive degree-14 workload
ADAPTIVE_Q = 14
ADAPTIVE_SEED_DEPTH, ADAPTIVE_MAX_DEPTH, ADAPTIVE_TOL = 5, 7, 1e-13
ADAPTIVE_SIGMA, ADAPTIVE_CENTRE = 0.04, (0.7, 0.5, 0.5)
ADAPTIVE_DT = 1e-3
ADAPTIVE_BC = 1  # periodic
ADAPTIVE_STEPS = 100

# uniform degree-8 workload
UNIFORM_Q, UNIFORM_DEPTH = 8, 4
UNIFORM_OMEGA = 2 * math.pi
UNIFORM_DT = 2 * math.pi / 200
UNIFORM_SIGMA, UNIFORM_CENTRE = 0.06, (0.7, 0.5, 0.5)
UNIFORM_BC = 0  # free space
UNIFORM_STEPS = 50


def taylor_green_tfac(t: float) -> float:
    """Time factor of the adaptive workload's velocity, cos(pi t)."""
    return math.cos(math.pi * t)


def adaptive_q14():
    """(anchor, depth, conc coeff [L][1][Nc], vel coeff [L][3][Nc])."""
    fn = lambda p: fields.gaussian(p, ADAPTIVE_CENTRE, ADAPTIVE_SIGMA)
    anchor, depth = adaptive_leaves(fn, ADAPTIVE_Q, 1, ADAPTIVE_TOL, max_depth=ADAPTIVE_MAX_DEPTH,
                                    min_depth=ADAPTIVE_SEED_DEPTH, bc=ADAPTIVE_BC)
    ccf = fit_function(fn, anchor, depth, ADAPTIVE_Q, 1)
    vcf = fit_function(fields.taylor_green_2d3, anchor, depth, ADAPTIVE_Q, 3)
    return anchor, depth, ccf, vcf


def uniform_q8():
    """(anchor, depth, conc coeff, vel coeff) of the uniform degree-8 case."""
    anchor, depth = uniform_leaves(UNIFORM_DEPTH)
    vcf = fit_function(lambda p: fields.rotation(p, UNIFORM_OMEGA), anchor, depth, UNIFORM_Q, 3)
    ccf = fit_function(lambda p: fields.gaussian(p, UNIFORM_CENTRE, UNIFORM_SIGMA), anchor, depth,
                       UNIFORM_Q, 1)
    return anchor, depth, ccf, vcf


def batch_problems(n: int = 256, seed: int = 33):
    """n problems (depth, degree, rotation axis): depth uniform in {3, 4, 5,
    6}, degree in {6, 8, 10}, a rigid rotation (omega = 2 pi) about a
    uniformly random axis through the centre, drawn from SplitMix64(seed) as
    (z, phi, depth, degree) per problem."""
    u = fields.splitmix64(seed, 4 * n).reshape(n, 4)
    out = []
    for p in range(n):
        z, phi = 2 * u[p, 0] - 1, 2 * math.pi * u[p, 1]
        r = math.sqrt(max(0.0, 1 - z * z))
        axis = (r * math.cos(phi), r * math.sin(phi), z)
        out.append((3 + int(4 * u[p, 2]), (6, 8, 10)[int(3 * u[p, 3])], axis))
    return out


def gaussian_grid_values(pts, centre=(0.6, 0.5, 0.5), sigma: float = 0.08):
    """The batch problems' initial blob at grid points (numpy or torch)."""
    if isinstance(pts, np.ndarray):
        return np.exp(-((pts - np.asarray(centre)) ** 2).sum(1) / (2 * sigma ** 2))
    import torch

    c = torch.tensor(centre, dtype=pts.dtype, device=pts.device)
    return torch.exp(-((pts - c) ** 2).sum(1) / (2 * sigma ** 2))
