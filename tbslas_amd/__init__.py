"""tbslas_amd: an MI355X-native semi-Lagrangian advection engine.

The hot path of arashb/tbslas (RK2 backward trajectories, octree leaf
location, Chebyshev interpolation and refit) as hand-written HIP kernels for
gfx950 behind a C ABI (include/tbx.h, libtbx.so).  This package is the
Python host binding of that ABI.
"""
from ._lib import (BC_FREESPACE, BC_PERIODIC, EXACT, FP32, FP64, MAX_DEPTH, MAX_Q, TbxError,
                   lib)
from .grid import Grid, grid_semilag_insitu, sample_function
from .tree import (Tree, Velocity, adaptive_leaves, balance21, first_kind_points, fit_function,
                   fit_leaves, ncoef, new_nodes, pt2coeff_matrix, refit_gram_inverse, semilag_insitu,
                   semilag_rk2, semilag_values, tail_norm, traj_rk2, uniform_leaves)

__all__ = [
    "BC_FREESPACE", "BC_PERIODIC", "EXACT", "FP32", "FP64", "Grid", "MAX_DEPTH", "MAX_Q",
    "TbxError", "grid_semilag_insitu", "sample_function",
    "Tree", "Velocity", "adaptive_leaves", "balance21", "first_kind_points", "fit_function",
    "fit_leaves", "lib", "ncoef", "new_nodes", "pt2coeff_matrix", "refit_gram_inverse",
    "semilag_insitu",
    "semilag_rk2", "semilag_values", "tail_norm", "traj_rk2", "uniform_leaves",
]
