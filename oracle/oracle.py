"""ctypes binding of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py (reference results / CPU baseline).  The engine
(tbslas_amd, libtbx.so) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

BC_FREESPACE = 0
BC_PERIODIC = 1
MAXD = 30

TFAC_FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


class OrcTree(C.Structure):
    _fields_ = [
        ("num_leaves", C.c_int64),
        ("anchor", C.c_void_p),
        ("depth", C.c_void_p),
        ("q", C.c_int32),
        ("dof", C.c_int32),
        ("coeff", C.c_void_p),
    ]


class OrcVelocity(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("t", C.c_void_p * 4),
        ("tt", C.c_double * 4),
        ("g", TFAC_FN),
        ("gctx", C.c_void_p),
    ]


VEL_TREE, VEL_EXTRAP, VEL_SET = 0, 1, 2


class OrcGrid(C.Structure):
    _fields_ = [
        ("num_leaves", C.c_int64),
        ("anchor", C.c_void_p),
        ("depth", C.c_void_p),
        ("n", C.c_int32),
        ("dof", C.c_int32),
        ("samples", C.c_void_p),
    ]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sig = {
            "orc_ncoef": (C.c_int, [C.c_int]),
            "orc_new_nodes": (None, [C.c_int, P]),
            "orc_cheb_nodes_first_kind": (None, [C.c_int, P]),
            "orc_cheb_poly": (None, [C.c_int, P, C.c_int64, P]),
            "orc_locate": (C.c_int, [P, P, C.c_int64, C.c_int, P]),
            "orc_eval_tree": (C.c_int, [P, P, C.c_int64, C.c_int, P]),
            "orc_traj_rk2": (C.c_int, [P, TFAC_FN, P, P, C.c_int64, C.c_double, C.c_double,
                                       C.c_int, C.c_int, P]),
            "orc_semilag_rk2": (C.c_int, [P, TFAC_FN, P, P, P, C.c_int64, C.c_int, C.c_double,
                                          C.c_int, C.c_int, P]),
            "orc_collect_grid_points": (C.c_int, [P, P]),
            "orc_pt2coeff": (C.c_int, [C.c_int, P]),
            "orc_set_grid_values": (C.c_int, [P, P, P]),
            "orc_semilag_insitu": (C.c_int, [P, TFAC_FN, P, P, P, C.c_int, C.c_double, C.c_int,
                                             C.c_int]),
            "orc_cheb_approx": (C.c_int, [C.c_int, C.c_int, P, P]),
            "orc_fast_interp": (C.c_int, [P, C.c_int, C.c_int, P, C.c_int64, P]),
            "orc_cubic1d": (C.c_double, [C.c_double, P, P]),
            "orc_field_set_eval": (C.c_int, [P, P, P, C.c_int64, C.c_double, C.c_int, P]),
            "orc_field_extrap_eval": (C.c_int, [P, P, P, C.c_int64, C.c_int, P]),
            "orc_grid_eval": (C.c_int, [P, P, C.c_int64, C.c_int, P]),
            "orc_grid_points": (C.c_int, [P, P]),
            "orc_grid_traj_rk2": (C.c_int, [P, TFAC_FN, P, P, C.c_int64, C.c_double, C.c_double,
                                            C.c_int, C.c_int, P]),
            "orc_grid_semilag_insitu": (C.c_int, [P, TFAC_FN, P, P, C.c_int, C.c_double,
                                                  C.c_int, C.c_int]),
            "orc_traj_rk2_v": (C.c_int, [P, P, C.c_int64, C.c_double, C.c_double, C.c_int,
                                         C.c_int, P]),
            "orc_semilag_insitu_v": (C.c_int, [P, P, P, C.c_int, C.c_double, C.c_int,
                                               C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def ncoef(q):
    return (q + 1) * (q + 2) * (q + 3) // 6


class Tree:
    """Host tree for the oracle; keeps its arrays alive."""

    def __init__(self, anchor, depth, q, dof, coeff=None):
        self.anchor = np.ascontiguousarray(anchor, dtype=np.uint32).reshape(-1, 3)
        self.depth = np.ascontiguousarray(depth, dtype=np.uint8).reshape(-1)
        self.q, self.dof = int(q), int(dof)
        self.L = len(self.depth)
        self.nc = ncoef(self.q)
        self.P = (self.q + 1) ** 3
        if coeff is None:
            coeff = np.zeros((self.L, self.dof, self.nc))
        self.coeff = np.ascontiguousarray(coeff, dtype=np.float64).reshape(self.L, self.dof,
                                                                           self.nc).copy()
        self.s = OrcTree(self.L, _p(self.anchor), _p(self.depth), self.q, self.dof,
                         _p(self.coeff))

    @property
    def ref(self):
        return C.byref(self.s)


def _tfac(g):
    if g is None:
        return TFAC_FN()
    return TFAC_FN(lambda t, ctx: float(g(t)))


def new_nodes(q):
    x = np.empty(q + 1)
    lib().orc_new_nodes(q, _p(x))
    return x


def cheb_nodes_first_kind(q):
    x = np.empty(q + 1)
    lib().orc_cheb_nodes_first_kind(q, _p(x))
    return x


def cheb_poly(q, x):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    out = np.empty((q + 1, x.size))
    lib().orc_cheb_poly(q, _p(x), x.size, _p(out))
    return out


def locate(tree, pts, bc=BC_FREESPACE):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    out = np.empty(len(pts), dtype=np.int32)
    _check(lib().orc_locate(tree.ref, _p(pts), len(pts), bc, _p(out)), "locate")
    return out


def eval_tree(tree, pts, bc=BC_FREESPACE, inplace=False):
    """EvalTree; returns (values [n,dof], wrapped points)."""
    p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    if not inplace:
        p = p.copy()
    out = np.empty((len(p), tree.dof))
    _check(lib().orc_eval_tree(tree.ref, _p(p), len(p), bc, _p(out)), "eval_tree")
    return out, p


def traj_rk2(vel, xinit, tinit, tfinal, num_rk_step, bc=BC_FREESPACE, g=None):
    x = np.ascontiguousarray(xinit, dtype=np.float64).reshape(-1, 3)
    out = np.empty_like(x)
    cb = _tfac(g)
    _check(lib().orc_traj_rk2(vel.ref, cb, None, _p(x), len(x), tinit, tfinal, num_rk_step, bc,
                              _p(out)), "traj_rk2")
    return out


def semilag_rk2(vel, conc, pts, timestep, dt, num_rk_step=1, bc=BC_FREESPACE, g=None):
    p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    out = np.empty((len(p), conc.dof))
    cb = _tfac(g)
    _check(lib().orc_semilag_rk2(vel.ref, cb, None, conc.ref, _p(p), len(p), timestep, dt,
                                 num_rk_step, bc, _p(out)), "semilag_rk2")
    return out


def collect_grid_points(tree):
    out = np.empty((tree.L * tree.P, 3))
    _check(lib().orc_collect_grid_points(tree.ref, _p(out)), "collect_grid_points")
    return out


def pt2coeff(q):
    M = np.empty(((q + 1) ** 3, ncoef(q)))
    _check(lib().orc_pt2coeff(q, _p(M)), "pt2coeff")
    return M


def set_grid_values(tree, M, vals):
    v = np.ascontiguousarray(vals, dtype=np.float64)
    _check(lib().orc_set_grid_values(tree.ref, _p(np.ascontiguousarray(M)), _p(v)),
           "set_grid_values")


def semilag_insitu(vel, conc, M, timestep, dt, num_rk_step=1, bc=BC_FREESPACE, g=None):
    cb = _tfac(g)
    _check(lib().orc_semilag_insitu(vel.ref, cb, None, conc.ref, _p(np.ascontiguousarray(M)),
                                    timestep, dt, num_rk_step, bc), "semilag_insitu")


def cheb_approx(q, dof, vals):
    v = np.ascontiguousarray(vals, dtype=np.float64)
    out = np.empty((dof, ncoef(q)))
    _check(lib().orc_cheb_approx(q, dof, _p(v), _p(out)), "cheb_approx")
    return out


def fast_interp(grid, dof, N, pts):
    g = np.ascontiguousarray(grid, dtype=np.float64)
    p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    out = np.empty((len(p), dof))
    _check(lib().orc_fast_interp(_p(g), dof, N, _p(p), len(p), _p(out)), "fast_interp")
    return out


def cubic1d(x, xx, pp):
    xx = np.ascontiguousarray(xx, dtype=np.float64)
    pp = np.ascontiguousarray(pp, dtype=np.float64)
    return lib().orc_cubic1d(x, _p(xx), _p(pp))


def field_set_eval(trees, tt, pts, t, bc=BC_FREESPACE):
    p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3).copy()
    arr = (C.c_void_p * 4)(*[C.addressof(tr.s) for tr in trees])
    tt = np.ascontiguousarray(tt, dtype=np.float64)
    out = np.empty((len(p), trees[0].dof))
    _check(lib().orc_field_set_eval(arr, _p(tt), _p(p), len(p), t, bc, _p(out)),
           "field_set_eval")
    return out


def field_extrap_eval(tp, tc, pts, bc=BC_FREESPACE):
    p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3).copy()
    out = np.empty((len(p), tc.dof))
    _check(lib().orc_field_extrap_eval(tp.ref, tc.ref, _p(p), len(p), bc, _p(out)),
           "field_extrap_eval")
    return out


def velocity(kind, trees, times=None, g=None):
    """A velocity functor: VEL_TREE (trees[0], optional g), VEL_EXTRAP
    (trees[0] current, trees[1] previous) or VEL_SET (4 trees at times)."""
    v = OrcVelocity()
    v.kind = kind
    for i, t in enumerate(trees):
        v.t[i] = C.addressof(t.s)
    if times is not None:
        for i, x in enumerate(times):
            v.tt[i] = x
    cb = _tfac(g)
    v.g = cb
    v._keep = (trees, cb)
    return v


def traj_rk2_v(vel, xinit, tinit, tfinal, num_rk_step, bc=BC_FREESPACE):
    x = np.ascontiguousarray(xinit, dtype=np.float64).reshape(-1, 3)
    out = np.empty_like(x)
    _check(lib().orc_traj_rk2_v(C.byref(vel), _p(x), len(x), tinit, tfinal, num_rk_step, bc,
                                _p(out)), "traj_rk2_v")
    return out


def semilag_insitu_v(vel, conc, M, timestep, dt, num_rk_step=1, bc=BC_FREESPACE):
    _check(lib().orc_semilag_insitu_v(C.byref(vel), conc.ref, _p(np.ascontiguousarray(M)),
                                      timestep, dt, num_rk_step, bc), "semilag_insitu_v")


class Grid:
    """Leaves with regular n^3 sample grids, samples [L][dof][n^3]."""

    def __init__(self, anchor, depth, n, dof, samples=None):
        self.anchor = np.ascontiguousarray(anchor, dtype=np.uint32).reshape(-1, 3)
        self.depth = np.ascontiguousarray(depth, dtype=np.uint8).reshape(-1)
        self.n, self.dof = int(n), int(dof)
        self.L = len(self.depth)
        self.P = self.n ** 3
        if samples is None:
            samples = np.zeros((self.L, self.dof, self.P))
        self.samples = np.ascontiguousarray(samples, dtype=np.float64).reshape(
            self.L, self.dof, self.P).copy()
        self.s = OrcGrid(self.L, _p(self.anchor), _p(self.depth), self.n, self.dof,
                         _p(self.samples))

    @property
    def ref(self):
        return C.byref(self.s)


def grid_eval(grid, pts, bc=BC_FREESPACE):
    p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3).copy()
    out = np.empty((len(p), grid.dof))
    _check(lib().orc_grid_eval(grid.ref, _p(p), len(p), bc, _p(out)), "grid_eval")
    return out


def grid_points(grid):
    out = np.empty((grid.L * grid.P, 3))
    _check(lib().orc_grid_points(grid.ref, _p(out)), "grid_points")
    return out


def grid_traj_rk2(vel, xinit, tinit, tfinal, num_rk_step, bc=BC_FREESPACE, g=None):
    x = np.ascontiguousarray(xinit, dtype=np.float64).reshape(-1, 3)
    out = np.empty_like(x)
    cb = _tfac(g)
    _check(lib().orc_grid_traj_rk2(vel.ref, cb, None, _p(x), len(x), tinit, tfinal, num_rk_step,
                                   bc, _p(out)), "grid_traj_rk2")
    return out


def grid_semilag_insitu(vel, conc, timestep, dt, num_rk_step=1, bc=BC_FREESPACE, g=None):
    cb = _tfac(g)
    _check(lib().orc_grid_semilag_insitu(vel.ref, cb, None, conc.ref, timestep, dt, num_rk_step,
                                         bc), "grid_semilag_insitu")
