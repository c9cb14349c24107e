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


# ---------------------------------------------------------------------------
# Adaptivity (RefineTree + Balance21, advection.cpp:305-314,
# tree_utils.h:117-123) restated on the host for the device's tbx_tree_adapt.
# PVFMM's subdivision test is not in the reference repository; the rule
# here is the engine's documented one: split a leaf whose top-shell tail
# (sum of |c_ijk| over i + j + k = q, largest over components) exceeds tol
# below max_depth; merge 8 sibling leaves that all stay, above min_depth,
# when their parent's fit has a tail <= tol; then split leaves more than one
# level coarser than an adjacent leaf until none is (found here by brute
# force over leaf pairs, an independent method from the device's probes).
# ---------------------------------------------------------------------------

def _shell_terms(q):
    out, m = [], 0
    for i in range(q + 1):
        for j in range(q + 1 - i):
            for k in range(q + 1 - i - j):
                if i + j + k == q:
                    out.append(m)
                m += 1
    return out


def tail_norm(coeff, q):
    """Top-shell |c| sums in packed order, one term after another (the
    device's order), largest over components: [L]."""
    c = np.asarray(coeff).reshape(coeff.shape[0], coeff.shape[1], -1)
    s = np.zeros(c.shape[:2])
    for m in _shell_terms(q):
        s = s + np.abs(c[:, :, m])
    return s.max(axis=1)


def _triangle(q):
    return [(i, j, k) for i in range(q + 1) for j in range(q + 1 - i) for k in range(q + 1 - i - j)]


def eval_in_leaf(coeff, anchor, depth, pts, q):
    """Each leaf's own expansion at its points: coeff [n][dof][Nc], pts
    [n][m][3] -> [n][dof][m]; leaf coordinates clamped to [-1, 1]."""
    h = 0.5 ** np.asarray(depth, dtype=np.float64)
    corner = np.asarray(anchor, dtype=np.float64) * h[:, None]
    xi = ((pts - corner[:, None, :]) * 2.0) / h[:, None, None] - 1.0
    xi = np.clip(xi, -1.0, 1.0)
    T = [np.polynomial.chebyshev.chebvander(xi[..., a], q) for a in range(3)]  # [n][m][d]
    tri = _triangle(q)
    B = np.stack([T[0][..., k] * T[1][..., j] * T[2][..., i] for i, j, k in tri], axis=-1)
    return np.einsum("nlc,nmc->nlm", coeff, B)


def _fit_leaves_from(q, dof, anchor, depth, vals):
    t = Tree(anchor, depth, q, dof)
    set_grid_values(t, pt2coeff(q), np.ascontiguousarray(vals))
    return t.coeff.copy()


def _resample(q, dof, coeff_src, anchor_src, depth_src, src, anchor_new, depth_new, children):
    """Coefficients of new leaves from old leaves' polynomials at the new
    leaves' stretched nodes (children: the node's octant picks one of the 8
    children starting at src[n])."""
    if len(src) == 0:
        return np.zeros((0, dof, ncoef(q)))
    x = new_nodes(q)
    d = q + 1
    m = np.arange(d ** 3)
    loc = np.stack([x[m % d], x[(m // d) % d], x[m // (d * d)]], axis=1)
    h = 0.5 ** np.asarray(depth_new, dtype=np.float64)
    pts = np.asarray(anchor_new, dtype=np.float64)[:, None, :] * h[:, None, None] + \
        h[:, None, None] * loc[None]
    out = np.empty((len(src), dof, d ** 3))
    oct_ = (loc[:, 0] >= 0.5) * 1 + (loc[:, 1] >= 0.5) * 2 + (loc[:, 2] >= 0.5) * 4
    for n in range(len(src)):
        if children:
            s = src[n] + oct_
            for o in range(8):
                sel = oct_ == o
                if sel.any():
                    out[n][:, sel] = eval_in_leaf(coeff_src[s[sel][:1]], anchor_src[s[sel][:1]],
                                                  depth_src[s[sel][:1]], pts[n][sel][None],
                                                  q)[0]
        else:
            out[n] = eval_in_leaf(coeff_src[src[n]:src[n] + 1], anchor_src[src[n]:src[n] + 1],
                                  depth_src[src[n]:src[n] + 1], pts[n][None], q)[0]
    return _fit_leaves_from(q, dof, anchor_new, depth_new, out)


def _adjacent(a0, d0, a1, d1, periodic):
    """Do the closed boxes of two leaves touch (periodic: across faces)?"""
    lo0, hi0 = a0 * 2.0 ** -d0, (a0 + 1) * 2.0 ** -d0
    h1 = (2.0 ** -np.asarray(d1, dtype=np.float64))[:, None]
    lo1, hi1 = a1 * h1, (a1 + 1) * h1
    ok = np.ones(len(d1), dtype=bool)
    for ax in range(3):
        touch = (lo1[:, ax] <= hi0[ax]) & (lo0[ax] <= hi1[:, ax])
        if periodic:
            touch |= (hi1[:, ax] == 1.0) & (lo0[ax] == 0.0)
            touch |= (hi0[ax] == 1.0) & (lo1[:, ax] == 0.0)
        ok &= touch
    return ok


def _split_list(anchor, depth, mask):
    reps = np.where(mask, 8, 1)
    idx = np.repeat(np.arange(len(depth)), reps)
    na = anchor[idx].astype(np.uint32)
    nd = depth[idx].astype(np.uint8)
    child = np.zeros(len(idx), dtype=np.int64)
    starts = np.cumsum(reps) - reps
    for c in range(8):
        child[starts[mask] + c] = c
    mm = np.repeat(mask, reps)
    for ax in range(3):
        na[mm, ax] = 2 * na[mm, ax] + ((child[mm] >> ax) & 1)
    nd[mm] += 1
    return na, nd, idx


def adapt(anchor, depth, coeff, q, tol, min_depth, max_depth, bc):
    """(anchor, depth, coeff) after one adaptation (see the section note)."""
    anchor = np.asarray(anchor, dtype=np.uint32).reshape(-1, 3)
    depth = np.asarray(depth, dtype=np.uint8)
    coeff = np.asarray(coeff, dtype=np.float64)
    L, dof = len(depth), coeff.shape[1]
    tail = tail_norm(coeff, q)
    refine = (tail > tol) & (depth < max_depth)
    cand = []
    j = 0
    while j + 8 <= L:
        f = anchor[j].astype(np.int64)
        ok = depth[j] > min_depth and not (f & 1).any()
        for c in range(8):
            if not ok:
                break
            g = anchor[j + c].astype(np.int64)
            ok = (depth[j + c] == depth[j] and g[0] == f[0] + (c & 1) and
                  g[1] == f[1] + ((c >> 1) & 1) and g[2] == f[2] + ((c >> 2) & 1) and
                  not refine[j + c])
        if ok:
            cand.append(j)
            j += 8
        else:
            j += 1
    cand = np.array(cand, dtype=np.int64)
    pa = anchor[cand] // 2 if len(cand) else np.zeros((0, 3), np.uint32)
    pd = depth[cand] - 1 if len(cand) else np.zeros(0, np.uint8)
    pcoef = _resample(q, dof, coeff, anchor, depth, cand, pa, pd, True)
    merge = tail_norm(pcoef, q) <= tol if len(cand) else np.zeros(0, dtype=bool)
    na, nd, nc = [], [], []
    k = 0
    j = 0
    while j < L:
        if k < len(cand) and cand[k] == j:
            if merge[k]:
                na.append(pa[k]), nd.append(pd[k]), nc.append(pcoef[k])
                j += 8
                k += 1
                continue
            k += 1
        if refine[j]:
            ca, cd, _ = _split_list(anchor[j:j + 1], depth[j:j + 1], np.array([True]))
            cc = _resample(q, dof, coeff, anchor, depth, np.full(8, j), ca, cd, False)
            for c in range(8):
                na.append(ca[c]), nd.append(cd[c]), nc.append(cc[c])
        else:
            na.append(anchor[j]), nd.append(depth[j]), nc.append(coeff[j])
        j += 1
    anchor = np.array(na, dtype=np.uint32)
    depth = np.array(nd, dtype=np.uint8)
    coeff = np.array(nc)
    while True:
        split = np.zeros(len(depth), dtype=bool)
        af = anchor.astype(np.float64)
        for m in range(len(depth)):
            adj = _adjacent(af[m], int(depth[m]), af, depth.astype(np.int64), bc == BC_PERIODIC)
            adj[m] = False
            split |= adj & (depth.astype(np.int64) + 1 < int(depth[m]))
        if not split.any():
            break
        ca, cd, src = _split_list(anchor, depth, split)
        newc = np.empty((len(cd), dof, coeff.shape[2]))
        keep = np.repeat(~split, np.where(split, 8, 1))
        newc[keep] = coeff[src[keep]]
        ch = ~keep
        newc[ch] = _resample(q, dof, coeff, anchor, depth, src[ch], ca[ch], cd[ch], False)
        anchor, depth, coeff = ca, cd, newc
    return anchor, depth, coeff
