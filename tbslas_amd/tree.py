"""Linear octrees with Chebyshev leaves on the device, and the operations of
the semi-Lagrangian hot path over the libtbx C ABI.

Layout (PVFMM Cheb_Node, include/tbx.h): leaves in Morton order, each with an
integer anchor in units of its own edge and a depth; coefficients
[L][dof][Nc] as packed triangles i(z) -> j(y) -> k(x), i+j+k <= q.

Device buffers are torch tensors on cuda; every call runs on torch's current
stream.  Host-side builders (uniform / adaptive leaves, 2:1 balance, initial
fit) are setup helpers in the same library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (BC_FREESPACE, BC_PERIODIC, MODES, VEL_EXTRAP, VEL_SET, VEL_TREE,
                   WRITE_BACK_WRAP, _Velocity, check, lib, tfac)


def ncoef(q: int) -> int:
    return (q + 1) * (q + 2) * (q + 3) // 6


def _ptr(a) -> int:
    if a is None:
        return 0
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


_device_ready = False


def _init_device() -> None:
    """Initialise torch's HIP context before libtbx makes its first device
    call: both share one HIP runtime, and torch's device counting does not
    initialise it."""
    global _device_ready
    if not _device_ready:
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: libtbx device calls need an MI355X")
        torch.cuda.init()
        torch.empty(1, device="cuda")
        _device_ready = True


def _stream() -> int:
    import torch

    _init_device()
    return torch.cuda.current_stream().cuda_stream


def _require_dev(t, dtype, name, shape_tail=None):
    import torch

    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a cuda tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if shape_tail is not None and tuple(t.shape[1:]) != tuple(shape_tail):
        raise ValueError(f"{name} must have shape [n,{shape_tail}], got {tuple(t.shape)}")


def _check_dep(dep, npts: int) -> None:
    """A departure-point output buffer, when given, must be a contiguous
    FP64 cuda tensor [npts][3]: the step kernels write every row of it."""
    if dep is None:
        return
    import torch

    _require_dev(dep, torch.float64, "dep", (3,))
    if dep.shape[0] != npts:
        raise ValueError(f"dep must have {npts} rows (one per grid point), got {dep.shape[0]}")


def _mode(mode: str) -> int:
    try:
        return MODES[mode]
    except KeyError:
        raise ValueError(f"mode must be one of {sorted(MODES)}") from None


class Tree:
    """One rank's octree on the GPU (the object NodeFieldFunctor points at,
    tree_functor.h:793-815)."""

    def __init__(self, anchor, depth, q: int, dof: int, coeff=None):
        anchor = np.ascontiguousarray(anchor, dtype=np.uint32).reshape(-1, 3)
        depth = np.ascontiguousarray(depth, dtype=np.uint8).reshape(-1)
        if anchor.shape[0] != depth.shape[0]:
            raise ValueError("anchor and depth lengths differ")
        self.anchor, self.depth = anchor, depth
        self.L = int(depth.shape[0])
        self.q, self.dof = int(q), int(dof)
        self.nc = ncoef(self.q)
        self.P = (self.q + 1) ** 3
        cf = None
        if coeff is not None:
            cf = np.ascontiguousarray(coeff, dtype=np.float64)
            if cf.size != self.L * self.dof * self.nc:
                raise ValueError("coefficient array has the wrong size")
        self._validate_host()
        _init_device()
        desc = _lib._TreeDesc(self.L, _ptr(anchor), _ptr(depth), self.q, self.dof, _ptr(cf))
        h = C.c_void_p()
        check(lib().tbx_tree_create(C.byref(desc), C.byref(h)))
        self._h = h

    def _validate_host(self):
        """Argument checks that need no device (the C ABI repeats them)."""
        if self.L == 0:
            raise _lib.TbxError(1, "tbx_tree_create: empty leaf list")
        if not (1 <= self.q <= _lib.MAX_Q):
            raise _lib.TbxError(7, f"tbx_tree_create: degree {self.q} outside 1..{_lib.MAX_Q}")

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().tbx_tree_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        L, q, dof, md = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().tbx_tree_info(self._h, C.byref(L), C.byref(q), C.byref(dof), C.byref(md)))
        return L.value, q.value, dof.value, md.value

    def set_coeffs(self, coeff):
        import torch

        if isinstance(coeff, torch.Tensor):
            _require_dev(coeff, torch.float64, "coeff")
            if coeff.numel() != self.L * self.dof * self.nc:
                raise ValueError("coefficient tensor has the wrong size")
            check(lib().tbx_tree_set_coeffs(self._h, _ptr(coeff), 1, _stream()))
        else:
            cf = np.ascontiguousarray(coeff, dtype=np.float64)
            if cf.size != self.L * self.dof * self.nc:
                raise ValueError("coefficient array has the wrong size")
            check(lib().tbx_tree_set_coeffs(self._h, _ptr(cf), 0, _stream()))

    def get_coeffs(self) -> np.ndarray:
        out = np.empty((self.L, self.dof, self.nc), dtype=np.float64)
        check(lib().tbx_tree_get_coeffs(self._h, _ptr(out), 0, _stream()))
        return out

    def get_coeffs_device(self):
        import torch

        out = torch.empty((self.L, self.dof, self.nc), dtype=torch.float64, device="cuda")
        check(lib().tbx_tree_get_coeffs(self._h, _ptr(out), 1, _stream()))
        return out

    def locate(self, pts, bc: int = BC_FREESPACE):
        import torch

        _require_dev(pts, torch.float64, "pts", (3,))
        out = torch.empty(pts.shape[0], dtype=torch.int32, device=pts.device)
        check(lib().tbx_locate(self._h, _ptr(pts), pts.shape[0], bc, _ptr(out), _stream()))
        return out

    def eval(self, pts, bc: int = BC_FREESPACE, mode: str = "fp64", write_back: bool = False):
        """EvalTree: values [n][dof] at pts [n][3]."""
        import torch

        _require_dev(pts, torch.float64, "pts", (3,))
        out = torch.empty((pts.shape[0], self.dof), dtype=torch.float64, device=pts.device)
        flags = _mode(mode) | (WRITE_BACK_WRAP if write_back else 0)
        check(lib().tbx_eval(self._h, _ptr(pts), pts.shape[0], bc, flags, _ptr(out), _stream()))
        return out

    def contraction_only(self, n: int, waves_per_leaf: int, mode: str = "fp64", out=None,
                         pts=None):
        """tbx_contraction_only: the leaf contraction alone at n hashed points,
        point i in leaf (i // 64 // waves_per_leaf) mod L; out[i] is the sum
        of the leaf's components there, pts (optional, [n][3]) receives the
        points.  Returns out."""
        import torch

        if out is None:
            out = torch.empty(n, dtype=torch.float64, device="cuda")
        _require_dev(out, torch.float64, "out")
        if out.shape[0] < n:
            raise ValueError(f"out holds {out.shape[0]} values, need {n}")
        if pts is not None:
            _require_dev(pts, torch.float64, "pts", (3,))
            if pts.shape[0] < n:
                raise ValueError(f"pts holds {pts.shape[0]} points, need {n}")
        check(lib().tbx_contraction_only(self._h, n, waves_per_leaf, _mode(mode), _ptr(out),
                                         _ptr(pts) if pts is not None else None, _stream()))
        return out

    def grid_points(self):
        import torch

        out = torch.empty((self.L * self.P, 3), dtype=torch.float64, device="cuda")
        check(lib().tbx_collect_grid_points(self._h, _ptr(out), _stream()))
        return out

    def first_kind_points(self):
        """First-kind Chebyshev points of every leaf on the device
        [L*(q+1)^3][3], where the initial field is sampled for fit()."""
        import torch

        out = torch.empty((self.L * self.P, 3), dtype=torch.float64, device="cuda")
        check(lib().tbx_tree_first_kind_points(self._h, _ptr(out), _stream()))
        return out

    def fit(self, vals):
        """InitTree's fit on the device: vals [L*(q+1)^3][dof] (point-major)
        at first_kind_points() become the coefficients."""
        import torch

        _require_dev(vals, torch.float64, "vals")
        if vals.numel() != self.L * self.P * self.dof:
            raise ValueError("vals must hold L*(q+1)^3*dof values")
        check(lib().tbx_tree_fit(self._h, _ptr(vals), _stream()))

    def fit_function(self, fn):
        """Fit fn (torch: points [n][3] -> values [n][dof] or [n]) on the
        device."""
        pts = self.first_kind_points()
        v = fn(pts).reshape(pts.shape[0], self.dof).contiguous()
        self.fit(v)

    def error_points(self):
        """The CheckChebOutput points [L*n^3][3], n = q/3 + 1."""
        import torch

        n = self.q // 3 + 1
        out = torch.empty((self.L * n ** 3, 3), dtype=torch.float64, device="cuda")
        check(lib().tbx_error_points(self._h, _ptr(out), _stream()))
        return out

    def check_cheb_output(self, ref, mode: str = "fp64"):
        """CheckChebOutput on the device against ref [L*n^3][dof] (the exact
        field at error_points()): (al2, rl2, ali, rli)."""
        import torch

        _require_dev(ref, torch.float64, "ref")
        n = self.q // 3 + 1
        if ref.numel() != self.L * n ** 3 * self.dof:
            raise ValueError("ref must hold L*n^3*dof values")
        res = np.zeros(4, dtype=np.float64)
        check(lib().tbx_check_cheb_output(self._h, _ptr(ref), _mode(mode), _ptr(res),
                                          _stream()))
        return tuple(float(v) for v in res)

    def adapt(self, tol: float, min_depth: int, max_depth: int, bc: int = BC_FREESPACE) -> int:
        """RefineTree + Balance21 on the device (tbx_tree_adapt): split
        leaves whose top-shell tail exceeds tol (below max_depth), merge
        sibling groups whose parent fit's tail is <= tol (above min_depth),
        then 2:1-balance.  Updates anchor, depth and L; returns L."""
        check(lib().tbx_tree_adapt(self._h, float(tol), int(min_depth), int(max_depth), bc,
                                   _stream()))
        L = C.c_int64()
        q, dof, md = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().tbx_tree_info(self._h, C.byref(L), C.byref(q), C.byref(dof), C.byref(md)))
        self.L = int(L.value)
        self.anchor = np.empty((self.L, 3), dtype=np.uint32)
        self.depth = np.empty(self.L, dtype=np.uint8)
        check(lib().tbx_tree_get_leaves(self._h, _ptr(self.anchor), _ptr(self.depth)))
        return self.L

    def write_vtu(self, prefix: str, order: int = 4) -> None:
        """Write2File: <prefix>.pvtu and <prefix>_000000.vtu with the tree's
        values on order^3 uniform points per leaf (VTK hexahedra)."""
        check(lib().tbx_tree_write_vtu(self._h, prefix.encode(), order, _stream()))

    def grid_values(self, out=None):
        """The tree's values at its own grid points, [L][dof][P] (each leaf's
        polynomial at its nodes, upper faces included); a float32 out
        receives them rounded to FP32."""
        import torch

        if out is None:
            out = torch.empty((self.L, self.dof, self.P), dtype=torch.float64, device="cuda")
        f32 = out.dtype == torch.float32
        _require_dev(out, torch.float32 if f32 else torch.float64, "out")
        if out.numel() != self.L * self.dof * self.P:
            raise ValueError("out must hold L*dof*P values")
        fn = lib().tbx_tree_grid_values_f32 if f32 else lib().tbx_tree_grid_values
        check(fn(self._h, _ptr(out), _stream()))
        return out

    def set_grid_values(self, vals):
        """SetTreeGridValues with vals [L][dof][P] (dof-major per leaf)."""
        import torch

        _require_dev(vals, torch.float64, "vals")
        if vals.numel() != self.L * self.dof * self.P:
            raise ValueError("vals must hold L*dof*P values")
        check(lib().tbx_set_grid_values(self._h, _ptr(vals), _stream()))


class Velocity:
    """A velocity functor over dof-3 device trees of one degree.

    Velocity.tree(t, g)        NodeFieldFunctor (tree_functor.h:793-815),
                               scaled by g(t) when g is given;
    Velocity.extrap(cur, prev) the extrapolating RK2 overloads
                               (traj.inc:70-115): cur for the first stage,
                               FieldExtrapFunctor(prev, cur) =
                               1.5 cur - 0.5 prev for the second
                               (tree_extrap_functor.h:47-78);
    Velocity.set(trees, times) FieldSetFunctor (tree_set_functor.h:49-79),
                               cubic Hermite in time over four trees.
    """

    def __init__(self, kind: int, trees, times=None, g=None):
        n = {VEL_TREE: 1, VEL_EXTRAP: 2, VEL_SET: 4}[kind]
        if len(trees) != n:
            raise ValueError(f"velocity kind {kind} takes {n} trees")
        self.trees = list(trees)
        self._cb = tfac(g)
        self._s = _Velocity()
        self._s.kind = kind
        for i, t in enumerate(self.trees):
            self._s.trees[i] = t.handle
        if times is not None:
            if len(times) != 4:
                raise ValueError("a velocity set takes four times")
            for i, x in enumerate(times):
                self._s.times[i] = float(x)
        self._s.g = self._cb

    @classmethod
    def tree(cls, t: "Tree", g=None) -> "Velocity":
        return cls(VEL_TREE, [t], g=g)

    @classmethod
    def extrap(cls, cur: "Tree", prev: "Tree") -> "Velocity":
        return cls(VEL_EXTRAP, [cur, prev])

    @classmethod
    def set(cls, trees, times) -> "Velocity":
        return cls(VEL_SET, trees, times)

    @property
    def ref(self):
        return C.byref(self._s)

    def eval(self, pts, t: float, bc: int = BC_FREESPACE, mode: str = "fp64",
             write_back: bool = False):
        """The functor's value at pts [n][3] at time t (the extrapolating
        functor gives its FieldExtrapFunctor value)."""
        import torch

        _require_dev(pts, torch.float64, "pts", (3,))
        out = torch.empty((pts.shape[0], 3), dtype=torch.float64, device=pts.device)
        flags = _mode(mode) | (WRITE_BACK_WRAP if write_back else 0)
        check(lib().tbx_velocity_eval(self.ref, _ptr(pts), pts.shape[0], t, bc, flags,
                                      _ptr(out), _stream()))
        return out


def _velocity(vel, g) -> Velocity:
    if isinstance(vel, Velocity):
        if g is not None:
            raise ValueError("pass g to Velocity.tree, not with a Velocity")
        return vel
    return Velocity.tree(vel, g)


def traj_rk2(vel, xinit, tinit: float, tfinal: float, num_rk_step: int,
             bc: int = BC_FREESPACE, mode: str = "fp64", g=None, out=None):
    """ComputeTrajRK2 (traj.inc:48-68, 97-115) through vel's functor (a Tree,
    optionally scaled by g, or a Velocity)."""
    import torch

    v = _velocity(vel, g)
    _require_dev(xinit, torch.float64, "xinit", (3,))
    xsol = torch.empty_like(xinit) if out is None else out
    check(lib().tbx_traj_rk2_v(v.ref, _ptr(xinit), xinit.shape[0], tinit, tfinal, num_rk_step,
                               bc, _mode(mode), _ptr(xsol), _stream()))
    return xsol


def semilag_rk2(vel: Tree, conc: Tree, pts, timestep: int, dt: float, num_rk_step: int = 1,
                bc: int = BC_FREESPACE, mode: str = "fp64", g=None):
    """SolveSemilagRK2 (semilag.inc:26-45): conc at the departure points."""
    import torch

    _require_dev(pts, torch.float64, "pts", (3,))
    vals = torch.empty((pts.shape[0], conc.dof), dtype=torch.float64, device=pts.device)
    cb = tfac(g)
    check(lib().tbx_semilag_rk2(vel.handle, cb, None, conc.handle, _ptr(pts), pts.shape[0],
                                timestep, dt, num_rk_step, bc, _mode(mode), _ptr(vals),
                                _stream()))
    return vals


def semilag_insitu(vel, conc: Tree, timestep: int, dt: float, num_rk_step: int = 1,
                   bc: int = BC_FREESPACE, mode: str = "fp64", g=None, dep=None):
    """SolveSemilagInSitu (tree_semilag.h:92-181): one fused step, conc refit
    in place.  dep (optional cuda tensor [L*P][3]) receives departure points."""
    v = _velocity(vel, g)
    _check_dep(dep, conc.L * conc.P)
    check(lib().tbx_semilag_insitu_v(v.ref, conc.handle, timestep, dt, num_rk_step, bc,
                                     _mode(mode), _ptr(dep), _stream()))


def semilag_values(vel, conc: Tree, timestep: int, dt: float, num_rk_step: int = 1,
                   bc: int = BC_FREESPACE, mode: str = "fp64", g=None, vals=None, dep=None):
    """SolveSemilagInSitu without the refit: conc at the departure points of
    its own grid points, returned as [L][dof][P] (the set_grid_values layout)."""
    import torch

    v = _velocity(vel, g)
    if vals is None:
        vals = torch.empty((conc.L, conc.dof, conc.P), dtype=torch.float64, device="cuda")
    _require_dev(vals, torch.float64, "vals")
    if vals.numel() != conc.L * conc.dof * conc.P:
        raise ValueError("vals must hold L*dof*P values")
    _check_dep(dep, conc.L * conc.P)
    check(lib().tbx_semilag_values_v(v.ref, conc.handle, timestep, dt, num_rk_step, bc,
                                     _mode(mode), _ptr(vals), _ptr(dep), _stream()))
    return vals


# ---------------------------------------------------------------------------
# Host-side setup helpers
# ---------------------------------------------------------------------------

def new_nodes(q: int) -> np.ndarray:
    x = np.empty(q + 1, dtype=np.float64)
    check(lib().tbx_new_nodes(q, _ptr(x)))
    return x


def pt2coeff_matrix(q: int) -> np.ndarray:
    M = np.empty(((q + 1) ** 3, ncoef(q)), dtype=np.float64)
    check(lib().tbx_pt2coeff_matrix(q, _ptr(M)))
    return M


def refit_gram_inverse(q: int) -> np.ndarray:
    """H = (B^T B)^-1 [Nc][Nc], the factor of pinv(B) = H B^T the device
    refit applies."""
    H = np.empty((ncoef(q), ncoef(q)), dtype=np.float64)
    check(lib().tbx_refit_gram_inverse(q, _ptr(H)))
    return H


def uniform_leaves(depth: int):
    L = 8 ** depth
    anchor = np.empty((L, 3), dtype=np.uint32)
    dep = np.empty(L, dtype=np.uint8)
    check(lib().tbx_uniform_leaves(depth, _ptr(anchor), _ptr(dep)))
    return anchor, dep


def first_kind_points(anchor, depth, q: int) -> np.ndarray:
    anchor = np.ascontiguousarray(anchor, dtype=np.uint32)
    depth = np.ascontiguousarray(depth, dtype=np.uint8)
    L = depth.shape[0]
    out = np.empty((L, (q + 1) ** 3, 3), dtype=np.float64)
    check(lib().tbx_first_kind_points(L, _ptr(anchor), _ptr(depth), q, _ptr(out)))
    return out


def fit_leaves(vals, q: int, dof: int) -> np.ndarray:
    """vals [L][P][dof] at first_kind_points -> coeff [L][dof][Nc]."""
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    P = (q + 1) ** 3
    L = vals.size // (P * dof)
    out = np.empty((L, dof, ncoef(q)), dtype=np.float64)
    check(lib().tbx_fit_leaves(L, q, dof, _ptr(vals), _ptr(out)))
    return out


def tail_norm(coeff, q: int, dof: int) -> np.ndarray:
    coeff = np.ascontiguousarray(coeff, dtype=np.float64)
    L = coeff.size // (dof * ncoef(q))
    out = np.empty(L, dtype=np.float64)
    check(lib().tbx_tail_norm(L, q, dof, _ptr(coeff), _ptr(out)))
    return out


def balance21(anchor, depth, bc: int = BC_FREESPACE, capacity: int | None = None):
    L = len(depth)
    cap = capacity or max(8 * L, L + 4096)
    a = np.zeros((cap, 3), dtype=np.uint32)
    d = np.zeros(cap, dtype=np.uint8)
    a[:L] = anchor
    d[:L] = depth
    n = C.c_int64(L)
    check(lib().tbx_balance21(C.byref(n), cap, _ptr(a), _ptr(d), bc))
    return a[: n.value].copy(), d[: n.value].copy()


def fit_function(fn, anchor, depth, q: int, dof: int, batch: int = 4096) -> np.ndarray:
    """Initial fit of fn (numpy: pts [n,3] -> values [n,dof]) on every leaf."""
    L = len(depth)
    out = np.empty((L, dof, ncoef(q)), dtype=np.float64)
    P = (q + 1) ** 3
    for a in range(0, L, batch):
        b = min(L, a + batch)
        pts = first_kind_points(anchor[a:b], depth[a:b], q)
        vals = np.asarray(fn(pts.reshape(-1, 3)), dtype=np.float64).reshape(b - a, P, dof)
        out[a:b] = fit_leaves(vals, q, dof)
    return out


def _split(anchor, depth, mask):
    """Replace every masked leaf by its 8 children, keeping Morton order."""
    reps = np.where(mask, 8, 1)
    idx = np.repeat(np.arange(len(depth)), reps)
    na = anchor[idx].astype(np.uint32)
    nd = depth[idx].astype(np.uint8)
    child = np.zeros(len(idx), dtype=np.int64)
    starts = np.cumsum(reps) - reps
    for c in range(8):
        sel = starts[mask] + c
        child[sel] = c
    m = np.repeat(mask, reps)
    na[m, 0] = 2 * na[m, 0] + (child[m] & 1)
    na[m, 1] = 2 * na[m, 1] + ((child[m] >> 1) & 1)
    na[m, 2] = 2 * na[m, 2] + ((child[m] >> 2) & 1)
    nd[m] = nd[m] + 1
    return na, nd


def adaptive_leaves(fn, q: int, dof: int, tol: float, max_depth: int, min_depth: int = 0,
                    bc: int = BC_FREESPACE, balance: bool = True, batch: int = 4096):
    """Refine from a uniform min_depth tree while a leaf's top-shell
    coefficient sum (tbx_tail_norm) exceeds tol, up to max_depth, then
    2:1-balance (the RefineTree + Balance21 roles, tree_utils.h:117-123)."""
    anchor, depth = uniform_leaves(min_depth)
    fresh = np.arange(len(depth))  # leaves not yet tested
    while True:
        cand = fresh[depth[fresh] < max_depth]
        if cand.size == 0:
            break
        mask = np.zeros(len(depth), dtype=bool)
        for a in range(0, cand.size, batch):
            sel = cand[a:a + batch]
            cf = fit_function(fn, anchor[sel], depth[sel], q, dof, batch=batch)
            mask[sel] = tail_norm(cf, q, dof) > tol
        if not mask.any():
            break
        reps = np.where(mask, 8, 1)
        anchor, depth = _split(anchor, depth, mask)
        fresh = np.nonzero(np.repeat(mask, reps))[0]
    if balance:
        anchor, depth = balance21(anchor, depth, bc)
    return anchor, depth
