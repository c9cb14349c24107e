"""The reference's analytic convergence drivers, run through the device
engine and measured by CheckChebOutput (metrics.py).

Test cases (advection.cpp:92-146), both with the rigid rotation
get_vorticity_field (fields.h:142-151, omega = 1 about (0.5, 0.5)) and
FreeSpace boundaries:

  1  rotating Gaussian, get_gaussian_field_atT (field_wrappers.h:31-56);
  2  Zalesak slotted sphere, get_slotted_cylinder_atT (field_wrappers.h:158-168:
     the slot turned by the angle t).

A run (advection.cpp:290-340) fits the concentration at t = 0, measures it
("In"), takes tn in-situ steps of dt (SolveSemilagInSitu, timestep 1..tn) and
measures the result against the analytic field at t = tn dt ("Out").  The
trajectory driver (traj.cpp:115-146) carries (0.99, 0.5, 0.5) backward over
tn RK2 steps of one revolution and reports |x_T - x_0|.

Sweeps (CLI: python -m tbslas_amd.convergence SWEEP):
  temporal  q = 14, dt = 1 .. 1/128, T = 1 (conv_adv.py:20-96), uniform depth 4
            in place of tol = 1e-9;
  spatial   q = 14, dt = 1e-3, one step, tol = 1e-1 .. 1e-8 (conv_adv.py:98-173);
  long      q = 3, uniform depth 5, 6, 7 with dt = 0.0628 / 1, 2, 4 and
            tn = 100, 200, 400 (conv_adv.py:212-266);
  zalesak   test 2, q = 3, 8, 14 at uniform depths (conv_zal.py:47-98 uses
            d = 5, 7 with tol = 1e-10 and 5 revolutions);
  traj      T = 2 pi, tn = 50 * 2^k, k = 0..7 (conv_traj.py:44-79).
  zalesak_adaptive
            the reference's table: test 2, tol = 1e-10, maximum depth 5 and 7,
            q = 3, 8, 14, dt = 0.0628, 5 revolutions, the concentration tree
            refined, merged and 2:1-balanced on the device after every step
            (advection.cpp:300-315, Tree.adapt).
The other sweeps keep the tree of t = 0 (uniform or tol-adaptive) fixed.

The RK2 departure map of the rotation is linear, so the semi-discrete
solution (exact interpolation, RK2 trajectories) is known in closed form:
one backward step maps z = (x - 0.5) + i (y - 0.5) to
((1 - dt^2 / 2) - i dt) z (traj.inc:36-42 with tau = -dt and
v = i (z) in complex form), and after n steps c(x) = c_0(M^n x).
rk2_solution returns it; the error against it is the spatial part alone.
"""
from __future__ import annotations

import math
import sys

import numpy as np

from . import fields, metrics
from ._lib import BC_FREESPACE
from .tree import (Tree, adaptive_leaves, fit_function, semilag_insitu, traj_rk2,
                   uniform_leaves)

OMEGA = 1.0
CENTRE = (0.5, 0.5, 0.5)


def concentration(test: int):
    """fn(pts, t) -> [n, 1]: the analytic concentration of a test case."""
    if test == 1:
        return lambda p, t: fields.gaussian_at(p, t).reshape(-1, 1)
    if test == 2:
        return lambda p, t: fields.slotted_sphere(p, a=OMEGA * t)
    raise ValueError(f"unknown test case {test}")


def rk2_map(pts, dt: float, n: int):
    """Departure points of n backward RK2 steps of dt through the rotation."""
    p = np.array(pts, dtype=np.float64, copy=True).reshape(-1, 3)
    z = (p[:, 0] - CENTRE[0]) + 1j * (p[:, 1] - CENTRE[1])
    z = z * ((1.0 - 0.5 * (OMEGA * dt) ** 2) - 1j * (OMEGA * dt)) ** n
    p[:, 0] = z.real + CENTRE[0]
    p[:, 1] = z.imag + CENTRE[1]
    return p


def rk2_solution(test: int, dt: float, n: int):
    """fn(pts) of the semi-discrete solution after n steps of dt."""
    c = concentration(test)
    return lambda p: c(rk2_map(p, dt, n), 0.0)


def velocity_tree(anchor, depth, q: int) -> Tree:
    """The rotation on every leaf (closed-form coefficients: the field is
    linear, so the fit is exact)."""
    cf = fields.rotation_coefficients(anchor, depth, q, (0.0, 0.0, 1.0), OMEGA, CENTRE)
    return Tree(anchor, depth, q, 3, cf)


def _evaluator(tree: Tree):
    import torch

    def ev(p):
        out = np.empty((len(p), tree.dof))
        for a in range(0, len(p), 1 << 22):
            x = torch.from_numpy(np.ascontiguousarray(p[a:a + (1 << 22)])).cuda()
            out[a:a + len(x)] = tree.eval(x, BC_FREESPACE, "fp64").cpu().numpy()
        return out

    return ev


def run_advection(test: int, anchor, depth, q: int, dt: float, tn: int,
                  mode: str = "fp64", keep_coeff: bool = False) -> dict:
    """One advection run; errors (al2, rl2, ali, rli) of the fit ("in"), of
    the result against the analytic field at tn dt ("out") and against the
    semi-discrete solution ("out_rk2"); the final coefficients under "coeff"
    when keep_coeff."""
    c = concentration(test)
    vel = velocity_tree(anchor, depth, q)
    conc = Tree(anchor, depth, q, 1, fit_function(lambda p: c(p, 0.0), anchor, depth, q, 1))
    ev = _evaluator(conc)
    e_in = metrics.check_cheb_output(anchor, depth, q, ev, lambda p: c(p, 0.0))
    for n in range(1, tn + 1):
        semilag_insitu(vel, conc, n, dt, 1, BC_FREESPACE, mode)
    e_out = metrics.check_cheb_output(anchor, depth, q, ev, lambda p: c(p, tn * dt))
    e_rk2 = metrics.check_cheb_output(anchor, depth, q, ev, rk2_solution(test, dt, tn))
    res = {"test": test, "L": len(depth), "max_depth": int(np.max(depth)), "q": q, "dt": dt,
           "tn": tn, "mode": mode, "in": e_in, "out": e_out, "out_rk2": e_rk2,
           "coeff": conc.get_coeffs() if keep_coeff else None}
    vel.close()
    conc.close()
    return res


def run_advection_adaptive(test: int, q: int, dt: float, tn: int, tol: float, max_depth: int,
                           min_depth: int = 3, mode: str = "fp64", progress=None) -> dict:
    """The reference's adaptive run (advection.cpp:254-341): the
    concentration tree built to tol, then after every in-situ step refined,
    merged and 2:1-balanced on the device (RefineTree + Balance21,
    advection.cpp:305-310); errors measured on the device
    (CheckChebOutput, tbx_check_cheb_output) at t = 0 and at tn dt."""
    import torch

    c = concentration(test)

    def ref_at(t):
        return lambda x: torch.from_numpy(
            np.ascontiguousarray(c(x.cpu().numpy(), t), dtype=np.float64)).cuda()

    anchor, depth = adaptive_leaves(lambda p: c(p, 0.0), q, 1, tol, max_depth, min_depth)
    conc = Tree(anchor, depth, q, 1, fit_function(lambda p: c(p, 0.0), anchor, depth, q, 1))
    va, vd = uniform_leaves(min_depth)
    vel = velocity_tree(va, vd, q)
    e_in = metrics.check_cheb_output_device(conc, ref_at(0.0))
    leaves = [conc.L]
    for n in range(1, tn + 1):
        semilag_insitu(vel, conc, n, dt, 1, BC_FREESPACE, mode)
        conc.adapt(tol, min_depth, max_depth, BC_FREESPACE)
        leaves.append(conc.L)
        if progress and n % progress == 0:
            print(f"# step {n}: {conc.L} leaves", flush=True)
    e_out = metrics.check_cheb_output_device(conc, ref_at(tn * dt))
    res = {"test": test, "L": conc.L, "max_depth": int(np.max(conc.depth)), "q": q, "dt": dt,
           "tn": tn, "mode": mode, "tol": tol, "in": e_in, "out": e_out,
           "leaves_min": min(leaves), "leaves_max": max(leaves), "L0": leaves[0]}
    vel.close()
    conc.close()
    return res


def run_traj(tn: int, q: int = 8, depth: int = 2, mode: str = "fp64") -> float:
    """traj.cpp:115-146: |x_T - x_0| after one backward revolution."""
    import torch

    anchor, dep = uniform_leaves(depth)
    vel = velocity_tree(anchor, dep, q)
    x0 = np.array([[0.99, 0.5, 0.5]])
    T = 2.0 * math.pi / OMEGA
    x = traj_rk2(vel, torch.from_numpy(x0).cuda(), T, 0.0, tn, BC_FREESPACE, mode).cpu().numpy()
    vel.close()
    return float(np.sqrt(((x - x0) ** 2).sum()))


def order(e_coarse: float, e_fine: float, ratio: float = 2.0) -> float:
    """Observed order between two runs whose resolution differs by ratio."""
    return math.log(e_coarse / e_fine) / math.log(ratio)


# ---------------------------------------------------------------------------
# Sweeps
# ---------------------------------------------------------------------------

def _row(r) -> str:
    i, o = r["in"], r["out"]
    return (f"test {r['test']} q {r['q']:2d} L {r['L']:8d} maxd {r['max_depth']} "
            f"dt {r['dt']:.4e} tn {r['tn']:4d} {r['mode']:5s} | in al2 {i[0]:.3e} rl2 "
            f"{i[1]:.3e} rli {i[3]:.3e} | out al2 {o[0]:.3e} rl2 {o[1]:.3e} ali {o[2]:.3e} "
            f"rli {o[3]:.3e} | vs rk2 rl2 {r['out_rk2'][1]:.3e} rli {r['out_rk2'][3]:.3e}")


def sweep_temporal(mode="fp64", depth=4, levels=8):
    anchor, dep = uniform_leaves(depth)
    rows = []
    for k in range(levels):
        dt = 0.5 ** k
        rows.append(run_advection(1, anchor, dep, 14, dt, int(round(1.0 / dt)), mode))
        yield rows[-1]


def sweep_spatial(mode="fp64", levels=8, max_depth=6):
    c = concentration(1)
    for k in range(levels):
        tol = 0.1 ** (k + 1)
        anchor, dep = adaptive_leaves(lambda p: c(p, 0.0), 14, 1, tol, max_depth)
        yield run_advection(1, anchor, dep, 14, 1e-3, 1, mode)


def sweep_long(mode="fp64", depths=(5, 6, 7)):
    for k, d in enumerate(depths):
        anchor, dep = uniform_leaves(d)
        yield run_advection(1, anchor, dep, 3, 0.0628 / 2 ** k, 100 * 2 ** k, mode)


def sweep_zalesak(mode="fp64", depths=(4, 5), degrees=(3, 8, 14), revolutions=1):
    dt = 0.0628
    tn = int(round(6.28 * revolutions / dt))
    for d in depths:
        anchor, dep = uniform_leaves(d)
        for q in degrees:
            yield run_advection(2, anchor, dep, q, dt, tn, mode)


def sweep_zalesak_adaptive(mode="fp64", depths=(5, 7), degrees=(3, 8, 14), revolutions=5,
                           tol=1e-10, progress=100):
    dt = 6.28e-2
    tn = int(round(6.28 * revolutions / dt))
    for d in depths:
        for q in degrees:
            yield run_advection_adaptive(2, q, dt, tn, tol, d, mode=mode, progress=progress)


def main(argv=None) -> int:
    import argparse

    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("sweep", choices=["temporal", "spatial", "long", "zalesak",
                                      "zalesak_adaptive", "traj"])
    ap.add_argument("--depths", default=None, help="comma list (zalesak_adaptive)")
    ap.add_argument("--degrees", default=None, help="comma list (zalesak_adaptive)")
    ap.add_argument("--revolutions", type=float, default=5)
    ap.add_argument("--progress", type=int, default=100,
                    help="print the leaf count every this many steps (zalesak_adaptive)")
    ap.add_argument("--mode", default="fp64")
    a = ap.parse_args(argv)
    if a.sweep == "traj":
        prev = None
        for k in range(8):
            tn = 50 * 2 ** k
            e = run_traj(tn, mode=a.mode)
            o = "" if prev is None else f" order {order(prev, e):.3f}"
            print(f"traj tn {tn:5d} dt {2 * math.pi / tn:.4e} err {e:.4e}{o}", flush=True)
            prev = e
        return 0
    if a.sweep == "zalesak_adaptive":
        kw = {"revolutions": a.revolutions, "progress": a.progress}
        if a.depths:
            kw["depths"] = tuple(int(x) for x in a.depths.split(","))
        if a.degrees:
            kw["degrees"] = tuple(int(x) for x in a.degrees.split(","))
        for r in sweep_zalesak_adaptive(a.mode, **kw):
            i, o = r["in"], r["out"]
            print(f"zalesak adaptive q {r['q']:2d} maxd {r['max_depth']} tol {r['tol']:.0e} "
                  f"dt {r['dt']:.4e} tn {r['tn']} leaves {r['L0']} -> {r['L']} (min "
                  f"{r['leaves_min']}, max {r['leaves_max']}) | in rl2 {i[1]:.3e} rli "
                  f"{i[3]:.3e} | out al2 {o[0]:.3e} rl2 {o[1]:.3e} ali {o[2]:.3e} rli {o[3]:.3e}",
                  flush=True)
        return 0
    gen = {"temporal": sweep_temporal, "spatial": sweep_spatial, "long": sweep_long,
           "zalesak": sweep_zalesak}[a.sweep](a.mode)
    prev = None
    for r in gen:
        o = ""
        if prev is not None and a.sweep in ("temporal", "long"):
            o = (f" | order rl2 {order(prev['out'][1], r['out'][1]):.2f} rli "
                 f"{order(prev['out'][3], r['out'][3]):.2f}")
        print(_row(r) + o, flush=True)
        prev = r
    return 0


if __name__ == "__main__":
    sys.exit(main())
