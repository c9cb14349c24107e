"""The reference's analytic convergence tests through the device engine,
measured by CheckChebOutput (utils.txx:118-297).

The reference stores no expected values: its sweeps (conv_adv.py,
conv_traj.py, conv_zal.py) print error tables whose slopes are read off by
hand.  The bars here are the slopes the mathematics fixes, checked on the
device results:

* trajectories (traj.cpp, conv_traj.py): second order in dt, and the end
  point equal to the closed-form RK2 map of the rotation;
* rotating Gaussian (advection.cpp test 1), against the analytic solution:
  second order in dt once the tree resolves the blob (conv_adv.py temporal),
  in FP64 and in the FP32-contraction mode with an FP64 head;
* against the semi-discrete solution (RK2 trajectories, exact
  interpolation; closed form for the rotation): spectral decay in the
  degree and order q + 1 in the leaf size (the spatial part alone);
* tol-adaptive trees (conv_adv.py spatial): the error after one step tracks
  the fit error of the initial tree;
* uniform depth x dt, one revolution (conv_adv.py long-time pairing
  d / dt / tn = d0 / 0.0628 / 100, +1 / half / double);
* Zalesak slotted sphere (advection.cpp test 2): a discontinuous field, so
  only monotone improvement with depth and degree;
* the harness on the oracle: the same errors at a small size.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.fixture(scope="module")
def cv(tb):
    from tbslas_amd import convergence

    return convergence


def test_trajectory_second_order(torch, tb, cv):
    import torch as _t

    errs = []
    for k in range(8):
        tn = 50 * 2 ** k
        errs.append(cv.run_traj(tn))
        # the end point is the closed-form RK2 map of the rotation
        anchor, dep = tb.uniform_leaves(2)
        vel = cv.velocity_tree(anchor, dep, 8)
        x0 = np.array([[0.99, 0.5, 0.5]])
        x = tb.traj_rk2(vel, _t.from_numpy(x0).cuda(), 2 * math.pi, 0.0, tn).cpu().numpy()
        assert np.abs(x - cv.rk2_map(x0, 2 * math.pi / tn, tn)).max() <= 1e-12, tn
        vel.close()
    orders = [cv.order(a, b) for a, b in zip(errs, errs[1:])]
    print("trajectory errors", ["%.3e" % e for e in errs], "orders",
          ["%.3f" % o for o in orders])
    assert all(1.99 <= o <= 2.01 for o in orders), orders


@pytest.mark.parametrize("mode,spatial", [("fp64", 1e-9), ("fp32h", 2e-7)])
def test_rotating_gaussian_second_order_in_dt(torch, tb, cv, mode, spatial):
    # conv_adv.py temporal: q = 14, T = 1, dt halved; uniform depth 4
    # resolves the blob to ~1e-12, so the error is RK2's
    anchor, depth = tb.uniform_leaves(4)
    rows = [cv.run_advection(1, anchor, depth, 14, 0.5 ** k, 2 ** k, mode) for k in range(2, 7)]
    for r in rows:
        print(cv._row(r))
        assert r["in"][1] <= 1e-11
        assert r["out_rk2"][1] <= spatial and r["out_rk2"][3] <= spatial
    for a, b in zip(rows, rows[1:]):
        o2, oi = cv.order(a["out"][1], b["out"][1]), cv.order(a["out"][3], b["out"][3])
        assert 1.95 <= o2 <= 2.1 and 1.95 <= oi <= 2.15, (a["dt"], o2, oi)


def test_rotating_gaussian_spectral_in_degree(torch, tb, cv):
    # depth 3, 16 steps of 1/16: the spatial error (against the semi-discrete
    # solution) falls by more than 6x per two degrees
    anchor, depth = tb.uniform_leaves(3)
    rows = [cv.run_advection(1, anchor, depth, q, 1 / 16, 16) for q in (4, 6, 8, 10, 12, 14)]
    e = [r["out_rk2"][1] for r in rows]
    print("degree sweep rl2 vs rk2", ["%.3e" % x for x in e])
    assert all(a / b >= 6.0 for a, b in zip(e, e[1:])), e
    assert e[-1] <= 2e-7


@pytest.mark.parametrize("q", [4, 6])
def test_rotating_gaussian_order_in_leaf_size(torch, tb, cv, q):
    # uniform depths 3, 4, 5 with 16 steps of 1/16: order -> q + 1
    rows = []
    for d in (3, 4, 5):
        anchor, depth = tb.uniform_leaves(d)
        rows.append(cv.run_advection(1, anchor, depth, q, 1 / 16, 16))
    e = [r["out_rk2"][1] for r in rows]
    orders = [cv.order(a, b) for a, b in zip(e, e[1:])]
    print(f"q {q} depth sweep rl2 vs rk2", ["%.3e" % x for x in e], "orders",
          ["%.2f" % o for o in orders])
    assert orders[0] >= q and orders[1] >= q + 0.5, orders


def test_adaptive_tolerance_sweep(torch, tb, cv):
    # conv_adv.py spatial: q = 14, one step of 1e-3 on tol-adaptive trees;
    # the output error is the fit error of the tree, carried one step
    prev = None
    c = cv.concentration(1)
    for tol in (1e-2, 1e-4, 1e-6, 1e-8):
        anchor, depth = tb.adaptive_leaves(lambda p: c(p, 0.0), 14, 1, tol, 6)
        r = cv.run_advection(1, anchor, depth, 14, 1e-3, 1)
        print(f"tol {tol:.0e}", cv._row(r))
        assert r["out"][1] <= 2.0 * r["in"][1] + 1e-12
        assert r["out"][1] <= 1e3 * tol
        if prev is not None:
            assert r["out"][1] < prev
        prev = r["out"][1]


def test_uniform_depth_and_dt_one_revolution(torch, tb, cv):
    # the long-time pairing of conv_adv.py (d, dt, tn) = (5, 0.0628, 100),
    # (6, /2, 200), (7, /4, 400), one depth lower: q = 3, one revolution
    rows = list(cv.sweep_long(depths=(4, 5, 6)))
    for r in rows:
        print(cv._row(r))
    orders = [cv.order(a["out"][1], b["out"][1]) for a, b in zip(rows, rows[1:])]
    assert all(o >= 2.5 for o in orders), orders


def test_zalesak_improves_with_depth_and_degree(torch, tb, cv):
    # conv_zal.py at one revolution and depths 4, 5 (the reference: 5 and 7,
    # five revolutions); the slotted sphere is discontinuous
    rows = {(r["max_depth"], r["q"]): r for r in cv.sweep_zalesak(depths=(4, 5), degrees=(3, 8))}
    for r in rows.values():
        print(cv._row(r))
        assert r["out"][1] >= 0.5 * r["in"][1]
    for d in (4, 5):
        assert rows[(d, 8)]["out"][1] < rows[(d, 3)]["out"][1]
    for q in (3, 8):
        assert rows[(5, q)]["out"][1] < rows[(4, q)]["out"][1]


def test_harness_errors_equal_the_oracle(torch, tb, orc, cv):
    # the same run (test 1, depth 2, q = 6, 5 steps of 0.1) through the
    # oracle's in-situ step: the same coefficients and the same errors
    from tbslas_amd import fields, metrics

    anchor, depth = tb.uniform_leaves(2)
    q, dt, tn = 6, 0.1, 5
    r = cv.run_advection(1, anchor, depth, q, dt, tn, "exact", keep_coeff=True)
    c = cv.concentration(1)
    ov = orc.Tree(anchor, depth, q, 3,
                  fields.rotation_coefficients(anchor, depth, q, (0, 0, 1), 1.0))
    oc = orc.Tree(anchor, depth, q, 1, tb.fit_function(lambda p: c(p, 0.0), anchor, depth, q, 1))
    M = orc.pt2coeff(q)
    for n in range(1, tn + 1):
        orc.semilag_insitu(ov, oc, M, n, dt, 1, 0)
    assert np.abs(r["coeff"] - oc.coeff).max() <= 1e-12 * np.abs(oc.coeff).max()
    ev = lambda p: orc.eval_tree(oc, p)[0]
    e_out = metrics.check_cheb_output(anchor, depth, q, ev, lambda p: c(p, tn * dt))
    assert np.allclose(e_out, r["out"], rtol=1e-9, atol=0), (e_out, r["out"])


def test_zalesak_adaptive_tree_follows_the_sphere(torch, tb, cv):
    # the reference's adaptive Zalesak run (conv_zal.py:47-98: tol 1e-10,
    # dt 0.0628, the tree refined and balanced after every step) at maximum
    # depth 5 for one quarter revolution: the tree changes from step to step
    # and stays balanced, and q = 8 beats q = 3 (the reference's table runs
    # depths 5 and 7 over 5 revolutions: `python -m tbslas_amd.convergence
    # zalesak_adaptive`)
    rows = {}
    for q in (3, 8):
        r = cv.run_advection_adaptive(2, q, 6.28e-2, 25, 1e-10, 5)
        print(r)
        assert r["leaves_max"] > r["leaves_min"]
        assert r["max_depth"] == 5
        rows[q] = r
    assert rows[8]["out"][1] < rows[3]["out"][1]
