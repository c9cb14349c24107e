"""The CheckChebOutput metric (utils.txx:118-297) on host trees: exact for
represented polynomials, blind to a constant offset (the mean error is
removed), and normalised so that a constant field's L2 norm is the constant
(the leaves tile the unit cube)."""
import numpy as np

from tbslas_amd import metrics


def test_metric_zero_for_represented_polynomial_and_offsets(orc, tb):
    anchor, depth = tb.uniform_leaves(2)
    q = 5
    f = lambda p: (p[:, 0] ** 2 * p[:, 1] - 0.3 * p[:, 2] ** 3 + 0.7)[:, None]
    cf = tb.fit_function(f, anchor, depth, q, 1)
    t = orc.Tree(anchor, depth, q, 1, cf)
    ev = lambda p: orc.eval_tree(t, p)[0]
    al2, rl2, ali, rli = metrics.check_cheb_output(anchor, depth, q, ev, f)
    assert ali <= 1e-13 and rl2 <= 1e-13
    # a constant offset is the mean error and is removed
    g = lambda p: f(p) - 0.25
    assert metrics.check_cheb_output(anchor, depth, q, ev, g)[2] <= 1e-13


def test_metric_normalisation():
    anchor = np.zeros((1, 3), dtype=np.uint32)
    depth = np.zeros(1, dtype=np.uint8)
    pts, w = metrics.error_points(anchor, depth, 8)
    assert len(pts) == 27 and abs(w.sum() - 1.0) <= 1e-15
    # error = a field with zero mean over the nodes: L-inf and L2 of it
    x = metrics.first_kind_nodes(3)
    assert np.allclose(sorted(set(np.round(pts[:, 0], 12))), np.sort(x))
    ev = lambda p: 2.0 + (p[:, 0] - 0.5)[:, None]
    al2, rl2, ali, rli = metrics.check_cheb_output(anchor, depth, 8, ev,
                                                   lambda p: np.full((len(p), 1), 2.0))
    e = x - 0.5
    assert abs(ali - np.abs(e).max()) <= 1e-15
    assert abs(al2 - np.sqrt((e * e).mean())) <= 1e-15
    assert abs(rl2 - al2 / 2.0) <= 1e-15 and abs(rli - ali / 2.0) <= 1e-15


def test_rk2_map_is_the_oracle_trajectory(orc, tb):
    # the closed-form semi-discrete map of the convergence harness against
    # the oracle's ComputeTrajRK2 through a fitted rotation tree
    from tbslas_amd import convergence as cv
    from tbslas_amd import fields

    anchor, depth = tb.uniform_leaves(2)
    q = 4
    vel = orc.Tree(anchor, depth, q, 3,
                   fields.rotation_coefficients(anchor, depth, q, (0, 0, 1), 1.0))
    rng = np.random.default_rng(5)
    x = 0.3 + 0.4 * rng.random((200, 3))
    for dt, n in ((0.1, 1), (0.0628, 20), (0.25, 8)):
        got = orc.traj_rk2(vel, x, n * dt, 0.0, n)
        assert np.abs(got - cv.rk2_map(x, dt, n)).max() <= 1e-13, (dt, n)
    # the exact solution: rk2_map approaches the rotation by -t as dt -> 0
    t = 1.0
    exact = cv.rk2_map(x, t, 0)
    c, s = np.cos(-t), np.sin(-t)
    exact[:, 0] = 0.5 + c * (x[:, 0] - 0.5) - s * (x[:, 1] - 0.5)
    exact[:, 1] = 0.5 + s * (x[:, 0] - 0.5) + c * (x[:, 1] - 0.5)
    e = [np.abs(cv.rk2_map(x, t / n, n) - exact).max() for n in (64, 128)]
    assert 3.9 <= e[0] / e[1] <= 4.1


def test_slotted_sphere_turns_with_the_rotation():
    # get_slotted_cylinder_atT: the slot at angle t is the slot at 0 carried
    # by the rotation for time t
    from tbslas_amd import convergence as cv

    rng = np.random.default_rng(6)
    x = rng.random((20000, 3))
    c = cv.concentration(2)
    t = 0.7
    back = x.copy()
    cs, sn = np.cos(-t), np.sin(-t)
    back[:, 0] = 0.5 + cs * (x[:, 0] - 0.5) - sn * (x[:, 1] - 0.5)
    back[:, 1] = 0.5 + sn * (x[:, 0] - 0.5) + cs * (x[:, 1] - 0.5)
    a, b = c(x, t), c(back, 0.0)
    assert (a != b).mean() <= 1e-3  # only points within rounding of a face
    assert 0.05 <= a.mean() <= 0.12
