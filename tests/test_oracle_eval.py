"""Pins the oracle's EvalTree (location + xi map + cheb_poly + triangular
contraction) against a long-double brute force and by polynomial
reproduction."""
import numpy as np
import pytest

from conftest import brute_locate, random_tree


def brute_eval(anchor, depth, coeff, q, pts, bc):
    """sum_{i+j+k<=q} c_ijk T_k(xi_x) T_j(xi_y) T_i(xi_z) in long double."""
    pts = np.asarray(pts, dtype=np.float64)
    leaf = brute_locate(anchor, depth, pts, bc)
    p = pts.copy()
    if bc == 1:
        p = np.where(p < 0.0, p + 1.0, p)
        p = np.where(p >= 1.0, p - 1.0, p)
    ld = np.longdouble
    s = ld(2.0) ** depth[leaf].astype(ld)
    c = anchor[leaf].astype(ld) / s[:, None]
    xi = (p.astype(ld) - c) * 2 * s[:, None] - 1
    Tx = np.polynomial.chebyshev.chebvander(xi[:, 0], q)
    Ty = np.polynomial.chebyshev.chebvander(xi[:, 1], q)
    Tz = np.polynomial.chebyshev.chebvander(xi[:, 2], q)
    for T, col in ((Tx, 0), (Ty, 1), (Tz, 2)):
        T[np.abs(xi[:, col]) > 1] = 0
    basis = []
    for i in range(q + 1):
        for j in range(q + 1 - i):
            for k in range(q + 1 - i - j):
                basis.append(Tx[:, k] * Ty[:, j] * Tz[:, i])
    B = np.stack(basis, axis=1)  # [n, nc]
    C = coeff[leaf].astype(ld)   # [n, dof, nc]
    return np.einsum("nm,ndm->nd", B, C).astype(np.float64)


@pytest.mark.parametrize("q,dof,bc", [(1, 1, 0), (3, 3, 0), (5, 1, 1), (6, 2, 0)])
def test_eval_matches_long_double_brute_force(orc, q, dof, bc):
    anchor, depth = random_tree(q + 20, max_depth=3)
    rng = np.random.default_rng(q)
    coeff = rng.standard_normal((len(depth), dof, orc.ncoef(q)))
    t = orc.Tree(anchor, depth, q, dof, coeff)
    pts = rng.random((3000, 3))
    if bc == 1:
        pts = pts * 1.2 - 0.1
    got, wrapped = orc.eval_tree(t, pts, bc)
    ref = brute_eval(anchor, depth, coeff, q, pts, bc)
    scale = np.abs(coeff).sum(axis=2).max()
    assert np.abs(got - ref).max() <= 1e-14 * scale
    if bc == 1:  # EvalTree wraps the caller's points in place
        assert (wrapped >= 0).all() and (wrapped < 1).all()


def test_eval_outside_freespace_is_zero(orc):
    anchor, depth = random_tree(5, max_depth=3)
    q = 4
    coeff = np.random.default_rng(0).standard_normal((len(depth), 1, orc.ncoef(q)))
    t = orc.Tree(anchor, depth, q, 1, coeff)
    pts = np.array([[1.2, 0.5, 0.5], [-0.3, 0.5, 0.5], [0.5, 0.5, 7.0], [np.nan, 0.5, 0.5]])
    got, _ = orc.eval_tree(t, pts, 0)
    assert (got == 0).all()


@pytest.mark.parametrize("q", [2, 5, 8])
def test_fit_then_eval_reproduces_polynomials(orc, tb, q):
    anchor, depth = random_tree(q, max_depth=3)
    rng = np.random.default_rng(q)
    terms = [(i, j, k) for i in range(q + 1) for j in range(q + 1 - i) for k in range(q + 1 - i - j)]
    a = rng.standard_normal(len(terms))

    def poly(p):
        p = np.asarray(p).reshape(-1, 3)
        return sum(a[m] * p[:, 0] ** k * p[:, 1] ** j * p[:, 2] ** i
                   for m, (i, j, k) in enumerate(terms))[:, None]

    coeff = tb.fit_function(poly, anchor, depth, q, 1)
    t = orc.Tree(anchor, depth, q, 1, coeff)
    pts = rng.random((4000, 3))
    got, _ = orc.eval_tree(t, pts, 0)
    ref = poly(pts)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(a).sum()


def test_cheb_approx_matches_library_fit(orc, tb):
    q, dof = 6, 2
    anchor, depth = tb.uniform_leaves(1)
    rng = np.random.default_rng(3)
    pts = tb.first_kind_points(anchor[:2], depth[:2], q)
    vals = np.sin(3 * pts.sum(axis=2, keepdims=True)) * np.array([1.0, -2.0])  # [2, P, 2]
    lib_cf = tb.fit_leaves(vals, q, dof)
    for j in range(2):
        orc_cf = orc.cheb_approx(q, dof, np.ascontiguousarray(vals[j].T))
        assert np.abs(orc_cf - lib_cf[j]).max() <= 1e-15 * 40
