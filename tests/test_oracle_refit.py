"""Pins the refit operator M = pinv(B)^T and SetTreeGridValues."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def basis_matrix(orc, q):
    d = q + 1
    x = orc.new_nodes(q) * 2.0 - 1.0
    T = np.polynomial.chebyshev.chebvander(x, q)  # [node, k]
    cols = []
    for i in range(d):
        for j in range(d - i):
            for k in range(d - i - j):
                idx = np.arange(d ** 3)
                cols.append(T[idx % d, k] * T[(idx // d) % d, j] * T[idx // (d * d), i])
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("q", [1, 2, 3, 4, 6, 8])
def test_pt2coeff_is_numpy_pinv(orc, q):
    M = orc.pt2coeff(q)
    B = basis_matrix(orc, q)
    Mn = np.linalg.pinv(B).T
    assert np.abs(M - Mn).max() <= 1e-13 * np.abs(Mn).max()
    # projection: M^T B = I on the triangle space
    assert np.abs(M.T @ B - np.eye(B.shape[1])).max() <= 1e-13


def test_pt2coeff_golden(orc):
    g = np.load(os.path.join(GOLDEN, "pinv_q1_4.npz"))
    for q in range(1, 5):
        assert np.abs(orc.pt2coeff(q) - g[f"q{q}"]).max() <= 1e-14


def test_library_pt2coeff_matches_oracle(orc, tb):
    for q in (2, 5, 8, 10):
        a, b = tb.pt2coeff_matrix(q), orc.pt2coeff(q)
        assert np.abs(a - b).max() <= 2e-15 * np.abs(b).max()


@pytest.mark.parametrize("q,dof", [(3, 1), (6, 2)])
def test_set_grid_values_reproduces_polynomials(orc, tb, q, dof):
    from conftest import random_tree

    anchor, depth = random_tree(40 + q, max_depth=3)
    t = orc.Tree(anchor, depth, q, dof)
    g = orc.collect_grid_points(t)  # [L*P, 3]
    rng = np.random.default_rng(q)
    terms = [(i, j, k) for i in range(q + 1) for j in range(q + 1 - i) for k in range(q + 1 - i - j)]
    a = rng.standard_normal((len(terms), dof))

    def poly(p):
        return sum(np.outer(p[:, 0] ** k * p[:, 1] ** j * p[:, 2] ** i, a[m])
                   for m, (i, j, k) in enumerate(terms))

    vals = poly(g).reshape(t.L, t.P, dof).transpose(0, 2, 1).copy()  # [L][dof][P]
    orc.set_grid_values(t, orc.pt2coeff(q), vals)
    pts = rng.random((3000, 3))
    got, _ = orc.eval_tree(t, pts, 0)
    assert np.abs(got - poly(pts)).max() <= 1e-12 * np.abs(a).sum()


def test_grid_points_layout(orc, tb):
    anchor, depth = tb.uniform_leaves(1)
    q = 3
    t = orc.Tree(anchor, depth, q, 1)
    g = orc.collect_grid_points(t).reshape(8, 4 ** 3, 3)
    x = orc.new_nodes(q)
    # x fastest, then y, then z (cheb.h:60-67); leaf corner + 2^-depth * node
    m = np.arange(64)
    for j in range(8):
        c = anchor[j] * 0.5
        assert (g[j, :, 0] == c[0] + 0.5 * x[m % 4]).all()
        assert (g[j, :, 1] == c[1] + 0.5 * x[(m // 4) % 4]).all()
        assert (g[j, :, 2] == c[2] + 0.5 * x[m // 16]).all()


@pytest.mark.parametrize("q", [1, 2, 4, 8, 10, 14])
def test_factored_refit_operator(orc, tb, q):
    # the device refit applies pinv(B) = H B^T: H must invert the Gram
    # matrix B^T B and H B^T must be the oracle's operator M^T
    B = basis_matrix(orc, q)
    H = tb.refit_gram_inverse(q)
    assert np.abs(H - H.T).max() <= 1e-14 * np.abs(H).max()
    assert np.abs(H @ (B.T @ B) - np.eye(B.shape[1])).max() <= 1e-12
    M = orc.pt2coeff(q)
    assert np.abs(H @ B.T - M.T).max() <= 1e-13 * np.abs(M).max()
