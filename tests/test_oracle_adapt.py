"""Pins the oracle's restatement of adaptivity (RefineTree + Balance21,
advection.cpp:305-314, tree_utils.h:117-123) by what the mathematics fixes:
a polynomial of total degree <= q survives refinement and merging exactly;
the result is 2:1 balanced, and its balance step gives the same leaves as
the library's independent host balance (tbx_balance21)."""
import numpy as np
import pytest

from conftest import random_tree


def poly_field(q, seed):
    rng = np.random.default_rng(seed)
    terms = [(i, j, k) for i in range(q + 1) for j in range(q + 1 - i) for k in range(q + 1 - i - j)]
    a = rng.standard_normal(len(terms))

    def f(p):
        return sum(a[m] * p[:, 0] ** k * p[:, 1] ** j * p[:, 2] ** i
                   for m, (i, j, k) in enumerate(terms))[:, None]
    return f


def balanced(anchor, depth, periodic):
    import oracle as orc

    af = anchor.astype(np.float64)
    d = depth.astype(np.int64)
    for m in range(len(depth)):
        adj = orc._adjacent(af[m], int(d[m]), af, d, periodic)
        adj[m] = False
        if (np.abs(d[adj] - d[m]) > 1).any():
            return False
    return True


@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("tol", [0.0, 1e9])
def test_adapt_keeps_polynomials_exact(orc, tb, bc, tol):
    q = 4
    anchor, depth = random_tree(31, max_depth=3, p_split=0.3, min_depth=2)
    f = poly_field(q, 5)
    cf = tb.fit_function(f, anchor, depth, q, 1)
    a, d, c = orc.adapt(anchor, depth, cf, q, tol, 1, 4, bc)
    if tol == 0.0:
        assert len(d) > len(depth) and d.max() == 4  # everything below 4 split
    else:
        assert len(d) < len(depth)  # sibling groups merged
    assert balanced(a, d, bc == 1)
    t = orc.Tree(a, d, q, 1, c)
    pts = np.random.default_rng(1).random((4000, 3))
    got, _ = orc.eval_tree(t, pts, 0)
    ref = f(pts)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("bc", [0, 1])
def test_adapt_balance_is_the_library_balance(orc, tb, bc):
    # a Gaussian refined at one spot to depth 5 in a depth-2 tree forces a
    # balance cascade; the leaves after adaptation equal the library's host
    # 2:1 balance of the refined list
    from tbslas_amd import fields

    q = 3
    anchor, depth = tb.uniform_leaves(2)
    fn = lambda p: fields.gaussian(p, (0.3, 0.62, 0.55), 0.03)
    a, d = anchor, depth
    c = tb.fit_function(fn, a, d, q, 1)
    for _ in range(3):
        a, d, c = orc.adapt(a, d, c, q, 1e-4, 2, 5, bc)
        c = tb.fit_function(fn, a, d, q, 1)
    assert d.max() >= 4 and balanced(a, d, bc == 1)
    ba, bd = tb.balance21(a, d, bc)
    assert np.array_equal(ba, a) and np.array_equal(bd, d)
