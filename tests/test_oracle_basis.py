"""Pins the oracle's nodes and Chebyshev recurrence by closed forms."""
import math

import numpy as np


def test_new_nodes_closed_form(orc):
    # cheb.h:53-58: x_i = -cos((i+1/2) pi / d) / cos(pi / (2d)) / 2 + 1/2
    # q = 0 is degenerate (scal makes its single node 0) and unused
    for q in range(1, 20):
        d = q + 1
        x = orc.new_nodes(q)
        ref = np.array([0.5 - 0.5 * math.cos((i + 0.5) * math.pi / d) / math.cos(math.pi / (2 * d))
                        for i in range(d)])
        assert np.abs(x - ref).max() <= 4e-16
        # stretched so the end nodes sit on the leaf faces
        assert abs(x[0]) <= 1e-16 and abs(x[-1] - 1.0) <= 2.3e-16
        # symmetric about 1/2 up to rounding
        assert np.abs(x + x[::-1] - 1.0).max() <= 5e-16
        assert (np.diff(x) > 0).all()


def test_new_nodes_exact_faces_for_config_degrees(orc):
    # q = 8 and q = 14 put arrival points exactly on leaf faces, which makes
    # the half-open location rule decide which leaf evaluates them
    for q in (8, 14):
        x = orc.new_nodes(q)
        assert x[0] == 0.0 and x[-1] == 1.0


def test_first_kind_nodes(orc):
    for q in range(0, 16):
        d = q + 1
        x = orc.cheb_nodes_first_kind(q)
        ref = 0.5 - 0.5 * np.cos((np.arange(d) + 0.5) * np.pi / d)
        assert np.abs(x - ref).max() <= 2e-16


def test_cheb_poly_cos_identity(orc):
    th = np.linspace(0.0, math.pi, 1001)
    x = np.cos(th)
    for q in (0, 1, 2, 5, 8, 14, 20):
        T = orc.cheb_poly(q, x)
        for k in range(q + 1):
            assert np.abs(T[k] - np.cos(k * th)).max() <= 1e-13 * max(1, k)


def test_cheb_poly_endpoints_bounds_and_numpy(orc):
    q = 16
    T = orc.cheb_poly(q, np.array([-1.0, 1.0]))
    k = np.arange(q + 1)
    assert (T[:, 0] == (-1.0) ** k).all() and (T[:, 1] == 1.0).all()
    x = np.random.default_rng(1).uniform(-1, 1, 5000)
    T = orc.cheb_poly(q, x)
    assert np.abs(T).max() <= 1.0 + 1e-14
    V = np.polynomial.chebyshev.chebvander(x, q).T
    assert np.abs(T - V).max() <= 1e-14


def test_cheb_poly_outside_is_zero(orc):
    x = np.array([1.0 + 2 ** -52, -1.0 - 2 ** -52, 3.0, -7.5, np.nan, np.inf])
    T = orc.cheb_poly(9, x)
    assert (T == 0.0).all()
