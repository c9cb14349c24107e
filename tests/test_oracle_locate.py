"""Pins the oracle's leaf location (EvalTree wrap + Morton keys + sort +
lower_bound partition) against a brute-force half-open box scan."""
import numpy as np
import pytest

from conftest import brute_locate, edge_points, random_tree


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("bc", [0, 1])
def test_locate_random_trees_brute_force(orc, seed, bc):
    anchor, depth = random_tree(seed, max_depth=4)
    t = orc.Tree(anchor, depth, 2, 1)
    rng = np.random.default_rng(100 + seed)
    pts = rng.random((8000, 3))
    if bc == 1:
        pts = pts * 1.4 - 0.2  # exercise the single wrap
    # points on faces and corners of random leaves
    j = rng.integers(0, len(depth), 1000)
    h = 0.5 ** depth[j].astype(np.float64)
    corner = anchor[j] * h[:, None]
    pts = np.concatenate([pts, corner, corner + h[:, None] * rng.integers(0, 2, (1000, 3))])
    got = orc.locate(t, pts, bc)
    assert (got == brute_locate(anchor, depth, pts, bc)).all()


@pytest.mark.parametrize("bc", [0, 1])
def test_locate_edge_table(orc, bc):
    anchor, depth = random_tree(11, max_depth=4)
    t = orc.Tree(anchor, depth, 2, 1)
    pts = edge_points()
    assert (orc.locate(t, pts, bc) == brute_locate(anchor, depth, pts, bc)).all()


def test_locate_uniform_closed_form(orc, tb):
    for D in (1, 2, 3):
        anchor, depth = tb.uniform_leaves(D)
        t = orc.Tree(anchor, depth, 1, 1)
        pts = np.random.default_rng(D).random((5000, 3))
        cell = np.floor(pts * 2 ** D).astype(np.int64)
        # Morton index with z above y above x
        key = np.zeros(len(pts), dtype=np.int64)
        for b in range(D):
            for a in range(3):
                key |= ((cell[:, a] >> b) & 1) << (3 * b + a)
        assert (orc.locate(t, pts, 0) == key).all()


def test_locate_specific_edges(orc, tb):
    anchor, depth = tb.uniform_leaves(1)  # 8 octants, Morton order
    t = orc.Tree(anchor, depth, 1, 1)
    P = np.array([
        [0.5, 0.5, 0.5],            # shared corner -> upper octant (7)
        [1.0, 1.0, 1.0],            # FreeSpace: c == 1.0 shifts into the last cell
        [np.nextafter(0.5, 0), 0.5, 0.0],
        [-0.0, -0.0, -0.0],
        [-1e-17, 0.2, 0.2],
    ])
    assert list(orc.locate(t, P, 0)) == [7, 7, 2, 0, 0]
    # periodic: 1.0 wraps to 0, -1e-17 + 1 rounds to 1.0 and wraps again to 0
    assert list(orc.locate(t, P, 1)) == [7, 0, 2, 0, 0]


def test_rejects_unsorted_leaves(orc, tb):
    anchor, depth = tb.uniform_leaves(1)
    t = orc.Tree(anchor[::-1], depth[::-1], 1, 1)
    with pytest.raises(RuntimeError):
        orc.locate(t, np.zeros((1, 3)), 0)
