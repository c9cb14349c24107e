import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtbx on cuda:0)")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def tb():
    import tbslas_amd

    tbslas_amd.lib()
    return tbslas_amd


def random_tree(seed: int, max_depth: int = 4, p_split: float = 0.35, min_depth: int = 1):
    """Random complete octree in Morton order (leaf splits drawn from seed)."""
    from tbslas_amd.tree import _split, uniform_leaves

    rng = np.random.default_rng(seed)
    anchor, depth = uniform_leaves(min_depth)
    for _ in range(max_depth - min_depth):
        mask = (rng.random(len(depth)) < p_split) & (depth < max_depth)
        if not mask.any():
            break
        anchor, depth = _split(anchor, depth, mask)
    return anchor, depth


def brute_locate(anchor, depth, pts, bc):
    """Leaf whose half-open box holds each point, after the documented wrap
    (one +-1 under periodic) and clamp (into [0, 1 - 2^-52], NaN -> 0)."""
    p = np.array(pts, dtype=np.float64, copy=True).reshape(-1, 3)
    if bc == 1:
        p = np.where(p < 0.0, p + 1.0, p)
        p = np.where(p >= 1.0, p - 1.0, p)
    p = np.where(np.isnan(p), 0.0, p)
    p = np.clip(p, 0.0, np.nextafter(1.0, 0.0))
    lo = anchor.astype(np.float64) * (0.5 ** depth.astype(np.float64))[:, None]
    hi = lo + (0.5 ** depth.astype(np.float64))[:, None]
    out = np.empty(len(p), dtype=np.int32)
    for s in range(0, len(p), 2048):
        q = p[s:s + 2048]
        inside = np.all((q[:, None, :] >= lo[None]) & (q[:, None, :] < hi[None]), axis=2)
        assert (inside.sum(axis=1) == 1).all()
        out[s:s + 2048] = np.argmax(inside, axis=1)
    return out


def edge_points():
    """Coordinates the reference's leaf location is sensitive to."""
    one_m = np.nextafter(1.0, 0.0)
    vals = [0.0, -0.0, 0.5, 0.25, 0.75, 0.125, 1.0, one_m, 1e-300, -1e-17, 1.5, -0.5, 2.5,
            0.5 - 2.0 ** -53, 0.5 + 2.0 ** -53, 0.0625, 0.9375]
    rng = np.random.default_rng(7)
    pts = [[a, b, c] for a in vals for b in (0.5, 0.3, 1.0) for c in (0.0, 0.7)]
    pts += [[v, v, v] for v in vals]
    pts += [list(rng.choice(vals, 3)) for _ in range(300)]
    return np.array(pts, dtype=np.float64)


def wrap_periodic(x):
    """The reference's periodic wrap, two sequential tests
    (tree_functor.h:445-446): c < 0 -> c + 1, then c >= 1 -> c - 1, so a
    tiny negative c whose c + 1 rounds to 1.0 ends at 0.0."""
    x = np.where(x < 0, x + 1, x)
    return np.where(x >= 1, x - 1, x)
