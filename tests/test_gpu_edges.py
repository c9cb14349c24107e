"""Edge cases of the hot path against the oracle, bit-exact in TBX_EXACT:

* a tree that is only its root (one leaf at depth 0);
* a chain refined toward one corner down to the maximum depth 21 (leaf
  edge 2^-21), where location leans on the full 21-bit keys and the
  binary search below the directory;
* empty point sets for every point-driven entry point;
* an odd number of components (5) through the fused in-situ step.
"""
import numpy as np
import pytest

from conftest import edge_points

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def corner_chain(tb, depth: int, corner=(1e-7, 2e-7, 3e-7)):
    """Uniform depth 1, then the leaf holding `corner` split again and again
    down to `depth` (Morton order kept by the splitter)."""
    from tbslas_amd.tree import _split

    anchor, dep = tb.uniform_leaves(1)
    for _ in range(depth - 1):
        s = 0.5 ** dep.astype(np.float64)
        lo = anchor * s[:, None]
        hit = np.all((np.asarray(corner)[None, :] >= lo) & (np.asarray(corner)[None, :] < lo + s[:, None]),
                     axis=1)
        anchor, dep = _split(anchor, dep, hit)
    return anchor, dep


def test_root_only_tree(torch, tb, orc):
    anchor = np.zeros((1, 3), dtype=np.uint32)
    depth = np.zeros(1, dtype=np.uint8)
    q = 6
    rng = np.random.default_rng(1)
    vc = rng.standard_normal((1, 3, tb.ncoef(q))) * 0.05
    cc = rng.standard_normal((1, 2, tb.ncoef(q)))
    gv, ov = tb.Tree(anchor, depth, q, 3, vc), orc.Tree(anchor, depth, q, 3, vc)
    gc, oc = tb.Tree(anchor, depth, q, 2, cc), orc.Tree(anchor, depth, q, 2, cc)
    pts = np.concatenate([rng.random((2000, 3)), edge_points()])
    for bc in (0, 1):
        assert np.array_equal(gv.locate(dev(torch, pts), bc).cpu().numpy(),
                              orc.locate(ov, pts, bc))
        assert np.array_equal(gc.eval(dev(torch, pts), bc, "exact").cpu().numpy(),
                              orc.eval_tree(oc, pts, bc)[0])
        x = rng.random((500, 3))
        got = tb.traj_rk2(gv, dev(torch, x), 0.4, 0.1, 3, bc, "exact").cpu().numpy()
        assert np.array_equal(got, orc.traj_rk2(ov, x, 0.4, 0.1, 3, bc))
    tb.semilag_insitu(gv, gc, 2, 0.05, 1, 1, "exact")
    orc.semilag_insitu(ov, oc, orc.pt2coeff(q), 2, 0.05, 1, 1)
    assert np.abs(gc.get_coeffs() - oc.coeff).max() <= 1e-12 * np.abs(oc.coeff).max()


def test_maximum_depth_chain(torch, tb, orc):
    anchor, depth = corner_chain(tb, 21)
    assert depth.max() == 21 and len(depth) == 1 + 7 * 21
    q = 4
    rng = np.random.default_rng(2)
    cc = rng.standard_normal((len(depth), 1, tb.ncoef(q)))
    g, o = tb.Tree(anchor, depth, q, 1, cc), orc.Tree(anchor, depth, q, 1, cc)
    # points inside and on the faces of the deepest leaves, and around them
    h = 0.5 ** 21
    base = rng.integers(0, 4, size=(4000, 3)) * h
    jitter = rng.random((4000, 3)) * h
    pts = np.concatenate([base, base + jitter, base + h, rng.random((2000, 3)) * 64 * h,
                          edge_points()])
    for bc in (0, 1):
        lg = g.locate(dev(torch, pts), bc).cpu().numpy()
        assert np.array_equal(lg, orc.locate(o, pts, bc))
        assert np.array_equal(g.eval(dev(torch, pts), bc, "exact").cpu().numpy(),
                              orc.eval_tree(o, pts, bc)[0])
    # every deep leaf is hit by its own cell centres
    s = 0.5 ** depth.astype(np.float64)
    centres = anchor * s[:, None] + 0.5 * s[:, None]
    assert np.array_equal(g.locate(dev(torch, centres), 0).cpu().numpy(),
                          np.arange(len(depth), dtype=np.int32))


def test_empty_point_sets(torch, tb):
    anchor, depth = tb.uniform_leaves(1)
    q = 3
    vel = tb.Tree(anchor, depth, q, 3, np.zeros((8, 3, tb.ncoef(q))))
    conc = tb.Tree(anchor, depth, q, 1, np.zeros((8, 1, tb.ncoef(q))))
    e = torch.empty((0, 3), dtype=torch.float64, device="cuda")
    assert tb.traj_rk2(vel, e, 1.0, 0.9, 2).shape == (0, 3)
    assert vel.locate(e).shape == (0,)
    assert tb.semilag_rk2(vel, conc, e, 1, 0.1).shape[0] == 0
    assert tb.Velocity.tree(vel).eval(e, 0.5).shape == (0, 3)


def test_odd_component_count_insitu(torch, tb, orc):
    from conftest import random_tree
    from tbslas_amd import fields

    anchor, depth = random_tree(17, max_depth=3)
    q, dof = 5, 5
    vc = tb.fit_function(lambda p: fields.taylor_green_2d3(p) + 0.2, anchor, depth, q, 3)
    rng = np.random.default_rng(5)
    cc = rng.standard_normal((len(depth), dof, tb.ncoef(q)))
    gv, ov = tb.Tree(anchor, depth, q, 3, vc), orc.Tree(anchor, depth, q, 3, vc)
    gc, oc = tb.Tree(anchor, depth, q, dof, cc), orc.Tree(anchor, depth, q, dof, cc)
    M = orc.pt2coeff(q)
    for n in (1, 2):
        tb.semilag_insitu(gv, gc, n, 0.03, 1, 1, "exact")
        orc.semilag_insitu(ov, oc, M, n, 0.03, 1, 1)
        assert np.abs(gc.get_coeffs() - oc.coeff).max() <= 1e-12 * np.abs(oc.coeff).max(), n


def test_departure_buffer_shape_dtype_layout_checked(tb):
    # wrong dtype, row count or layout of dep raises in the wrapper; the
    # kernels never see such a buffer
    import torch

    anchor, depth = tb.uniform_leaves(1)
    vel, conc = tb.Tree(anchor, depth, 2, 3), tb.Tree(anchor, depth, 2, 1)
    n = conc.L * conc.P
    bad = [torch.zeros((n, 3), dtype=torch.float32, device="cuda"),
           torch.zeros((n - 1, 3), dtype=torch.float64, device="cuda"),
           torch.zeros((3, n), dtype=torch.float64, device="cuda").t()]
    for dep in bad:
        with pytest.raises((TypeError, ValueError)):
            tb.semilag_insitu(vel, conc, 1, 0.01, 1, 0, "fp64", dep=dep)
        with pytest.raises((TypeError, ValueError)):
            tb.semilag_values(vel, conc, 1, 0.01, 1, 0, "fp64", dep=dep)
    g = tb.Grid(anchor, depth, 4, 1)
    gv = tb.Grid(anchor, depth, 4, 3)
    with pytest.raises(ValueError):
        tb.grid_semilag_insitu(gv, g, 1, 0.01, 1, 0, "fp64",
                               dep=torch.zeros((g.L * 64 + 1, 3), dtype=torch.float64,
                                               device="cuda"))
