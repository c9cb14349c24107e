"""Grid (tricubic) leaves on the device against the oracle's restatement of
fast_interp (tree_functor.h:89-153) applied per leaf.

Bars: TBX_EXACT bit-identical values, departure points and samples; TBX_FP64
within 1e-13 relative; TBX_FP32 within 2e-6 relative (single-precision
samples and sums).  At the full size (262,144 leaves of 8^3 samples,
134,217,728 departure points): a constant field stays constant and a linear
field is advected exactly (both reproduced by the Lagrange stencil).
"""
import math

import numpy as np
import pytest

from conftest import edge_points, random_tree, wrap_periodic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rel(a, b):
    return np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("n,bc", [(8, 0), (8, 1), (5, 0), (12, 1)])
def test_grid_eval_against_oracle(torch, tb, orc, n, bc):
    anchor, depth = random_tree(70 + n, max_depth=4)
    rng = np.random.default_rng(n)
    samples = rng.standard_normal((len(depth), 2, n ** 3))
    g = tb.Grid(anchor, depth, n, 2, samples)
    o = orc.Grid(anchor, depth, n, 2, samples)
    x = np.concatenate([rng.random((30000, 3)) * (1.2 if bc else 1.0) - (0.1 if bc else 0.0),
                        edge_points()])
    ref = orc.grid_eval(o, x, bc)
    assert np.array_equal(g.eval(dev(torch, x), bc, "exact").cpu().numpy(), ref)
    assert rel(g.eval(dev(torch, x), bc, "fp64").cpu().numpy(), ref) <= 1e-13
    assert rel(g.eval(dev(torch, x), bc, "fp32").cpu().numpy(), ref) <= 2e-6


@pytest.mark.parametrize("n,bc,timed", [(8, 1, True), (6, 0, False)])
def test_grid_step_against_oracle(torch, tb, orc, n, bc, timed):
    from tbslas_amd import fields

    anchor, depth = random_tree(80 + n, max_depth=3)
    vs = tb.sample_function(fields.taylor_green_2d3, anchor, depth, n, 3)
    cs = tb.sample_function(lambda p: fields.gaussian(p, (0.55, 0.5, 0.5), 0.12), anchor, depth,
                            n, 1)
    fac = (lambda t: math.cos(math.pi * t)) if timed else None
    gv, gc = tb.Grid(anchor, depth, n, 3, vs), tb.Grid(anchor, depth, n, 1, cs)
    g64 = tb.Grid(anchor, depth, n, 1, cs)
    ov, oc = orc.Grid(anchor, depth, n, 3, vs), orc.Grid(anchor, depth, n, 1, cs)
    dt = 0.03
    pts = orc.grid_points(oc)
    dep = torch.empty((len(pts), 3), dtype=torch.float64, device="cuda")
    for step in range(1, 4):
        ref_dep = orc.grid_traj_rk2(ov, pts, step * dt, step * dt - dt, 2, bc, g=fac)
        if bc == 1:
            ref_dep = wrap_periodic(ref_dep)
        orc.grid_semilag_insitu(ov, oc, step, dt, 2, bc, g=fac)
        tb.grid_semilag_insitu(gv, gc, step, dt, 2, bc, "exact", g=fac, dep=dep)
        tb.grid_semilag_insitu(gv, g64, step, dt, 2, bc, "fp64", g=fac)
        assert np.array_equal(dep.cpu().numpy(), ref_dep), step
        assert np.array_equal(gc.get_samples(), oc.samples), step
        assert rel(g64.get_samples(), oc.samples) <= 1e-12, step


def test_tricubic_full_size_step_properties(torch, tb):
    # the large tricubic workload: uniform depth 6, 8^3 samples per leaf, Lamb-Oseen pair
    # (core 0.05, centres (0.4, 0.5), (0.6, 0.5)), dt = 5e-4, periodic.
    # The Lagrange stencil reproduces constants and linear functions, so a
    # constant stays constant and c = x + 2y - z moves with the departure
    # points exactly.
    from tbslas_amd import fields

    anchor, depth = tb.uniform_leaves(6)
    n = 8
    vel = tb.Grid(anchor, depth, n, 3, tb.sample_function(fields.lamb_oseen_pair, anchor, depth,
                                                          n, 3))
    lin = lambda p: (p[:, 0] + 2 * p[:, 1] - p[:, 2])[:, None]
    pts = tb.grid.grid_points(anchor, depth, n)
    for mode, tol in (("fp64", 1e-13), ("fp32", 2e-6)):
        conc = tb.Grid(anchor, depth, n, 1, np.ones((len(depth), 1, n ** 3)))
        tb.grid_semilag_insitu(vel, conc, 1, 5e-4, 1, 1, mode)
        s = conc.get_samples()
        assert np.abs(s - 1.0).max() <= tol, mode
    conc = tb.Grid(anchor, depth, n, 1, tb.sample_function(lin, anchor, depth, n, 1))
    dep = torch.empty((len(pts), 3), dtype=torch.float64, device="cuda")
    tb.grid_semilag_insitu(vel, conc, 1, 5e-4, 1, 0, "fp64", dep=dep)
    d = dep.cpu().numpy()
    want = lin(d)[:, 0].reshape(len(depth), n ** 3)
    inside = np.all((d >= 0) & (d <= 1), axis=1).reshape(len(depth), n ** 3)
    got = conc.get_samples()[:, 0, :]
    assert inside.mean() > 0.99
    assert np.abs(got[inside] - want[inside]).max() <= 1e-12


@pytest.mark.parametrize("bc", [0, 1])
def test_grid_step_lds_blocks_against_oracle(torch, tb, orc, bc):
    # a uniform tree of 8^3-sample leaves shared by velocity and concentration
    # runs the sibling-block kernel (samples staged in LDS; TBX_OPT_GRID_LDS 2
    # takes it in every mode): TBX_EXACT bit-identical with the oracle, and
    # every mode bit-identical with the global-memory kernel (option 0)
    from tbslas_amd import _lib, fields

    n = 8
    anchor, depth = tb.uniform_leaves(3)
    vs = tb.sample_function(lambda p: fields.taylor_green_2d3(p) + 0.3, anchor, depth, n, 3)
    cs = tb.sample_function(lambda p: fields.gaussian(p, (0.55, 0.5, 0.45), 0.1), anchor, depth,
                            n, 1)
    fac = lambda t: math.cos(math.pi * t)
    gv = tb.Grid(anchor, depth, n, 3, vs)
    ov, oc = orc.Grid(anchor, depth, n, 3, vs), orc.Grid(anchor, depth, n, 1, cs)
    dt = 0.02
    pts = orc.grid_points(oc)
    ref_dep = orc.grid_traj_rk2(ov, pts, dt, 0.0, 2, bc, g=fac)
    if bc == 1:
        ref_dep = wrap_periodic(ref_dep)
    orc.grid_semilag_insitu(ov, oc, 1, dt, 2, bc, g=fac)
    out = {}
    try:
        for lds in (2, 0):
            _lib.set_option(_lib.OPT_GRID_LDS, lds)
            for mode in ("exact", "fp64", "fp32"):
                gc = tb.Grid(anchor, depth, n, 1, cs)
                dep = torch.empty((len(pts), 3), dtype=torch.float64, device="cuda")
                tb.grid_semilag_insitu(gv, gc, 1, dt, 2, bc, mode, g=fac, dep=dep)
                out[(lds, mode)] = (dep.cpu().numpy(), gc.get_samples())
    finally:
        _lib.set_option(_lib.OPT_GRID_LDS, 1)
    assert np.array_equal(out[(2, "exact")][0], ref_dep)
    assert np.array_equal(out[(2, "exact")][1], oc.samples)
    for mode in ("exact", "fp64", "fp32"):
        assert np.array_equal(out[(2, mode)][0], out[(0, mode)][0]), mode
        assert np.array_equal(out[(2, mode)][1], out[(0, mode)][1]), mode
