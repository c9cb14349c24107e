"""The first RK2 stage from the velocity tree's values at its own grid nodes
(TBX_OPT_NODE_VELOCITY): against the same step with every stage evaluated,
and against the FP64 CPU oracle.

The arrival points of a step are the concentration's grid nodes; where the
velocity leaf a node is located in has the arrival leaf's depth, the node is
one of that leaf's grid nodes and the step takes the first stage's velocity
from the expansion at the nodes.  The two paths differ only in summation
order (the expansion's three one-axis contractions against the leaf's
triangular contraction at the point), so:

* TBX_FP64: departure points within 1e-15 (absolute, in [0,1)^3) and values
  within 1e-13 relative L-inf of the evaluated path, and within 1e-12 of the
  oracle;
* TBX_FP32 and TBX_FP32_HEAD64: the expansion is FP64, so the first stage is
  more accurate than the FP32 contraction (~1e-7 relative); with a velocity
  the tree represents exactly (continuous across faces), departure points
  within 1e-8 and values within 1e-6 of the evaluated path;
* TBX_EXACT never takes it: bit-identical with the option on and off;
* every launch shape gives the same bits with the option on.

Cases: velocity and concentration on the same mixed-depth tree (most nodes
take the nodal values, nodes on faces towards a leaf of another depth are
evaluated), the velocity on a coarser uniform tree (only the concentration
leaves of the velocity's depth take them), both boundary conditions, two
substeps (the second substep evaluates), and a 16-component concentration
on the matrix-core step; the oracle comparison and the launch shapes use
the Taylor-Green velocity.
"""
import math

import numpy as np
import pytest

from conftest import random_tree, wrap_periodic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def rel(a, b):
    b = np.asarray(b)
    return np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300)


def periodic_diff(a, b):
    d = np.abs(a - b)
    return np.minimum(d, 1 - d).max()


def drift(p):
    """A drift of total degree 2, (0.3 + 0.1 y(1-y), -0.2 + 0.1 z(1-z),
    0.25 + 0.1 x(1-x)): represented exactly at q >= 2, so continuous across
    leaf faces and (its components equal at 0 and 1) across the periodic
    boundary, and without a vanishing component.  The FP32 steps then differ
    from the FP64 ones by rounding alone.  (Taylor-Green's fit jumps across
    faces by ~1e-5 at q = 8, so a midpoint that FP32 rounding moves across a
    face changes the departure point by tau times that jump; and under free
    space its v_x vanishes on x = 0, where a first stage rounded to either
    sign puts the midpoint inside or outside the domain.)"""
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    return np.stack([0.3 + 0.1 * y * (1 - y), -0.2 + 0.1 * z * (1 - z),
                     0.25 + 0.1 * x * (1 - x)], axis=1)


def _setup(tb, q, vel_tree, dof=1, seed=11, field="drift"):
    from tbslas_amd import fields

    anchor, depth = random_tree(seed, max_depth=3, min_depth=2)
    if vel_tree == "same":
        va, vd = anchor, depth
    else:
        va, vd = tb.uniform_leaves(2)
    vcf = tb.fit_function(drift if field == "drift" else fields.taylor_green_2d3, va, vd, q, 3)
    if dof == 1:
        fn = lambda p: fields.gaussian(p, (0.55, 0.5, 0.45), 0.12)
    else:
        fn = lambda p: np.concatenate(
            [fields.gaussian(p, (0.3 + 0.025 * k, 0.5, 0.5), 0.1 + 0.01 * k) for k in range(dof)],
            axis=1)
    ccf = tb.fit_function(fn, anchor, depth, q, dof)
    return (va, vd, vcf), (anchor, depth, ccf)


def _run(torch, tb, vel, conc, mode, bc, nrk, on, dt=0.02):
    from tbslas_amd import _lib

    old = _lib.get_option(_lib.OPT_NODE_VELOCITY)
    _lib.set_option(_lib.OPT_NODE_VELOCITY, 1 if on else 0)
    try:
        dep = torch.empty((conc.L * conc.P, 3), dtype=torch.float64, device="cuda")
        fac = lambda t: math.cos(math.pi * t)
        vals = tb.semilag_values(vel, conc, 2, dt, nrk, bc, mode, g=fac, dep=dep)
        return vals.cpu().numpy(), dep.cpu().numpy()
    finally:
        _lib.set_option(_lib.OPT_NODE_VELOCITY, old)


@pytest.mark.parametrize("vel_tree", ["same", "coarse"])
@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("nrk", [1, 2])
def test_nodal_first_stage_matches_evaluated(torch, tb, vel_tree, bc, nrk):
    q = 8
    (va, vd, vcf), (anchor, depth, ccf) = _setup(tb, q, vel_tree)
    vel = tb.Tree(va, vd, q, 3, vcf)
    conc = tb.Tree(anchor, depth, q, 1, ccf)
    for mode, dtol, vtol in (("fp64", 1e-15, 1e-13), ("fp32", 1e-8, 1e-6),
                             ("fp32h", 1e-8, 1e-6)):
        v1, d1 = _run(torch, tb, vel, conc, mode, bc, nrk, True)
        v0, d0 = _run(torch, tb, vel, conc, mode, bc, nrk, False)
        dd = periodic_diff(d1, d0) if bc else np.abs(d1 - d0).max()
        assert dd <= dtol, (mode, dd)
        assert rel(v1, v0) <= vtol, (mode, rel(v1, v0))
    e1, f1 = _run(torch, tb, vel, conc, "exact", bc, nrk, True)
    e0, f0 = _run(torch, tb, vel, conc, "exact", bc, nrk, False)
    assert np.array_equal(e1, e0) and np.array_equal(f1, f0)


@pytest.mark.parametrize("vel_tree", ["same", "coarse"])
def test_nodal_first_stage_against_oracle(torch, tb, orc, vel_tree):
    q = 8
    (va, vd, vcf), (anchor, depth, ccf) = _setup(tb, q, vel_tree, seed=12, field="tg")
    vel = tb.Tree(va, vd, q, 3, vcf)
    conc = tb.Tree(anchor, depth, q, 1, ccf)
    ov, oc = orc.Tree(va, vd, q, 3, vcf), orc.Tree(anchor, depth, q, 1, ccf)
    dt = 0.02
    fac = lambda t: math.cos(math.pi * t)
    pts = orc.collect_grid_points(oc)
    ref_dep = wrap_periodic(orc.traj_rk2(ov, pts, 2 * dt, dt, 1, 1, g=fac))
    ref_val, _ = orc.eval_tree(oc, ref_dep, 1)
    P = (q + 1) ** 3
    ref_val = ref_val.reshape(len(depth), P, 1).transpose(0, 2, 1)
    v, d = _run(torch, tb, vel, conc, "fp64", 1, 1, True, dt=dt)
    assert periodic_diff(d, ref_dep) <= 1e-14
    assert rel(v, ref_val) <= 1e-12


def test_every_shape_same_bits_with_nodal_first_stage(torch, tb):
    from tbslas_amd import _lib

    q = 8
    (va, vd, vcf), (anchor, depth, ccf) = _setup(tb, q, "same", seed=13, field="tg")
    vel = tb.Tree(va, vd, q, 3, vcf)
    conc = tb.Tree(anchor, depth, q, 1, ccf)
    old = _lib.get_option(_lib.OPT_STEP_SHAPE)
    try:
        ref = {}
        for shape in range(11):
            _lib.set_option(_lib.OPT_STEP_SHAPE, shape)
            for m in ("fp64", "fp32", "fp32h"):
                v, d = _run(torch, tb, vel, conc, m, 1, 1, True)
                if m not in ref:
                    ref[m] = (v, d)
                assert np.array_equal(v, ref[m][0]) and np.array_equal(d, ref[m][1]), (shape, m)
    finally:
        _lib.set_option(_lib.OPT_STEP_SHAPE, old)


@pytest.mark.parametrize("mode", ["fp64", "fp32"])
def test_nodal_first_stage_on_matrix_core_step(torch, tb, mode):
    q = 8
    (va, vd, vcf), (anchor, depth, ccf) = _setup(tb, q, "same", dof=16, seed=14)
    vel = tb.Tree(va, vd, q, 3, vcf)
    conc = tb.Tree(anchor, depth, q, 16, ccf)
    v1, d1 = _run(torch, tb, vel, conc, mode, 1, 1, True)
    v0, d0 = _run(torch, tb, vel, conc, mode, 1, 1, False)
    assert periodic_diff(d1, d0) <= (1e-15 if mode == "fp64" else 1e-8)
    assert rel(v1, v0) <= (1e-13 if mode == "fp64" else 1e-6)


def test_node_velocity_option_checked(tb):
    from tbslas_amd import _lib

    with pytest.raises(tb.TbxError):
        _lib.set_option(_lib.OPT_NODE_VELOCITY, 2)
    assert _lib.get_option(_lib.OPT_NODE_VELOCITY) == 1


@pytest.mark.parametrize("q,dof", [(3, 1), (8, 3), (14, 2)])
def test_grid_values_against_oracle(torch, tb, orc, q, dof):
    # each leaf's own polynomial at its grid nodes (upper-face nodes
    # included), against the oracle's per-leaf Chebyshev evaluation
    anchor, depth = random_tree(15, max_depth=3)
    nc = (q + 1) * (q + 2) * (q + 3) // 6
    cf = np.random.default_rng(q).standard_normal((len(depth), dof, nc))
    t = tb.Tree(anchor, depth, q, dof, cf)
    o = orc.Tree(anchor, depth, q, dof, cf)
    P = (q + 1) ** 3
    pts = orc.collect_grid_points(o).reshape(len(depth), P, 3)
    ref = orc.eval_in_leaf(cf, anchor, depth, pts, q)
    got = t.grid_values().cpu().numpy()
    assert got.shape == ref.shape
    assert rel(got, ref) <= 1e-13, rel(got, ref)
    # the FP32 variant: the same values rounded once
    g32 = t.grid_values(torch.empty((t.L, dof, P), dtype=torch.float32, device="cuda"))
    assert np.array_equal(g32.cpu().numpy(), got.astype(np.float32))
    # the refit of its own grid values gives the coefficients back
    t.set_grid_values(t.grid_values())
    assert rel(t.get_coeffs(), cf) <= 1e-12


@pytest.mark.parametrize("bc", [0, 1])
def test_single_leaf_tree(torch, tb, bc):
    # depth 0: the leaf is its own periodic neighbour, so its upper-face
    # nodes (wrapped to 0) must not read the node at the upper face; a
    # velocity that differs across the periodic boundary shows it
    q = 8
    anchor, depth = tb.uniform_leaves(0)
    vcf = tb.fit_function(lambda p: drift(p) + 0.1 * p, anchor, depth, q, 3)
    ccf = tb.fit_function(lambda p: np.exp(-((p - 0.4) ** 2).sum(1, keepdims=True) / 0.05),
                          anchor, depth, q, 1)
    vel = tb.Tree(anchor, depth, q, 3, vcf)
    conc = tb.Tree(anchor, depth, q, 1, ccf)
    v1, d1 = _run(torch, tb, vel, conc, "fp64", bc, 1, True)
    v0, d0 = _run(torch, tb, vel, conc, "fp64", bc, 1, False)
    dd = periodic_diff(d1, d0) if bc else np.abs(d1 - d0).max()
    assert dd <= 1e-15, dd
    assert rel(v1, v0) <= 1e-13


def test_nodal_first_stage_extrapolated_velocity(torch, tb):
    # the extrapolating functor's first stage is its current tree alone
    # (traj.inc:80): it reads that tree's nodal values; the second stage
    # extrapolates as before
    q = 8
    (va, vd, vcf), (anchor, depth, ccf) = _setup(tb, q, "same", seed=16)
    cur = tb.Tree(va, vd, q, 3, vcf)
    prev = tb.Tree(va, vd, q, 3, 0.9 * vcf)
    conc = tb.Tree(anchor, depth, q, 1, ccf)
    vel = tb.Velocity.extrap(cur, prev)
    from tbslas_amd import _lib

    out = {}
    for on in (True, False):
        old = _lib.get_option(_lib.OPT_NODE_VELOCITY)
        _lib.set_option(_lib.OPT_NODE_VELOCITY, 1 if on else 0)
        try:
            for mode in ("fp64", "fp32"):
                dep = torch.empty((conc.L * conc.P, 3), dtype=torch.float64, device="cuda")
                v = tb.semilag_values(vel, conc, 2, 0.02, 1, 1, mode, dep=dep)
                out[(on, mode)] = (v.cpu().numpy(), dep.cpu().numpy())
        finally:
            _lib.set_option(_lib.OPT_NODE_VELOCITY, old)
    for mode, dtol, vtol in (("fp64", 1e-15, 1e-13), ("fp32", 1e-8, 1e-6)):
        (v1, d1), (v0, d0) = out[(True, mode)], out[(False, mode)]
        assert periodic_diff(d1, d0) <= dtol, mode
        assert rel(v1, v0) <= vtol, mode
