"""Parity of the HIP path (libtbx.so through its C ABI) with the FP64 CPU
oracle on the same inputs.

Bars: leaf ids bit-exact; TBX_EXACT values, departure points and grid
points bit-identical; TBX_FP64 (fused multiply-add) within 1e-12 relative
L-inf; TBX_FP32 contraction within 2e-6 relative L-inf of the coefficient
scale (single-precision rounding of ~Nc terms).
"""
import math

import numpy as np
import pytest

from conftest import edge_points, random_tree, wrap_periodic

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-6


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def dev(torch, a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def pair(tb, orc, anchor, depth, q, dof, coeff):
    return tb.Tree(anchor, depth, q, dof, coeff), orc.Tree(anchor, depth, q, dof, coeff)


def rel(a, b):
    b = np.asarray(b)
    return np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("max_depth", [6, 8])
def test_locate_bit_exact(torch, tb, orc, seed, bc, max_depth):
    # max depth 6: the directory names the leaf; 8: directory + search
    anchor, depth = random_tree(seed, max_depth=max_depth, p_split=0.3)
    g, o = pair(tb, orc, anchor, depth, 2, 1, None)
    rng = np.random.default_rng(seed)
    pts = rng.random((40000, 3)) * (1.3 if bc else 1.0) - (0.15 if bc else 0.0)
    j = rng.integers(0, len(depth), 5000)
    h = 0.5 ** depth[j].astype(np.float64)
    corner = anchor[j] * h[:, None]
    pts = np.concatenate([pts, edge_points(), corner,
                          corner + h[:, None] * rng.integers(0, 2, (5000, 3))])
    got = g.locate(dev(torch, pts), bc).cpu().numpy()
    assert (got == orc.locate(o, pts, bc)).all()


def test_locate_full_size_property(torch, tb):
    # 2^24 points on the 262144-leaf uniform tree: every point lies in its
    # leaf's half-open box (size-independent)
    anchor, depth = tb.uniform_leaves(6)
    g = tb.Tree(anchor, depth, 1, 1)
    torch.manual_seed(0)
    pts = torch.rand((1 << 24, 3), dtype=torch.float64, device="cuda")
    leaf = g.locate(pts, 0).long()
    lo = dev(torch, anchor, torch.float64)[leaf] / 64.0
    assert bool(((pts >= lo) & (pts < lo + 1.0 / 64)).all())


@pytest.mark.parametrize("q,dof,bc", [(1, 1, 0), (4, 3, 0), (6, 1, 1), (8, 3, 1), (14, 1, 0),
                                      (16, 2, 0)])
def test_eval_modes_against_oracle(torch, tb, orc, q, dof, bc):
    anchor, depth = random_tree(q, max_depth=4)
    rng = np.random.default_rng(q)
    coeff = rng.standard_normal((len(depth), dof, tb.ncoef(q))) / (1 + np.arange(tb.ncoef(q)))
    g, o = pair(tb, orc, anchor, depth, q, dof, coeff)
    pts = rng.random((20000, 3)) * 1.1 - 0.05
    pts = np.concatenate([pts, edge_points()])
    ref, wrapped = orc.eval_tree(o, pts, bc)
    exact = g.eval(dev(torch, pts), bc, "exact").cpu().numpy()
    assert np.array_equal(exact, ref), rel(exact, ref)
    fma = g.eval(dev(torch, pts), bc, "fp64").cpu().numpy()
    assert rel(fma, ref) <= FP64_TOL
    f32 = g.eval(dev(torch, pts), bc, "fp32").cpu().numpy()
    scale = np.abs(coeff).sum(axis=2).max()
    assert np.abs(f32 - ref).max() <= FP32_TOL * scale
    # EvalTree's in-place wrap, on request
    p = dev(torch, pts)
    g.eval(p, bc, "exact", write_back=True)
    assert np.array_equal(p.cpu().numpy(), wrapped if bc else pts)


def test_eval_every_degree_bitwise(torch, tb, orc):
    anchor, depth = random_tree(99, max_depth=3)
    rng = np.random.default_rng(5)
    pts = rng.random((3000, 3))
    for q in range(1, 17):
        coeff = rng.standard_normal((len(depth), 3, tb.ncoef(q)))
        g, o = pair(tb, orc, anchor, depth, q, 3, coeff)
        ref, _ = orc.eval_tree(o, pts, 0)
        assert np.array_equal(g.eval(dev(torch, pts), 0, "exact").cpu().numpy(), ref), q


def test_eval_empty_and_errors(torch, tb):
    anchor, depth = tb.uniform_leaves(1)
    g = tb.Tree(anchor, depth, 3, 1)
    out = g.eval(torch.empty((0, 3), dtype=torch.float64, device="cuda"))
    assert out.shape == (0, 1)
    with pytest.raises(tb.TbxError):
        g.eval(torch.zeros((4, 3), dtype=torch.float64, device="cuda"), bc=7)


def rotation_trees(tb, orc, depth, q, omega):
    from tbslas_amd import fields

    anchor, dep = tb.uniform_leaves(depth)
    cf = tb.fit_function(lambda p: fields.rotation(p, omega), anchor, dep, q, 3)
    return pair(tb, orc, anchor, dep, q, 3, cf)


def tg_trees(tb, orc, anchor, dep, q):
    from tbslas_amd import fields

    cf = tb.fit_function(fields.taylor_green_2d3, anchor, dep, q, 3)
    return pair(tb, orc, anchor, dep, q, 3, cf)


@pytest.mark.parametrize("nrk,bc,timed", [(1, 0, False), (3, 1, True), (2, 0, True)])
def test_traj_bit_identical(torch, tb, orc, nrk, bc, timed):
    anchor, depth = random_tree(7, max_depth=4)
    g, o = tg_trees(tb, orc, anchor, depth, 6)
    rng = np.random.default_rng(nrk)
    x = rng.random((30000, 3))
    fac = (lambda t: math.cos(math.pi * t)) if timed else None
    ref = orc.traj_rk2(o, x, 0.7, 0.6, nrk, bc, g=fac)
    got = tb.traj_rk2(g, dev(torch, x), 0.7, 0.6, nrk, bc, "exact", g=fac).cpu().numpy()
    assert np.array_equal(got, ref)
    got = tb.traj_rk2(g, dev(torch, x), 0.7, 0.6, nrk, bc, "fp64", g=fac).cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-13
    # in place, as ComputeTrajRK2 runs (xsol aliases xinit)
    xi = dev(torch, x)
    tb.traj_rk2(g, xi, 0.7, 0.6, nrk, bc, "exact", g=fac, out=xi)
    assert np.array_equal(xi.cpu().numpy(), ref)


def test_semilag_rk2_against_oracle(torch, tb, orc):
    from tbslas_amd import fields

    gv, ov = rotation_trees(tb, orc, 3, 6, 2 * math.pi)
    anchor, dep = tb.uniform_leaves(3)
    cf = tb.fit_function(lambda p: fields.gaussian(p, (0.7, 0.5, 0.5), 0.06), anchor, dep, 6, 1)
    gc, oc = pair(tb, orc, anchor, dep, 6, 1, cf)
    x = np.random.default_rng(1).random((20000, 3))
    ref = orc.semilag_rk2(ov, oc, x, 3, 2 * math.pi / 200, 1, 0)
    got = tb.semilag_rk2(gv, gc, dev(torch, x), 3, 2 * math.pi / 200, 1, 0, "exact").cpu().numpy()
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("q,dof", [(7, 2), (14, 3), (16, 2)])
def test_grid_points_and_refit(torch, tb, orc, q, dof):
    # q = 7: four rows per refit workgroup, groups spanning leaves; q = 14:
    # one row, values staged through LDS; q = 16: one row, loaded per thread
    anchor, depth = random_tree(12, max_depth=4 if q < 10 else 2)
    g, o = pair(tb, orc, anchor, depth, q, dof, None)
    gp = g.grid_points().cpu().numpy()
    assert np.array_equal(gp, orc.collect_grid_points(o))
    vals = np.random.default_rng(3).standard_normal((len(depth), dof, (q + 1) ** 3))
    g.set_grid_values(dev(torch, vals))
    orc.set_grid_values(o, orc.pt2coeff(q), vals)
    assert rel(g.get_coeffs(), o.coeff) <= 1e-13


@pytest.mark.parametrize("mode,bc,setup", [("exact", 0, "coarse"), ("fp64", 1, "coarse"),
                                           ("fp32", 1, "fine")])
def test_insitu_steps_against_oracle(torch, tb, orc, mode, bc, setup):
    # "coarse": random tree of depth <= 3, q = 6, a Gaussian blob; the fields
    # jump across leaf faces by their fit error (~1e-3).  "fine": uniform
    # depth 3, q = 10, a smooth periodic concentration; jumps ~1e-7.  FP32
    # moves departure points by ~1e-8, and a departure point that sits on a
    # leaf face or on the periodic boundary (a zero velocity component
    # there: Taylor-Green's v_x on x = 0) can then be evaluated on the other
    # side, changing the result by the jump (or by the field's mismatch
    # across the periodic boundary), so FP32 is held to its bar on the
    # well-resolved periodic setup.
    from tbslas_amd import fields

    if setup == "coarse":
        q = 6
        anchor, depth = random_tree(21, max_depth=3)
        fn = lambda p: fields.gaussian(p, (0.6, 0.45, 0.5), 0.1)
    else:
        q = 10
        anchor, depth = tb.uniform_leaves(3)
        tp = 2 * math.pi
        fn = lambda p: (0.5 + 0.25 * np.sin(tp * (p[:, 0] + p[:, 1])) * np.cos(tp * p[:, 2]))[:, None]
    gv, ov = tg_trees(tb, orc, anchor, depth, q)
    cf = tb.fit_function(fn, anchor, depth, q, 1)
    gc, oc = pair(tb, orc, anchor, depth, q, 1, cf)
    M = orc.pt2coeff(q)
    fac = lambda t: math.cos(math.pi * t)
    dt = 0.02
    dep = torch.empty((gc.L * gc.P, 3), dtype=torch.float64, device="cuda")
    for step in range(1, 4):
        # departure points of the oracle's step at the same state
        pts = orc.collect_grid_points(oc)
        ref_dep = orc.traj_rk2(ov, pts, step * dt, step * dt - dt, 2, bc, g=fac)
        if bc == 1:
            ref_dep = wrap_periodic(ref_dep)
        orc.semilag_insitu(ov, oc, M, step, dt, 2, bc, g=fac)
        tb.semilag_insitu(gv, gc, step, dt, 2, bc, mode, g=fac, dep=dep)
        got_dep = dep.cpu().numpy()
        if mode == "exact":
            assert np.array_equal(got_dep, ref_dep)
        elif mode == "fp64":
            assert np.abs(got_dep - ref_dep).max() <= 1e-14
        else:
            d = np.abs(got_dep - ref_dep)
            d = np.minimum(d, 1 - d)  # periodic images
            assert d.max() <= 1e-8
        err = np.abs(gc.get_coeffs() - oc.coeff).max() / np.abs(oc.coeff).max()
        assert err <= (FP64_TOL if mode != "fp32" else 1e-5), (step, err)


def test_uniform_depth4_full_size_step(torch, tb, orc):
    # the CPU-runnable workload at full size: uniform depth 4 (4096 leaves),
    # q = 8, v = 2 pi (0.5 - y, x - 0.5, 0), Gaussian blob sigma 0.06 at
    # (0.7, 0.5, 0.5), dt = 2 pi / 200; one step of 2,985,984 departure points
    from tbslas_amd import fields

    q, dt = 8, 2 * math.pi / 200
    gv, ov = rotation_trees(tb, orc, 4, q, 2 * math.pi)
    anchor, depth = tb.uniform_leaves(4)
    cf = tb.fit_function(lambda p: fields.gaussian(p, (0.7, 0.5, 0.5), 0.06), anchor, depth, q, 1)
    gc, oc = pair(tb, orc, anchor, depth, q, 1, cf)
    pts = orc.collect_grid_points(oc)
    ref_dep = orc.traj_rk2(ov, pts, dt, 0.0, 1, 0)
    orc.semilag_insitu(ov, oc, orc.pt2coeff(q), 1, dt, 1, 0)
    dep = torch.empty((gc.L * gc.P, 3), dtype=torch.float64, device="cuda")
    tb.semilag_insitu(gv, gc, 1, dt, 1, 0, "exact", dep=dep)
    assert np.array_equal(dep.cpu().numpy(), ref_dep)
    leaf_ref = orc.locate(oc, ref_dep, 0)
    assert (gc.locate(dep, 0).cpu().numpy() == leaf_ref).all()
    assert rel(gc.get_coeffs(), oc.coeff) <= FP64_TOL


def _vel_trees(tb, orc, anchor, depth, q, n, seed):
    # n distinct smooth dof-3 fields on one tree: Taylor-Green scaled and
    # shifted per tree, so every tree differs
    from tbslas_amd import fields

    out = []
    for k in range(n):
        s, o = 1.0 - 0.15 * k, 0.05 * (k + seed)
        cf = tb.fit_function(lambda p, s=s, o=o: s * fields.taylor_green_2d3(p) + o, anchor, depth,
                             q, 3)
        out.append(pair(tb, orc, anchor, depth, q, 3, cf))
    return [g for g, _ in out], [o for _, o in out]


@pytest.mark.parametrize("kind", ["extrap", "set"])
def test_velocity_functors_against_oracle(torch, tb, orc, kind):
    anchor, depth = random_tree(31, max_depth=4)
    q = 6
    n = 2 if kind == "extrap" else 4
    gs, os_ = _vel_trees(tb, orc, anchor, depth, q, n, 1)
    tt = [0.1, 0.2, 0.35, 0.5]
    if kind == "extrap":
        gvel, ovel = tb.Velocity.extrap(gs[0], gs[1]), orc.velocity(orc.VEL_EXTRAP, os_)
    else:
        gvel, ovel = tb.Velocity.set(gs, tt), orc.velocity(orc.VEL_SET, os_, tt)
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.random((20000, 3)), edge_points()])
    for bc in (0, 1):
        xx = x * 1.2 - 0.1 if bc else x
        # functor values
        if kind == "extrap":
            ref = orc.field_extrap_eval(os_[1], os_[0], xx, bc)
        else:
            ref = orc.field_set_eval(os_, tt, xx, 0.27, bc)
        got = gvel.eval(dev(torch, xx), 0.27, bc, "exact").cpu().numpy()
        assert np.array_equal(got, ref)
        got = gvel.eval(dev(torch, xx), 0.27, bc, "fp64").cpu().numpy()
        assert rel(got, ref) <= FP64_TOL
        # trajectories, both stages
        ref = orc.traj_rk2_v(ovel, xx, 0.4, 0.3, 2, bc)
        got = tb.traj_rk2(gvel, dev(torch, xx), 0.4, 0.3, 2, bc, "exact").cpu().numpy()
        assert np.array_equal(got, ref)
        got = tb.traj_rk2(gvel, dev(torch, xx), 0.4, 0.3, 2, bc, "fp64").cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-13


@pytest.mark.parametrize("kind,mode", [("extrap", "exact"), ("extrap", "fp64"), ("set", "exact"),
                                       ("set", "fp64")])
def test_insitu_velocity_functors_against_oracle(torch, tb, orc, kind, mode):
    from tbslas_amd import fields

    q = 6
    anchor, depth = random_tree(22, max_depth=3)
    n = 2 if kind == "extrap" else 4
    gs, os_ = _vel_trees(tb, orc, anchor, depth, q, n, 2)
    tt = [0.0, 0.03, 0.06, 0.09]
    if kind == "extrap":
        gvel, ovel = tb.Velocity.extrap(gs[0], gs[1]), orc.velocity(orc.VEL_EXTRAP, os_)
    else:
        gvel, ovel = tb.Velocity.set(gs, tt), orc.velocity(orc.VEL_SET, os_, tt)
    cf = tb.fit_function(lambda p: fields.gaussian(p, (0.6, 0.45, 0.5), 0.1), anchor, depth, q, 1)
    gc, oc = pair(tb, orc, anchor, depth, q, 1, cf)
    M = orc.pt2coeff(q)
    dt = 0.02
    dep = torch.empty((gc.L * gc.P, 3), dtype=torch.float64, device="cuda")
    for step in range(2, 5):
        pts = orc.collect_grid_points(oc)
        ref_dep = orc.traj_rk2_v(ovel, pts, step * dt, step * dt - dt, 2, 1)
        ref_dep = wrap_periodic(ref_dep)
        orc.semilag_insitu_v(ovel, oc, M, step, dt, 2, 1)
        tb.semilag_insitu(gvel, gc, step, dt, 2, 1, mode, dep=dep)
        if mode == "exact":
            assert np.array_equal(dep.cpu().numpy(), ref_dep)
        else:
            assert np.abs(dep.cpu().numpy() - ref_dep).max() <= 1e-14
        err = np.abs(gc.get_coeffs() - oc.coeff).max() / np.abs(oc.coeff).max()
        assert err <= FP64_TOL, (step, err)


def test_velocity_errors(torch, tb):
    anchor, depth = tb.uniform_leaves(1)
    v3 = tb.Tree(anchor, depth, 4, 3)
    v3b = tb.Tree(anchor, depth, 5, 3)
    s1 = tb.Tree(anchor, depth, 4, 1)
    x = torch.zeros((4, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(tb.TbxError):   # dof 1 velocity
        tb.traj_rk2(tb.Velocity.extrap(v3, s1), x, 0.1, 0.0, 1)
    with pytest.raises(tb.TbxError):   # degrees differ
        tb.traj_rk2(tb.Velocity.extrap(v3, v3b), x, 0.1, 0.0, 1)
    with pytest.raises(ValueError):
        tb.Velocity.set([v3, v3], [0, 1, 2, 3])


@pytest.mark.parametrize("q", [8, 5])
def test_every_step_shape_matches_oracle(torch, tb, orc, q):
    # the launch shapes (arrival order, or points regrouped by leaf inside
    # each workgroup) change which lane evaluates a point, never its value;
    # q = 5 exercises the fallback of the shapes built for q = 8 and 14 only
    from tbslas_amd import _lib, fields

    anchor, depth = random_tree(41, max_depth=3)
    gv, ov = tg_trees(tb, orc, anchor, depth, q)
    two = lambda p: np.concatenate([fields.gaussian(p, (0.55, 0.5, 0.45), 0.12),
                                    fields.gaussian(p, (0.4, 0.6, 0.5), 0.2)], axis=1)
    cf = tb.fit_function(two, anchor, depth, q, 2)
    oc = orc.Tree(anchor, depth, q, 2, cf)
    dt = 0.03
    pts = orc.collect_grid_points(oc)
    # tinit = timestep * dt, tfinal = tinit - dt (semilag.inc:34-35)
    ref_dep = orc.traj_rk2(ov, pts, 2 * dt, 2 * dt - dt, 2, 1)
    ref_dep = wrap_periodic(ref_dep)
    ref_val, _ = orc.eval_tree(oc, ref_dep, 1)
    P = (q + 1) ** 3
    ref_val = ref_val.reshape(len(depth), P, 2).transpose(0, 2, 1)
    gc = tb.Tree(anchor, depth, q, 2, cf)
    dep = torch.empty((gc.L * P, 3), dtype=torch.float64, device="cuda")
    old = _lib.get_option(_lib.OPT_STEP_SHAPE)
    v32 = {m: tb.semilag_values(gv, gc, 2, dt, 2, 1, m).cpu().numpy() for m in ("fp32", "fp32h")}
    try:
        for shape in range(11):
            _lib.set_option(_lib.OPT_STEP_SHAPE, shape)
            dep.zero_()
            vals = tb.semilag_values(gv, gc, 2, dt, 2, 1, "exact", dep=dep)
            assert np.array_equal(dep.cpu().numpy(), ref_dep), shape
            assert np.array_equal(vals.cpu().numpy(), ref_val), shape
            v64 = tb.semilag_values(gv, gc, 2, dt, 2, 1, "fp64").cpu().numpy()
            assert rel(v64, ref_val) <= FP64_TOL, shape
            # the FP32 modes: the same bits from every shape
            for m in ("fp32", "fp32h"):
                assert np.array_equal(tb.semilag_values(gv, gc, 2, dt, 2, 1, m).cpu().numpy(),
                                      v32[m]), (shape, m)
    finally:
        _lib.set_option(_lib.OPT_STEP_SHAPE, old)


def test_fp32_head64_mode(torch, tb, orc):
    # coefficients decaying 10x per total degree, like a well-resolved field:
    # the FP64 head leaves only the FP32 rounding of the degree > 1 tail,
    # bounded by a few FP32 ulps of the tail's coefficient sum; the pure
    # FP32 mode rounds the whole value
    q = 10
    anchor, depth = random_tree(51, max_depth=3)
    deg = np.array([i + j + k for i in range(q + 1) for j in range(q + 1 - i)
                    for k in range(q + 1 - i - j)])
    rng = np.random.default_rng(9)
    coeff = rng.standard_normal((len(depth), 2, tb.ncoef(q))) * 0.1 ** deg
    g, o = pair(tb, orc, anchor, depth, q, 2, coeff)
    x = np.random.default_rng(8).random((100000, 3))
    ref, _ = orc.eval_tree(o, x, 0)
    scale = np.abs(ref).max()
    e32 = np.abs(g.eval(dev(torch, x), 0, "fp32").cpu().numpy() - ref).max() / scale
    e32h = np.abs(g.eval(dev(torch, x), 0, "fp32h").cpu().numpy() - ref).max() / scale
    tail = np.abs(coeff[..., deg > 1]).sum(-1).max() / scale
    assert e32 <= FP32_TOL
    assert e32h <= 8 * 2.0 ** -24 * tail + 1e-14, (e32h, tail)
    # the tail is 3 % of the coefficient sum here: the head removes most of
    # the FP32 rounding (about 30x; 10x asserted)
    assert e32h * 10 <= e32, (e32, e32h)
    # slowly decaying coefficients still meet the FP32 bar
    coeff = rng.standard_normal(coeff.shape) / (1 + np.arange(coeff.shape[-1]))
    g.set_coeffs(coeff)
    o2 = orc.Tree(anchor, depth, q, 2, coeff)
    ref, _ = orc.eval_tree(o2, x, 0)
    assert rel(g.eval(dev(torch, x), 0, "fp32h").cpu().numpy(), ref) <= FP32_TOL


@pytest.mark.parametrize("q,dof", [(6, 3), (14, 1)])
def test_contraction_only_matches_oracle(torch, tb, orc, q, dof):
    # the stand-alone contraction: TBX_EXACT gives, at the points it reports,
    # the oracle's component values summed in order; FP64 within 1e-12
    anchor, depth = random_tree(61 + q, max_depth=3)
    rng = np.random.default_rng(q)
    cf = rng.standard_normal((len(depth), dof, tb.ncoef(q)))
    g, o = tb.Tree(anchor, depth, q, dof, cf), orc.Tree(anchor, depth, q, dof, cf)
    n, wpl = 5000, 2
    pts = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    out = g.contraction_only(n, wpl, "exact", pts=pts)
    p = pts.cpu().numpy()
    # point i lies in leaf (i // 64 // wpl) mod L
    leaf = orc.locate(o, p, 0)
    assert np.array_equal(leaf, (np.arange(n) // 64 // wpl) % len(depth))
    vals, _ = orc.eval_tree(o, p, 0)
    want = np.zeros(n)
    for c in range(dof):
        want = want + vals[:, c]
    assert np.array_equal(out.cpu().numpy(), want)
    out = g.contraction_only(n, wpl, "fp64")
    assert rel(out.cpu().numpy(), want) <= 1e-12
