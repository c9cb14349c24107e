"""Many tracers carried by one trajectory set (the 64-tracer workload at a reduced size): a
concentration tree with 64 components, one per tracer, each a seeded
Gaussian blob (centres in [0.3, 0.7]^3, sigma in [0.03, 0.08], SplitMix64
seed 3), advected by the rigid rotation of the CPU workload.

Bars: departure points bit-identical with the oracle; TBX_EXACT values of
all 64 tracers bit-identical with the oracle's EvalTree at those points and
with 64 single-tracer runs; the refit within 1e-12 of the oracle's in-situ
step; TBX_FP32 within 2e-6 relative.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def test_64_tracers_share_one_trajectory_set(torch, tb, orc):
    from tbslas_amd import fields

    q, T = 8, 64
    anchor, depth = tb.uniform_leaves(3)
    omega = 2 * math.pi
    vcf = tb.fit_function(lambda p: fields.rotation(p, omega), anchor, depth, q, 3)
    centres, sigma = fields.tracer_blobs(T, 3)
    ccf = tb.fit_function(lambda p: fields.gaussian_blobs(p, centres, sigma), anchor, depth, q, T)
    vel = tb.Tree(anchor, depth, q, 3, vcf)
    conc = tb.Tree(anchor, depth, q, T, ccf)
    ov, oc = orc.Tree(anchor, depth, q, 3, vcf), orc.Tree(anchor, depth, q, T, ccf)
    dt = 2 * math.pi / 200 / omega
    P = (q + 1) ** 3
    L = len(depth)
    dep = torch.empty((L * P, 3), dtype=torch.float64, device="cuda")
    vals = tb.semilag_values(vel, conc, 1, dt, 1, 0, "exact", dep=dep).cpu().numpy()
    pts = orc.collect_grid_points(oc)
    ref_dep = orc.traj_rk2(ov, pts, dt, dt - dt, 1, 0)
    assert np.array_equal(dep.cpu().numpy(), ref_dep)
    ref, _ = orc.eval_tree(oc, ref_dep, 0)
    assert np.array_equal(vals, ref.reshape(L, P, T).transpose(0, 2, 1))
    for k in (0, 17, 63):
        one = tb.Tree(anchor, depth, q, 1, ccf[:, k:k + 1])
        v1 = tb.semilag_values(vel, one, 1, dt, 1, 0, "exact").cpu().numpy()
        assert np.array_equal(v1[:, 0, :], vals[:, k, :])
    v32 = tb.semilag_values(vel, conc, 1, dt, 1, 0, "fp32").cpu().numpy()
    assert np.abs(v32 - vals).max() <= 2e-6 * np.abs(vals).max()
    # the in-situ step (refit of all 64 tracers) against the oracle's
    tb.semilag_insitu(vel, conc, 1, dt, 1, 0, "exact")
    orc.semilag_insitu(ov, oc, orc.pt2coeff(q), 1, dt, 1, 0)
    got = conc.get_coeffs()
    assert np.abs(got - oc.coeff).max() <= 1e-12 * np.abs(oc.coeff).max()


def _small_problems(tb, n=12, seed=33):
    # a reduced 256-problem batch: depths 2..3, degrees 6/8/10, rigid rotations about
    # random axes through the centre, a Gaussian blob each
    from tbslas_amd import fields

    u = fields.splitmix64(seed, 4 * n).reshape(n, 4)
    probs = []
    for p in range(n):
        depth_u = 2 + p % 2
        q = (6, 8, 10)[p % 3]
        z = 2 * u[p, 0] - 1
        phi = 2 * math.pi * u[p, 1]
        axis = (math.sqrt(1 - z * z) * math.cos(phi), math.sqrt(1 - z * z) * math.sin(phi), z)
        anchor, depth = tb.uniform_leaves(depth_u)
        vcf = tb.fit_function(lambda x, a=axis: fields.rotation_axis(x, a, 2 * math.pi), anchor,
                              depth, q, 3)
        ccf = tb.fit_function(lambda x: fields.gaussian(x, (0.6, 0.5, 0.5), 0.12), anchor, depth,
                              q, 1)
        probs.append((anchor, depth, q, vcf, ccf, 0.005 * (1 + p % 3)))
    return probs


@pytest.mark.parametrize("nodal", [1, 0])
def test_batch_matches_per_problem_steps(torch, tb, nodal):
    from tbslas_amd.batch import CONFIGS, Batch

    probs = _small_problems(tb)
    vel = [tb.Tree(a, d, q, 3, v) for a, d, q, v, c, dt in probs]
    conc = [tb.Tree(a, d, q, 1, c) for a, d, q, v, c, dt in probs]
    dts = [dt for *_, dt in probs]
    # per-problem references in the same arithmetic (FP32 contraction,
    # FP64 refit), three steps (timestep 1: a stationary velocity makes
    # every step the same map).  Against FP64 the difference would be the
    # face flips of FP32 departure points on these coarse problems (leaf
    # faces the rotation axis crosses, fit jumps ~1e-4), not rounding.
    # TBX_OPT_NODE_VELOCITY is the same for both: on (the default), the
    # batch, made while it is on, and the references take the first RK2
    # stage from the velocities' FP32 nodal values; off, both evaluate it.
    from tbslas_amd import _lib

    refs = []
    results = []
    old = _lib.get_option(_lib.OPT_NODE_VELOCITY)
    _lib.set_option(_lib.OPT_NODE_VELOCITY, nodal)
    try:
        for (a, d, q, v, c, dt), tv in zip(probs, vel):
            t = tb.Tree(a, d, q, 1, c)
            for _ in range(3):
                tb.semilag_insitu(tv, t, 1, dt, 1, 0, "fp32")
            refs.append(t.get_coeffs())
        for cfg in range(len(CONFIGS)):
            b = Batch(vel, conc, dts, 0)
            b.set_config(cfg)
            b.step(3)  # graph replays
            results.append([b.coeffs(p) for p in range(len(probs))])
            b.close()
        b = Batch(vel, conc, dts, 0)
    finally:
        _lib.set_option(_lib.OPT_NODE_VELOCITY, old)
    for cfg in range(1, len(CONFIGS)):  # configurations change the work split only
        for p in range(len(probs)):
            assert np.array_equal(results[cfg][p], results[0][p]), (cfg, p)
    for p in range(len(probs)):  # the batch keeps its state in FP32
        err = np.abs(results[0][p] - refs[p]).max() / np.abs(refs[p]).max()
        assert err <= 1e-6, (p, err)
    ms = b.tune()
    assert (ms > 0).all() and b.config == int(np.argmin(ms))
    b.step(3)
    for p in range(len(probs)):
        assert np.array_equal(b.coeffs(p), results[0][p])


def _drift(p):
    # total degree 2, exact at q >= 2: continuous across leaf faces and the
    # periodic boundary, no vanishing component (FP32 rounding of the first
    # stage moves departure points by rounding only)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    return np.stack([0.3 + 0.1 * y * (1 - y), -0.2 + 0.1 * z * (1 - z),
                     0.25 + 0.1 * x * (1 - x)], axis=1)


def test_batch_nodal_first_stage_on_mixed_depths(torch, tb):
    # The batch's first RK2 stage from the velocities' nodal values where the
    # velocity leaf holding an arrival node has the arrival leaf's depth, and
    # evaluated elsewhere (nodes on faces towards a leaf of another depth, and
    # every node under a velocity of another depth).  Problems: velocity and
    # concentration on one mixed depth-2/3 tree, and a uniform depth-2
    # velocity under a mixed concentration; degrees 6 and 8; periodic.
    # Against per-problem FP32 steps that take the same first stage (each
    # point's result is the same in both) at 1e-6, and against the batch
    # made with TBX_OPT_NODE_VELOCITY off (the first stage evaluated in FP32:
    # rounding differences only) at 2e-6.
    from conftest import random_tree
    from tbslas_amd import _lib, fields
    from tbslas_amd.batch import Batch

    probs = []
    for p, (q, coarse) in enumerate(((6, False), (8, False), (6, True), (8, True))):
        anchor, depth = random_tree(50 + p, max_depth=3, min_depth=2)
        va, vd = tb.uniform_leaves(2) if coarse else (anchor, depth)
        vcf = tb.fit_function(_drift, va, vd, q, 3)
        ccf = tb.fit_function(lambda x: fields.gaussian(x, (0.55, 0.5, 0.45), 0.12), anchor,
                              depth, q, 1)
        probs.append(((va, vd, vcf), (anchor, depth, ccf), q, 0.004 * (1 + p % 2)))
    vel = [tb.Tree(v[0], v[1], q, 3, v[2]) for v, c, q, dt in probs]
    dts = [dt for *_, dt in probs]
    got = {}
    refs = []
    old = _lib.get_option(_lib.OPT_NODE_VELOCITY)
    try:
        for on in (1, 0):
            _lib.set_option(_lib.OPT_NODE_VELOCITY, on)
            conc = [tb.Tree(c[0], c[1], q, 1, c[2]) for v, c, q, dt in probs]
            b = Batch(vel, conc, dts, 1)
            assert b.node_velocity == bool(on)
            b.step(2)
            got[on] = [b.coeffs(p) for p in range(len(probs))]
            b.close()
        _lib.set_option(_lib.OPT_NODE_VELOCITY, 1)
        for (v, c, q, dt), tv in zip(probs, vel):
            t = tb.Tree(c[0], c[1], q, 1, c[2])
            for _ in range(2):
                tb.semilag_insitu(tv, t, 1, dt, 1, 1, "fp32")
            refs.append(t.get_coeffs())
    finally:
        _lib.set_option(_lib.OPT_NODE_VELOCITY, old)
    for p in range(len(probs)):
        scale = np.abs(refs[p]).max()
        assert np.abs(got[1][p] - refs[p]).max() <= 1e-6 * scale, p
        assert np.abs(got[1][p] - got[0][p]).max() <= 2e-6 * scale, p


@pytest.mark.parametrize("q,dof,bc", [(8, 16, 0), (6, 48, 1), (5, 80, 0), (14, 32, 1)])
def test_tracers_on_matrix_cores(torch, tb, orc, q, dof, bc):
    # components in multiples of 16 run on the matrix cores (TBX_FP64 and
    # TBX_FP32): FP64 against TBX_EXACT (the oracle's arithmetic, vector
    # path) and against the vector contraction within 1e-13 (another
    # summation order); FP32 against the FP32 vector contraction (the same
    # departure points) within 5e-6: two FP32 summation orders over up to
    # 680 terms of a slowly decaying random expansion (2.3e-6 measured at
    # q = 14, below 1e-6 at q <= 8).  The concentration is random per leaf
    # (discontinuous across faces), so FP32 is not compared with FP64: an
    # FP32 departure point within 1e-8 of a face may take the other leaf.
    # The velocity drifts (no stagnation on leaf faces).  A random adaptive
    # tree puts several leaves in many waves.
    from conftest import random_tree
    from tbslas_amd import _lib, fields

    anchor, depth = random_tree(90 + dof, max_depth=3)
    drift = np.array([0.31, 0.17, 0.23])
    vcf = tb.fit_function(lambda p: fields.taylor_green_2d3(p) + drift, anchor, depth, q, 3)
    rng = np.random.default_rng(dof)
    ccf = rng.standard_normal((len(depth), dof, tb.ncoef(q))) / (1.0 + np.arange(tb.ncoef(q)))
    vel, conc = tb.Tree(anchor, depth, q, 3, vcf), tb.Tree(anchor, depth, q, dof, ccf)
    dt = 0.02
    ref = tb.semilag_values(vel, conc, 2, dt, 1, bc, "exact").cpu().numpy()
    scale = np.abs(ref).max()
    got = {}
    try:
        for sw in (1, 0):
            _lib.set_option(_lib.OPT_TRACER_MFMA, sw)
            for mode in ("fp64", "fp32"):
                got[(sw, mode)] = tb.semilag_values(vel, conc, 2, dt, 1, bc, mode).cpu().numpy()
    finally:
        _lib.set_option(_lib.OPT_TRACER_MFMA, 1)
    assert np.abs(got[(1, "fp64")] - ref).max() <= 1e-13 * scale
    assert np.abs(got[(1, "fp64")] - got[(0, "fp64")]).max() <= 1e-13 * scale
    assert np.abs(got[(1, "fp32")] - got[(0, "fp32")]).max() <= 5e-6 * scale
    # the in-situ step (refit of every component) through the same path
    o = tb.Tree(anchor, depth, q, dof, ccf)
    tb.semilag_insitu(vel, o, 2, dt, 1, bc, "fp64")
    e = tb.Tree(anchor, depth, q, dof, ccf)
    tb.semilag_insitu(vel, e, 2, dt, 1, bc, "exact")
    assert np.abs(o.get_coeffs() - e.get_coeffs()).max() <= 1e-12 * np.abs(e.get_coeffs()).max()


@pytest.mark.parametrize("steps", [1, 3])
def test_batch_against_oracle(torch, tb, orc, steps):
    # The persistent batch against the oracle's SolveSemilagInSitu on each
    # problem: degrees 6/8/10 at depths 2 and 3, periodic, a Gaussian blob,
    # and a constant velocity whose components are exact in FP32, so the
    # batch's FP32 velocity gives the oracle's departure points bit for bit
    # and the comparison sees the FP32 state alone: coefficients stored and
    # contracted in FP32, refit summed in FP64; bar 2e-6 of the largest
    # coefficient (the FP32 rounding of up to 286 terms).  Rotations are
    # compared with per-problem steps in the same arithmetic above: under
    # FP32 velocity a free-space trajectory from a grid point on the domain
    # boundary can end inside or outside (where the velocity is cut to 0),
    # a discontinuity of the reference scheme that rounding the oracle's own
    # velocity to FP32 shows too (1.7e-4 on a depth-3, degree-6 rotation).
    from tbslas_amd.batch import Batch

    from tbslas_amd import fields

    probs = []
    for p in range(9):
        q, depth_u = (6, 8, 10)[p % 3], 2 + (p // 3) % 2
        anchor, depth = tb.uniform_leaves(depth_u)
        vcf = np.zeros((len(depth), 3, tb.ncoef(q)))
        vcf[:, :, 0] = np.array([0.25, -0.125, 0.375]) * (0.5, 1.0, 2.0)[p % 3]
        ccf = tb.fit_function(lambda x: fields.gaussian(x, (0.6, 0.45, 0.5), 0.15), anchor,
                              depth, q, 1)
        probs.append((anchor, depth, q, vcf, ccf, 0.0037 * (1 + p % 4)))
    vel = [tb.Tree(a, d, q, 3, v) for a, d, q, v, c, dt in probs]
    conc = [tb.Tree(a, d, q, 1, c) for a, d, q, v, c, dt in probs]
    b = Batch(vel, conc, [pr[-1] for pr in probs], 1)
    b.step(steps)
    for p, (a, d, q, v, c, dt) in enumerate(probs):
        ov, oc = orc.Tree(a, d, q, 3, v), orc.Tree(a, d, q, 1, c)
        M = orc.pt2coeff(q)
        for _ in range(steps):
            orc.semilag_insitu(ov, oc, M, 1, dt, 1, 1)
        err = np.abs(b.coeffs(p) - oc.coeff).max() / np.abs(oc.coeff).max()
        assert err <= 2e-6, (p, q, err)
    b.close()


def test_batch_preserves_a_constant_on_the_largest_problem(torch, tb):
    # a depth-6, degree-10 problem (262,144 leaves, 348,913,664 departure
    # points: the largest in the 256-problem batch) beside a small one; a
    # constant concentration stays constant: every coefficient but c_000
    # stays 0 and c_000 stays 1, to FP32 rounding
    from tbslas_amd import fields
    from tbslas_amd.batch import Batch

    vel, conc = [], []
    for depth_u, q in ((6, 10), (3, 6)):
        anchor, depth = tb.uniform_leaves(depth_u)
        vel.append(tb.Tree(anchor, depth, q, 3,
                           fields.rotation_coefficients(anchor, depth, q, (0.48, 0.6, 0.64),
                                                        2 * math.pi)))
        c = np.zeros((len(depth), 1, tb.ncoef(q)))
        c[:, 0, 0] = 1.0
        conc.append(tb.Tree(anchor, depth, q, 1, c))
    b = Batch(vel, conc, [1.0 / 200, 1.0 / 200], 1)
    b.step(2)
    for p in range(2):
        cf = b.coeffs(p)
        assert np.abs(cf[:, 0, 0] - 1.0).max() <= 1e-6
        assert np.abs(cf[:, 0, 1:]).max() <= 1e-6
    b.close()


@pytest.mark.parametrize("mode", ["fp32", "fp64"])
def test_attached_tracers_ride_the_batch_graph(torch, tb, mode):
    # a 16-tracer concentration (matrix-core contraction) attached to a batch
    # of small problems: every replay of the one step graph advances the
    # tracers exactly as an eager in-situ step does and the problems exactly
    # as the batch alone does; the replays alternate with eager work and
    # synchronisations, and every replay is checked
    from tbslas_amd import fields
    from tbslas_amd.batch import Batch

    probs = _small_problems(tb, n=6)
    vel = [tb.Tree(a, d, q, 3, v) for a, d, q, v, c, dt in probs]
    dts = [dt for *_, dt in probs]
    q, T = 8, 16
    anchor, depth = tb.uniform_leaves(3)
    tv = tb.Tree(anchor, depth, q, 3, tb.fit_function(
        lambda p: fields.rotation(p, 2 * math.pi), anchor, depth, q, 3))
    centres, sigma = fields.tracer_blobs(T, 3)
    ccf = tb.fit_function(lambda p: fields.gaussian_blobs(p, centres, sigma), anchor, depth, q, T)
    tc = tb.Tree(anchor, depth, q, T, ccf)
    ref = tb.Tree(anchor, depth, q, T, ccf)
    alone = Batch(vel, [tb.Tree(a, d, q_, 1, c) for a, d, q_, v, c, dt in probs], dts, 0)
    b = Batch(vel, [tb.Tree(a, d, q_, 1, c) for a, d, q_, v, c, dt in probs], dts, 0)
    b.attach_step(tv, tc, 0.005, 1, mode)
    probe = torch.rand((4096, 3), dtype=torch.float64, device="cuda")
    for r in range(3):
        b.step(1)
        alone.step(1)
        tb.semilag_insitu(tv, ref, 1, 0.005, 1, 0, mode)
        ref.eval(probe.clone(), 0, "fp64")  # eager work between replays
        torch.cuda.synchronize()
        assert np.array_equal(tc.get_coeffs(), ref.get_coeffs()), r
        for p in range(len(probs)):
            assert np.array_equal(b.coeffs(p), alone.coeffs(p)), (r, p)
    b.close()
    alone.close()
