"""Many tracers carried by one trajectory set (C3a at reduced size): a
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
