"""The initial fit (InitTree's cheb_approx role, tree_utils.h:132-178) and
the CheckChebOutput metric (utils.txx:118-297) on the device, against the
oracle's cheb_approx and the host restatement of the metric evaluated
through the oracle's EvalTree.

Bars: first-kind points bit-identical with the host's; coefficients within
1e-14 of the largest (the transform's sums run in another order, with fused
multiply-adds); the metric's four numbers within 1e-12 relative.
"""
import numpy as np
import pytest

from conftest import random_tree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def field(p, dof):
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    cols = [np.exp(-((x - 0.6) ** 2 + (y - 0.45) ** 2 + (z - 0.5) ** 2) / 0.02),
            np.sin(3 * x) * np.cos(2 * y + z), x * y * y - z ** 3]
    return np.stack(cols[:dof], axis=1)


@pytest.mark.parametrize("q,dof", [(3, 1), (8, 3), (14, 1), (16, 2)])
def test_device_fit_against_oracle(torch, tb, orc, q, dof):
    anchor, depth = random_tree(q + 5, max_depth=3)
    t = tb.Tree(anchor, depth, q, dof)
    pts = t.first_kind_points()
    hp = tb.first_kind_points(anchor, depth, q).reshape(-1, 3)
    assert np.array_equal(pts.cpu().numpy(), hp)
    vals = field(hp, dof)  # [L*P][dof]
    t.fit(torch.from_numpy(vals).cuda())
    got = t.get_coeffs()
    P = (q + 1) ** 3
    ref = np.stack([orc.cheb_approx(q, dof, np.ascontiguousarray(vals[j * P:(j + 1) * P].T))
                    for j in range(len(depth))])
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    assert np.abs(got - tb.fit_leaves(vals.reshape(len(depth), P, dof), q, dof)).max() <= \
        1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("q,dof,mode", [(4, 1, "fp64"), (8, 3, "exact"), (14, 1, "fp64"),
                                        (8, 2, "fp32")])
def test_device_metric_against_host_metric(torch, tb, orc, q, dof, mode):
    from tbslas_amd import metrics

    anchor, depth = random_tree(q + 11, max_depth=3)
    cf = tb.fit_function(lambda p: field(p, dof), anchor, depth, q, dof)
    # a perturbed tree, so the error has a mean and a shape
    rng = np.random.default_rng(q)
    cf = cf + 1e-3 * rng.standard_normal(cf.shape) / (1 + np.arange(cf.shape[2]))
    t = tb.Tree(anchor, depth, q, dof, cf)
    o = orc.Tree(anchor, depth, q, dof, cf)
    host = metrics.check_cheb_output(anchor, depth, q, lambda p: orc.eval_tree(o, p)[0],
                                     lambda p: field(p, dof))
    ep = t.error_points()
    hp, _ = metrics.error_points(anchor, depth, q)
    assert np.abs(ep.cpu().numpy() - hp).max() <= 1e-15  # numpy and libm cos: 1 ulp
    dev = t.check_cheb_output(torch.from_numpy(field(hp, dof)).cuda(), mode)
    tol = 1e-12 if mode != "fp32" else 1e-4
    for a, b in zip(dev, host):
        assert abs(a - b) <= tol * abs(b), (dev, host)
