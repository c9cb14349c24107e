"""The hand-written adjoint of the in-situ step (no counterpart in the
reference), pinned by what the forward fixes:

* the concentration map c -> c' is linear, so <lam, F c> = <F^T lam, c> to
  rounding (dot-product test);
* the velocity cotangent agrees with central finite differences of the
  forward;
* the 20-step problem (uniform depth 5, q = 8, stationary random
  Chebyshev velocity, checkpoints every 5 steps): loss equal to the plain
  forward's, gradients equal to central differences (eps = 1e-6) along 32
  seeded directions (16 in c_0, 16 smooth velocity perturbations);
* the torch.autograd wrapper gives the C++ driver's gradients.
"""
import math

import numpy as np
import pytest

from conftest import random_tree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def _vel_field(p):
    from tbslas_amd import fields

    # Taylor-Green plus a drift: no stagnation on leaf faces, so a small
    # velocity perturbation moves no departure point across a face
    return fields.taylor_green_2d3(p) + np.array([0.31, 0.17, 0.23])


def _setup(tb, seed, q, dof):
    from tbslas_amd import fields

    anchor, depth = random_tree(seed, max_depth=3)
    vel = tb.Tree(anchor, depth, q, 3, tb.fit_function(_vel_field, anchor, depth, q, 3))
    blob = lambda p: np.concatenate(
        [fields.gaussian(p, (0.5 + 0.05 * k, 0.5, 0.45), 0.15) for k in range(dof)], axis=1)
    c0 = tb.fit_function(blob, anchor, depth, q, dof)
    conc = tb.Tree(anchor, depth, q, dof, c0)
    return anchor, depth, vel, conc, c0


def _forward(tb, vel, conc, c, timestep, dt, bc, g=None, nsub=1):
    conc.set_coeffs(c)
    tb.semilag_insitu(vel, conc, timestep, dt, nsub, bc, "fp64", g=g)
    return conc.get_coeffs_device()


@pytest.mark.parametrize("nsub,bc,timed", [(2, 0, False), (3, 1, True)])
def test_vjp_several_substeps(torch, tb, nsub, bc, timed):
    # a forward step of nsub RK2 substeps: the dot-product identity for the
    # concentration and central differences for the velocity cotangent,
    # which runs back through every substep
    from tbslas_amd.adjoint import semilag_insitu_vjp

    g = (lambda t: math.cos(math.pi * t)) if timed else None
    _, _, vel, conc, c0 = _setup(tb, 63, 6, 1)
    gen = torch.Generator(device="cuda").manual_seed(3)
    c = torch.randn(c0.shape, dtype=torch.float64, device="cuda", generator=gen)
    lam = torch.randn(c0.shape, dtype=torch.float64, device="cuda", generator=gen)
    fc = _forward(tb, vel, conc, c, 2, 0.04, bc, g, nsub)
    conc.set_coeffs(c)
    ftl = semilag_insitu_vjp(vel, conc, 2, 0.04, lam, bc, "fp64", g=g, num_rk_step=nsub)
    lhs, rhs = float((lam * fc).sum()), float((ftl * c).sum())
    assert abs(lhs - rhs) <= 1e-13 * float(lam.norm() * fc.norm()), (lhs, rhs)
    v0 = vel.get_coeffs_device().clone()
    c = torch.from_numpy(c0).cuda()
    gv = torch.zeros_like(v0)
    conc.set_coeffs(c)
    semilag_insitu_vjp(vel, conc, 2, 0.04, lam, bc, "fp64", g=g, grad_vel=gv, num_rk_step=nsub)

    def loss(v):
        vel.set_coeffs(v)
        return float((lam * _forward(tb, vel, conc, c, 2, 0.04, bc, g, nsub)).sum())

    eps = 1e-6
    for k in range(4):
        d = torch.randn(v0.shape, dtype=torch.float64, device="cuda", generator=gen)
        d = d / d.abs().max()
        fd = (loss(v0 + eps * d) - loss(v0 - eps * d)) / (2 * eps)
        an = float((gv * d).sum())
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-9 * float(gv.norm() * d.norm()), (k, fd, an)
    vel.set_coeffs(v0)


@pytest.mark.parametrize("bc,timed", [(0, False), (1, True)])
def test_vjp_dot_product(torch, tb, bc, timed):
    from tbslas_amd.adjoint import semilag_insitu_vjp

    g = (lambda t: math.cos(math.pi * t)) if timed else None
    _, _, vel, conc, c0 = _setup(tb, 61, 6, 2)
    gen = torch.Generator(device="cuda").manual_seed(1)
    c = torch.randn(c0.shape, dtype=torch.float64, device="cuda", generator=gen)
    lam = torch.randn(c0.shape, dtype=torch.float64, device="cuda", generator=gen)
    fc = _forward(tb, vel, conc, c, 3, 0.03, bc, g)
    conc.set_coeffs(c)
    ftl = semilag_insitu_vjp(vel, conc, 3, 0.03, lam, bc, "fp64", g=g)
    lhs = float((lam * fc).sum())
    rhs = float((ftl * c).sum())
    scale = float(lam.norm() * fc.norm())
    assert abs(lhs - rhs) <= 1e-13 * scale, (lhs, rhs)


@pytest.mark.parametrize("bc,timed", [(0, False), (1, True)])
def test_vjp_velocity_finite_differences(torch, tb, bc, timed):
    from tbslas_amd.adjoint import semilag_insitu_vjp

    g = (lambda t: math.cos(math.pi * t)) if timed else None
    _, _, vel, conc, c0 = _setup(tb, 62, 6, 1)
    v0 = vel.get_coeffs_device().clone()
    c = torch.from_numpy(c0).cuda()
    gen = torch.Generator(device="cuda").manual_seed(2)
    lam = torch.randn(c0.shape, dtype=torch.float64, device="cuda", generator=gen)
    gv = torch.zeros_like(v0)
    conc.set_coeffs(c)
    semilag_insitu_vjp(vel, conc, 2, 0.02, lam, bc, "fp64", g=g, grad_vel=gv)

    def loss(v):
        vel.set_coeffs(v)
        return float((lam * _forward(tb, vel, conc, c, 2, 0.02, bc, g)).sum())

    eps = 1e-6
    for k in range(4):
        d = torch.randn(v0.shape, dtype=torch.float64, device="cuda", generator=gen)
        d = d / d.abs().max()
        fd = (loss(v0 + eps * d) - loss(v0 - eps * d)) / (2 * eps)
        an = float((gv * d).sum())
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-9 * float(gv.norm() * d.norm()), (k, fd, an)
    vel.set_coeffs(v0)


def _c4(tb, torch):
    from tbslas_amd import fields

    q, depth_u = 8, 5
    anchor, depth = tb.uniform_leaves(depth_u)
    vf = fields.random_poly_velocity(q, seed=4)
    vc = tb.fit_function(vf, anchor, depth, q, 3)
    blob = lambda p: fields.gaussian(p, (0.5, 0.5, 0.5), 0.1)
    c0 = tb.fit_function(blob, anchor, depth, q, 1)
    return anchor, depth, q, vc, c0


def test_adjoint_20_steps_against_finite_differences(torch, tb):
    # 20 steps of dt = 0.01 (|v| <= 1: at most a third of a leaf per
    # step), FreeSpace, target = the blob advected by the 10 % faster field
    import time

    from tbslas_amd import fields
    from tbslas_amd.adjoint import adjoint_run

    anchor, depth, q, vc, c0 = _c4(tb, torch)
    T, dt, bc = 20, 0.01, 0
    vel = tb.Tree(anchor, depth, q, 3, vc)
    conc = tb.Tree(anchor, depth, q, 1, c0)
    vel_t = tb.Tree(anchor, depth, q, 3, 1.1 * vc)
    for n in range(1, T + 1):
        tb.semilag_insitu(vel_t, conc, n, dt, 1, bc, "fp64")
    target = conc.get_coeffs_device().clone()
    c0d = torch.from_numpy(c0).cuda()
    v0d = torch.from_numpy(vc).cuda()

    def forward_loss(c, v):
        conc.set_coeffs(c)
        vel.set_coeffs(v)
        for n in range(1, T + 1):
            tb.semilag_insitu(vel, conc, n, dt, 1, bc, "fp64")
        return 0.5 * float(((conc.get_coeffs_device() - target) ** 2).sum())

    conc.set_coeffs(c0d)
    vel.set_coeffs(v0d)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss, gc, gv = adjoint_run(vel, conc, T, dt, target, bc, "fp64", checkpoint_every=5)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    plain = forward_loss(c0d, v0d)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    print(f"20-step adjoint: loss {loss:.6e}, adjoint run {t_adj * 1e3:.1f} ms "
          f"(forward {T} steps + backward), plain forward {t_fwd * 1e3:.1f} ms")
    assert abs(loss - plain) <= 1e-12 * plain
    # smooth velocity directions: seeded combinations of six fitted
    # trigonometric fields
    tp = 2 * math.pi
    smooth = []
    for (a, b, c) in ((1, 2, 1), (2, 1, 3), (3, 1, 2), (1, 3, 1), (2, 2, 1), (1, 1, 3)):
        f = lambda p, a=a, b=b, c=c: np.stack(
            [np.sin(tp * (a * p[:, 0] + 0.1 * b)), np.cos(tp * (b * p[:, 1] + 0.2 * c)),
             np.sin(tp * (c * p[:, 2] + 0.3 * a))], axis=1)
        smooth.append(torch.from_numpy(tb.fit_function(f, anchor, depth, q, 3)).cuda())
    u = fields.splitmix64(44, 32)
    eps = 1e-6
    gen = torch.Generator(device="cuda")
    worst = 0.0
    for k in range(32):
        gen.manual_seed(int(u[k] * 2 ** 31))
        if k < 16:
            d = torch.randn(c0d.shape, dtype=torch.float64, device="cuda", generator=gen)
            d = d / d.abs().max()
            fd = (forward_loss(c0d + eps * d, v0d) - forward_loss(c0d - eps * d, v0d)) / (2 * eps)
            an = float((gc * d).sum())
            gn = float(gc.norm())
        else:
            # a smooth perturbation (a random global polynomial, continuous
            # across leaves): with a per-leaf random one, V + eps d jumps by
            # O(eps) across leaf faces, the O(eps) points that cross a face
            # change L by O(eps) each, and the difference quotient is off by
            # O(1) (it converges to the gradient only as eps -> 0)
            w = torch.from_numpy(2.0 * fields.splitmix64(4400 + k, len(smooth)) - 1.0)
            d = sum(float(w[b]) * smooth[b] for b in range(len(smooth)))
            fd = (forward_loss(c0d, v0d + eps * d) - forward_loss(c0d, v0d - eps * d)) / (2 * eps)
            an = float((gv * d).sum())
            gn = float(gv.norm())
        err = abs(fd - an) / abs(an)
        worst = max(worst, err)
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-10 * gn * float(d.norm()), (k, fd, an)
    print(f"20-step adjoint: worst relative FD mismatch over 32 directions {worst:.2e}")


def test_autograd_function_matches_driver(torch, tb):
    from tbslas_amd.adjoint import adjoint_run, semilag_step

    anchor, depth, vel, conc, c0 = _setup(tb, 63, 5, 1)
    dt, bc, T = 0.02, 1, 3
    v0 = vel.get_coeffs_device().clone()
    target = torch.from_numpy(c0).cuda() * 0.9
    c = torch.from_numpy(c0).cuda().requires_grad_(True)
    v = v0.clone().requires_grad_(True)
    x = c
    for n in range(1, T + 1):
        x = semilag_step(x, v, vel, conc, n, dt, bc)
    loss = 0.5 * ((x - target) ** 2).sum()
    loss.backward()
    conc.set_coeffs(torch.from_numpy(c0).cuda())
    vel.set_coeffs(v0)
    l2, gc, gv = adjoint_run(vel, conc, T, dt, target, bc, "fp64", checkpoint_every=2)
    assert abs(float(loss) - l2) <= 1e-12 * l2
    assert float((c.grad - gc).abs().max()) <= 1e-11 * float(gc.abs().max())
    assert float((v.grad - gv).abs().max()) <= 1e-11 * float(gv.abs().max())


@pytest.mark.parametrize("bc,timed,K", [(1, True, 2), (0, True, 3), (0, False, 2)])
def test_run_equals_chained_vjp(torch, tb, bc, timed, K):
    # adjoint_run keeps the trajectories of the recomputed steps (of step 1
    # for a stationary velocity) and reads them back; the chain of single
    # vjp steps integrates every trajectory again.  Same gradients.
    from tbslas_amd.adjoint import adjoint_run, semilag_insitu_vjp

    g = (lambda t: math.cos(math.pi * t)) if timed else None
    _, _, vel, conc, c0 = _setup(tb, 64, 5, 1)
    dt, T = 0.03, 5
    c0d = torch.from_numpy(c0).cuda()
    target = 0.8 * c0d
    states = [c0d.clone()]
    conc.set_coeffs(c0d)
    for n in range(1, T + 1):
        tb.semilag_insitu(vel, conc, n, dt, 1, bc, "fp64", g=g)
        states.append(conc.get_coeffs_device().clone())
    lam = states[-1] - target
    gv = torch.zeros((vel.L, 3, vel.nc), dtype=torch.float64, device="cuda")
    for n in range(T, 0, -1):
        conc.set_coeffs(states[n - 1])
        lam = semilag_insitu_vjp(vel, conc, n, dt, lam, bc, "fp64", g=g, grad_vel=gv)
    conc.set_coeffs(c0d)
    loss, gc, gvr = adjoint_run(vel, conc, T, dt, target, bc, "fp64", g=g, checkpoint_every=K)
    assert abs(loss - 0.5 * float(((states[-1] - target) ** 2).sum())) <= 1e-13 * loss
    assert float((gc - lam).abs().max()) <= 1e-12 * float(lam.abs().max())
    assert float((gvr - gv).abs().max()) <= 1e-11 * float(gv.abs().max())


@pytest.mark.parametrize("bc,timed", [(0, False), (1, True)])
def test_vjp_with_stored_trajectories(torch, tb, bc, timed):
    # the departure points the forward step wrote, read back by the VJP,
    # give the VJP that integrates them again (atomic sums: to rounding)
    from tbslas_amd.adjoint import semilag_insitu_vjp

    g = (lambda t: math.cos(math.pi * t)) if timed else None
    _, _, vel, conc, c0 = _setup(tb, 65, 6, 2)
    c = torch.from_numpy(c0).cuda()
    dep = torch.empty((conc.L * conc.P, 3), dtype=torch.float64, device="cuda")
    conc.set_coeffs(c)
    tb.semilag_insitu(vel, conc, 3, 0.03, 1, bc, "fp64", g=g, dep=dep)
    gen = torch.Generator(device="cuda").manual_seed(7)
    lam = torch.randn(c0.shape, dtype=torch.float64, device="cuda", generator=gen)
    out = []
    for d in (None, dep):
        conc.set_coeffs(c)
        gv = torch.zeros((vel.L, 3, vel.nc), dtype=torch.float64, device="cuda")
        out.append((semilag_insitu_vjp(vel, conc, 3, 0.03, lam, bc, "fp64", g=g, grad_vel=gv,
                                       dep=d), gv))
    (l0, g0), (l1, g1) = out
    assert float((l0 - l1).abs().max()) <= 1e-13 * float(l0.abs().max())
    assert float((g0 - g1).abs().max()) <= 1e-13 * float(g0.abs().max())
