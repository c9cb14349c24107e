"""tbx_semilag_insitu_noalloc (a step that allocates no device memory: its
value buffer, refit workspace and the velocity's nodal values live in the
library's static scratch) called through the C ABI, against
tbx_semilag_insitu on the same inputs: bit-identical coefficients and
departure points in every arithmetic mode, on two calls in a row, the
second with a smaller tree; a step whose nodal values do not fit, which
evaluates its first stage instead; and a step larger than the scratch
refused before any work is queued."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def _trees(tb, depth_u, q):
    from tbslas_amd import fields

    anchor, depth = tb.uniform_leaves(depth_u)
    vcf = tb.fit_function(lambda p: fields.taylor_green_2d3(p) + 0.2, anchor, depth, q, 3)
    ccf = tb.fit_function(lambda p: fields.gaussian(p, (0.6, 0.5, 0.45), 0.1), anchor, depth, q,
                          1)
    return anchor, depth, vcf, ccf


def _noalloc(tb, torch, vel, conc, step, dt, bc, mode, g, dep):
    from tbslas_amd import _lib
    from tbslas_amd.tree import _mode, _ptr, _stream

    need = C.c_size_t(0)
    rc = tb.lib().tbx_semilag_insitu_noalloc(vel.handle, _lib.tfac(g), None, conc.handle, step,
                                             dt, 1, bc, _mode(mode), _ptr(dep), C.byref(need),
                                             _stream())
    return rc, need.value


@pytest.mark.parametrize("mode", ["exact", "fp64", "fp32h", "fp32"])
def test_noalloc_step_equals_allocating_step(torch, tb, mode):
    g = lambda t: math.cos(math.pi * t)
    for depth_u, q, bc in ((3, 8, 1), (2, 6, 0)):  # the second call is smaller
        anchor, depth, vcf, ccf = _trees(tb, depth_u, q)
        vel = tb.Tree(anchor, depth, q, 3, vcf)
        a = tb.Tree(anchor, depth, q, 1, ccf)
        b = tb.Tree(anchor, depth, q, 1, ccf)
        n = a.L * a.P
        da = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        db = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        for step in (1, 2):
            tb.semilag_insitu(vel, a, step, 0.01, 1, bc, mode, g=g, dep=da)
            rc, need = _noalloc(tb, torch, vel, b, step, 0.01, bc, mode, g, db)
            assert rc == 0, tb.lib().tbx_last_error()
            assert 0 < need <= 256 << 20
            torch.cuda.synchronize()
            assert torch.equal(da, db)
            assert np.array_equal(a.get_coeffs(), b.get_coeffs()), (depth_u, q, step)


def test_noalloc_refuses_a_step_beyond_the_scratch(torch, tb):
    # 32,768 leaves at q = 14 need 1.06 GB of values and workspace
    anchor, depth = tb.uniform_leaves(5)
    vel = tb.Tree(anchor, depth, 14, 3)
    conc = tb.Tree(anchor, depth, 14, 1)
    before = conc.get_coeffs_device().clone()
    rc, need = _noalloc(tb, torch, vel, conc, 1, 0.01, 0, "fp64", None, None)
    assert rc == 6 and need > 256 << 20  # TBX_EOVERFLOW
    assert b"scratch" in tb.lib().tbx_last_error()
    assert torch.equal(conc.get_coeffs_device(), before)


def test_noalloc_step_without_room_for_nodal_velocities(torch, tb):
    # 4,096 leaves at q = 14 in FP64: values and workspace (133 MB) fit the
    # 256 MiB scratch, the velocity's nodal values (332 MB) do not, so the
    # step evaluates its first RK2 stage: the same step as the allocating
    # one (which reads the nodal values) to rounding, and as the allocating
    # step with the nodal first stage switched off bit for bit.
    from tbslas_amd import _lib

    g = lambda t: math.cos(math.pi * t)
    anchor, depth, vcf, ccf = _trees(tb, 4, 14)
    vel = tb.Tree(anchor, depth, 14, 3, vcf)
    a = tb.Tree(anchor, depth, 14, 1, ccf)
    b = tb.Tree(anchor, depth, 14, 1, ccf)
    c = tb.Tree(anchor, depth, 14, 1, ccf)
    tb.semilag_insitu(vel, a, 1, 0.01, 1, 1, "fp64", g=g)
    rc, need = _noalloc(tb, torch, vel, b, 1, 0.01, 1, "fp64", g, None)
    assert rc == 0, tb.lib().tbx_last_error()
    nodal = 4096 * 3 * 15 ** 3 * 8
    assert need <= 256 << 20 and need + nodal > 256 << 20
    torch.cuda.synchronize()
    ref = a.get_coeffs()
    assert np.abs(b.get_coeffs() - ref).max() <= 1e-12 * np.abs(ref).max()
    old = _lib.get_option(_lib.OPT_NODE_VELOCITY)
    _lib.set_option(_lib.OPT_NODE_VELOCITY, 0)
    try:
        tb.semilag_insitu(vel, c, 1, 0.01, 1, 1, "fp64", g=g)
    finally:
        _lib.set_option(_lib.OPT_NODE_VELOCITY, old)
    assert np.array_equal(c.get_coeffs(), b.get_coeffs())
