"""Host-side checks of libtbx.so that need no GPU: the library loads and
exports every entry point of include/tbx.h, tree validation, and the setup
helpers (nodes, refit operator, fits, 2:1 balance, adaptive refinement)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, random_tree


def declared_functions():
    names = []
    for h in ("tbx.h",):
        src = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"^\s*(?:int|const char \*|double)\s*\*?\s*(tbx_\w+)\s*\(", src, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(tb):
    names = declared_functions()
    assert len(names) >= 20
    L = C.CDLL(os.path.join(ROOT, "tbslas_amd", "libtbx.so"))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_nodes_bit_identical_to_oracle(orc, tb):
    for q in range(1, 17):
        assert (tb.new_nodes(q) == orc.new_nodes(q)).all()


def test_uniform_leaves_morton_order(tb):
    anchor, depth = tb.uniform_leaves(2)
    assert len(depth) == 64 and (depth == 2).all()
    key = np.zeros(64, dtype=np.int64)
    for b in range(2):
        for a in range(3):
            key |= ((anchor[:, a].astype(np.int64) >> b) & 1) << (3 * b + a)
    assert (key == np.arange(64)).all()


def _create_error(tb, anchor, depth, q=2, dof=1):
    """tbx_tree_create on a layout it must reject before touching a device."""
    from tbslas_amd import _lib

    anchor = np.ascontiguousarray(anchor, dtype=np.uint32)
    depth = np.ascontiguousarray(depth, dtype=np.uint8)
    desc = _lib._TreeDesc(len(depth), anchor.ctypes.data, depth.ctypes.data, q, dof, None)
    h = C.c_void_p()
    rc = tb.lib().tbx_tree_create(C.byref(desc), C.byref(h))
    assert rc != 0 and not h.value
    assert tb.lib().tbx_last_error()
    return rc


def test_tree_create_rejects_bad_layouts(tb):
    anchor, depth = tb.uniform_leaves(1)
    assert _create_error(tb, anchor[::-1], depth[::-1]) == 4  # first leaf not at origin
    a2 = anchor.copy()
    a2[[3, 4]] = a2[[4, 3]]
    assert _create_error(tb, a2, depth) == 4  # leaf 2 no longer meets its successor
    a3 = anchor.copy()
    a3[4] = a3[3]
    assert _create_error(tb, a3, depth) == 3  # repeated leaf: keys not increasing
    assert _create_error(tb, anchor[:-1], depth[:-1]) == 4  # gap at the far corner
    assert _create_error(tb, anchor, depth, q=17) == 7  # degree not built
    assert _create_error(tb, anchor, depth, q=0) == 7
    d2 = depth.copy()
    d2[0] = 22
    assert _create_error(tb, anchor, d2) == 7  # deeper than the keys resolve
    assert _create_error(tb, np.zeros((0, 3)), np.zeros(0)) == 1


def _neighbour_depths_ok(anchor, depth, bc):
    lo = anchor * (0.5 ** depth.astype(float))[:, None]
    h = 0.5 ** depth.astype(float)
    for j in range(len(depth)):
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    if dx == dy == dz == 0:
                        continue
                    p = lo[j] + h[j] * (0.5 + np.array([dx, dy, dz]))
                    if bc == 1:
                        p = p % 1.0
                    elif (p < 0).any() or (p >= 1).any():
                        continue
                    inside = np.all((p >= lo) & (p < lo + h[:, None]), axis=1)
                    k = int(np.argmax(inside))
                    if depth[k] < depth[j] - 1:
                        return False
    return True


@pytest.mark.parametrize("bc", [0, 1])
def test_balance21(tb, bc):
    anchor, depth = random_tree(5, max_depth=5, p_split=0.25)
    a, d = tb.balance21(anchor, depth, bc)
    assert len(d) >= len(depth)
    assert _neighbour_depths_ok(a, d, bc)


def test_adaptive_refinement_meets_tolerance(tb):
    from tbslas_amd import fields

    fn = lambda p: fields.gaussian(p, (0.6, 0.5, 0.5), 0.06)
    q, tol = 6, 1e-4
    a, d = tb.adaptive_leaves(fn, q, 1, tol, max_depth=5, balance=False)
    cf = tb.fit_function(fn, a, d, q, 1)
    tail = tb.tail_norm(cf, q, 1)
    assert (tail[d < 5] <= tol).all()
    assert d.max() == 5 and d.min() <= 2
    # leaves still tile the unit cube
    assert abs((0.125 ** d.astype(float)).sum() - 1.0) < 1e-12


def test_tail_norm_is_top_shell_sum(tb):
    q = 3
    cf = np.zeros((1, 2, tb.ncoef(q)))
    m = 0
    for i in range(q + 1):
        for j in range(q + 1 - i):
            for k in range(q + 1 - i - j):
                cf[0, 0, m] = (m + 1) if i + j + k == q else 100.0
                cf[0, 1, m] = -0.5 if i + j + k == q else 0.0
                m += 1
    top = [m + 1 for m, (i, j, k) in enumerate(
        (i, j, k) for i in range(q + 1) for j in range(q + 1 - i) for k in range(q + 1 - i - j))
        if i + j + k == q]
    assert tb.tail_norm(cf, q, 2)[0] == sum(top)


def test_splitmix64_known_answers():
    # the first outputs of SplitMix64 from seed 0 (0xe220a8397b1dcdaf,
    # 0x6e789e6aa1b965f4, 0x06c45d188009454f), as u = (z >> 11) * 2^-53
    from tbslas_amd import fields

    u = fields.splitmix64(0, 3)
    ref = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert list(u) == [(z >> 11) * 2.0 ** -53 for z in ref]


def test_random_poly_velocity_is_bounded_and_polynomial():
    import tbslas_amd as tb
    from tbslas_amd import fields

    f = fields.random_poly_velocity(4, seed=4)
    p = np.random.default_rng(0).random((2000, 3))
    v = f(p)
    assert np.abs(v).max() <= 1.0
    # a degree-4 polynomial is reproduced exactly by a degree-4 fit
    anchor, depth = tb.uniform_leaves(1)
    cf = tb.fit_function(f, anchor, depth, 4, 3)
    cf2 = tb.fit_function(f, anchor, depth, 6, 3)
    assert np.abs(cf).max() > 0 and np.isfinite(cf2).all()


def test_rotation_coefficients_closed_form():
    # the closed-form coefficients of a rigid rotation equal the fitted ones
    import tbslas_amd as tb
    from tbslas_amd import fields

    anchor, depth = tb.uniform_leaves(2)
    axis = (0.3, -0.5, 0.81)
    for q in (1, 6):
        fit = tb.fit_function(lambda p: fields.rotation_axis(p, axis, 2.5), anchor, depth, q, 3)
        cf = fields.rotation_coefficients(anchor, depth, q, axis, 2.5)
        assert np.abs(fit - cf).max() <= 1e-13


def test_step_shape_option(tb):
    # -1 (the default) resolves per mode and degree; explicit shapes are
    # taken where built for the degree and fall back to 6 elsewhere
    from tbslas_amd import _lib
    from tbslas_amd._lib import TbxError

    old = _lib.get_option(_lib.OPT_STEP_SHAPE)
    try:
        _lib.set_option(_lib.OPT_STEP_SHAPE, -1)
        assert _lib.get_option(_lib.OPT_STEP_SHAPE) == -1
        assert [_lib.step_shape(m, 14) for m in ("fp64", "exact", "fp32", "fp32h")] == [6, 6, 8, 10]
        assert [_lib.step_shape(m, 8) for m in ("fp64", "fp32", "fp32h")] == [6, 6, 6]
        _lib.set_option(_lib.OPT_STEP_SHAPE, 7)
        assert _lib.step_shape("fp32h", 14) == 7 and _lib.step_shape("fp32h", 5) == 6
        for bad in (-2, 11):
            with pytest.raises(TbxError):
                _lib.set_option(_lib.OPT_STEP_SHAPE, bad)
        with pytest.raises(TbxError):
            _lib.step_shape("fp64", 0)
    finally:
        _lib.set_option(_lib.OPT_STEP_SHAPE, old)


def test_departure_buffer_checked_before_any_launch(tb):
    # a host array or a CPU tensor handed as the departure-point buffer is
    # refused before the step kernel could write through it
    import torch

    from tbslas_amd.tree import _check_dep

    _check_dep(None, 10)
    with pytest.raises(TypeError):
        _check_dep(np.zeros((10, 3)), 10)
    with pytest.raises(TypeError):
        _check_dep(torch.zeros((10, 3), dtype=torch.float64), 10)
