"""Write2File (the tree writer the reference's drivers call,
advection.cpp:184-188, 317-324) as VTK XML: the files are read back here
and checked against the tree layout and the oracle's EvalTree.

Bars: point coordinates bit-identical with the uniform per-leaf points;
connectivity, offsets and cell types exact; values within 1e-12 relative of
the oracle (the writer evaluates with fused multiply-adds) at the points
pulled into their leaf by 1 - 2^-40.
"""
import os
import re

import numpy as np
import pytest

from conftest import random_tree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def read_vtu(path):
    """(header text, [array per DataArray in file order]) of a raw-appended
    VTK XML file with UInt64 block headers."""
    raw = open(path, "rb").read()
    mark = b'<AppendedData encoding="raw">\n_'
    k = raw.index(mark)
    head = raw[:k].decode()
    data = raw[k + len(mark):]
    arrays = []
    for m in re.finditer(r'<DataArray type="(\w+)"[^>]*offset="(\d+)"', head):
        typ, off = m.group(1), int(m.group(2))
        nbytes = int(np.frombuffer(data[off:off + 8], dtype="<u8")[0])
        dt = {"Float64": "<f8", "Int64": "<i8", "UInt8": "u1"}[typ]
        arrays.append(np.frombuffer(data[off + 8:off + 8 + nbytes], dtype=dt))
    return head, arrays


@pytest.mark.parametrize("q,dof,order", [(4, 1, 3), (6, 3, 5)])
def test_vtu_round_trip(torch, tb, orc, tmp_path, q, dof, order):
    anchor, depth = random_tree(q + 21, max_depth=3)
    rng = np.random.default_rng(q)
    cf = rng.standard_normal((len(depth), dof, tb.ncoef(q))) / (1 + np.arange(tb.ncoef(q)))
    t = tb.Tree(anchor, depth, q, dof, cf)
    prefix = str(tmp_path / "val_T0003_P")
    t.write_vtu(prefix, order)
    pv = open(prefix + ".pvtu").read()
    assert 'Source="val_T0003_P_000000.vtu"' in pv
    head, (val, xyz, conn, offs, types) = read_vtu(prefix + "_000000.vtu")
    L, n = len(depth), order
    assert f'NumberOfPoints="{L * n ** 3}"' in head
    assert f'NumberOfCells="{L * (n - 1) ** 3}"' in head
    u = np.arange(n) / (n - 1)
    m = np.arange(n ** 3)
    loc = np.stack([u[m % n], u[(m // n) % n], u[m // (n * n)]], axis=1)
    ln = 1.0 / (1 << depth.astype(np.int64)).astype(np.float64)
    exp = anchor.astype(np.float64)[:, None, :] * ln[:, None, None] + ln[:, None, None] * loc[None]
    assert np.array_equal(xyz.reshape(-1, 3), exp.reshape(-1, 3))
    inner = anchor.astype(np.float64)[:, None, :] * ln[:, None, None] + \
        ln[:, None, None] * (loc * (1.0 - 2.0 ** -40))[None]
    o = orc.Tree(anchor, depth, q, dof, cf)
    ref, _ = orc.eval_tree(o, inner.reshape(-1, 3), 0)
    assert np.abs(val.reshape(-1, dof) - ref).max() <= 1e-12 * np.abs(ref).max()
    nc = L * (n - 1) ** 3
    assert (types == 12).all() and len(types) == nc
    assert np.array_equal(offs, 8 * np.arange(1, nc + 1))
    first = conn[:8]
    assert list(first) == [0, 1, n + 1, n, n * n, n * n + 1, n * n + n + 1, n * n + n]
    assert conn.max() == L * n ** 3 - 1 and conn.min() == 0
