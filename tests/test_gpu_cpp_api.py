"""The reference's functor/step API (include/tbx/tbslas.hpp) from a C++
program: tests/cpp/advection_driver (built by __graft_entry__.build())."""
import os
import re
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_functor_api_driver():
    exe = os.path.join(ROOT, "tests", "cpp", "advection_driver")
    assert os.path.exists(exe), "build the driver with __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    vals = dict((k, float(v)) for k, v in re.findall(r"(\w+)=([-+0-9.e]+)", r.stdout))
    # fused device trajectories == the reference's host loop over the same
    # functor (TBX_EXACT on both sides)
    assert vals["traj_diff"] == 0.0
    assert vals["rk2_diff"] == 0.0
    # rotating blob after 20 steps (1/10 revolution), q = 8, depth 3
    assert vals["linf"] < 2e-3
    # extrapolating and four-tree functors: fused == host loop
    assert vals["extrap_diff"] == 0.0
    assert vals["set_diff"] == 0.0
    assert vals["insitu_extrap_diff"] < 1e-12
