"""Build recipe for libtbx.so (HIP kernels for gfx950 + the C ABI).

Every kernel translation unit is compiled by hipcc in parallel into
tbslas_amd/_build/, then linked into tbslas_amd/libtbx.so next to this file.
The per-degree kernels come from one source compiled once per degree
(-DTBX_Q=1..16).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtbx.so")
MAX_Q = 16

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # the reference's arithmetic is reproduced exactly only without
    # contraction of a*b+c into fma (TBX_EXACT); FMA paths call fma explicitly
    "-ffp-contract=off",
    # no packed FP32 math on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}",
    f"-I{CSRC}",
    "-Wall",
    "-Wno-unused-function",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        ln for ln in r.stdout.splitlines()
        if "packed-fp32-ops" not in ln and "not a recognized feature" not in ln
    )
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")


def build(jobs: int | None = None, verbose: bool = False) -> str:
    """Compile every kernel unit and link libtbx.so; return its path."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(os.cpu_count() or 4, 16)
    units = []
    for q in range(1, MAX_Q + 1):
        units.append((os.path.join(CSRC, "tbx_kernels.hip"),
                      os.path.join(BUILD, f"tbx_kernels_q{q}.o"), [f"-DTBX_Q={q}"]))
        units.append((os.path.join(CSRC, "tbx_adjoint.hip"),
                      os.path.join(BUILD, f"tbx_adjoint_q{q}.o"), [f"-DTBX_Q={q}"]))
    units.append((os.path.join(CSRC, "tbx_common.hip"), os.path.join(BUILD, "tbx_common.o"), []))
    units.append((os.path.join(CSRC, "tbx_grid.hip"), os.path.join(BUILD, "tbx_grid.o"), []))
    units.append((os.path.join(CSRC, "tbx_batch.hip"), os.path.join(BUILD, "tbx_batch.o"), []))
    units.append((os.path.join(CSRC, "tbx_refit.hip"), os.path.join(BUILD, "tbx_refit.o"), []))
    units.append((os.path.join(CSRC, "tbx_adapt.hip"), os.path.join(BUILD, "tbx_adapt.o"), []))
    units.append((os.path.join(CSRC, "tbx_noalloc.hip"), os.path.join(BUILD, "tbx_noalloc.o"),
                  []))
    units.append((os.path.join(CSRC, "tbx_host.cpp"), os.path.join(BUILD, "tbx_host.o"),
                  ["-x", "hip"]))

    def compile_unit(u):
        src, obj, extra = u
        cmd = [HIPCC, *COMMON, *extra, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_unit, units))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


def build_drivers() -> list[str]:
    """C++ programs that use the engine through include/tbx/tbslas.hpp."""
    out = []
    cdir = os.path.join(ROOT, "tests", "cpp")
    for name in ("advection_driver",):
        exe = os.path.join(cdir, name)
        _run([HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-ffp-contract=off",
              f"-I{os.path.join(ROOT, 'include')}", os.path.join(cdir, name + ".cpp"),
              f"-L{HERE}", "-ltbx", "-Wl,-rpath,$ORIGIN/../../tbslas_amd", "-Wl,-rpath,/opt/rocm/lib",
              "-o", exe])
        out.append(exe)
    return out


def build_oracle() -> str:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-s"])
    return os.path.join(odir, "liboracle.so")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_drivers())
    print(build_oracle())
