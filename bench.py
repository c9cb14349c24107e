#!/usr/bin/env python3
"""Benchmark of the semi-Lagrangian step on one MI355X.

Workload (adaptive Chebyshev octree, q = 14): a 2:1-balanced octree seeded
at uniform depth 5 and refined (up to depth 7) where the top-shell Chebyshev
tail of a Gaussian blob (sigma 0.04, centre (0.7, 0.5, 0.5)) exceeds 1e-13;
36,324 leaves, 122,593,500 departure points per step.  Velocity: the
Taylor-Green field (sin 2pi x cos 2pi y cos 2pi z, -cos 2pi x sin 2pi y
cos 2pi z, 0) times cos(pi t), sampled on the same tree; periodic; dt = 1e-3;
one RK2 substep.  A step is SolveSemilagInSitu: arrival points -> RK2 through
the velocity tree -> concentration at the departure points -> refit.

Reported: whole-job departure-point interpolations/s (3 per point per step:
two velocity evaluations and one concentration evaluation) of three
arithmetic modes of the same step: plain FP64, FP32 (coefficients and
accumulation in FP32) and FP32 with the degree <= 1 head of each
contraction in FP64 (headline `value`); the L-inf distance of each FP32
mode from the FP64 path after the first step and after all steps; the
roofline of the dominant kernel measured with HIP events on the launch
stream; and the FP64 CPU oracle timed on one full step of the same
workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "departure-pt interps/s on 1 GPU: batched FP32 vs plain FP64 at L∞≤1e-6"
from tbslas_amd import workloads as W  # noqa: E402

Q = W.ADAPTIVE_Q
DT = W.ADAPTIVE_DT
PEAK = {"fp64": 78.6, "fp32": 157.3}  # TFLOP/s, MI355X published dense
# The kernels issue no packed FP32 instructions, so an FP32 vector kernel
# issues one FMA per lane per clock, the FP64 rate (64 FMA/clock/CU).
PEAK_FP32_UNPACKED = 78.6
MFMA_PEAK_F64, MFMA_PEAK_F32 = PEAK["fp64"], PEAK["fp32"]  # dense matrix cores
HBM_PEAK = 8000.0  # GB/s


def flops_per_point(q: int) -> int:
    """Algorithmic flops of one departure point in the step kernel: two
    velocity evaluations (dof 3), one concentration evaluation, RK2 updates
    (tree_functor.h:389-394 counts 3*d*3 + 2*Nc*dof per evaluation;
    traj.inc:44 counts 18 per point per substep)."""
    d = q + 1
    nc = d * (d + 1) * (d + 2) // 6
    return 2 * (9 * d + 2 * nc * 3) + (9 * d + 2 * nc * 1) + 18


def executed_flops_per_point(q: int) -> int:
    """Flops the step kernel itself executes per departure point when the
    first RK2 stage reads the velocity's nodal values (TBX_OPT_NODE_VELOCITY):
    the reference's count less one velocity evaluation (the few nodes on a
    face towards a leaf of another depth, 0.2 % on the benchmark tree, still
    evaluate and are not counted)."""
    d = q + 1
    nc = d * (d + 1) * (d + 2) // 6
    return flops_per_point(q) - (9 * d + 2 * nc * 3)


def expansion_flops_per_point(q: int) -> float:
    """Flops k_refit_expand executes per grid point for the velocity's
    expansion at its nodes: per leaf component D^3 (D+1)/2 multiply-adds for
    the z contraction over the (j, k) triangle and D^4 each for y and x;
    3 components, D^3 points per leaf."""
    d = q + 1
    return 2.0 * (d ** 3 * (d + 1) / 2 + 2 * d ** 4) * 3 / d ** 3


def build_workload():
    return W.adaptive_q14()


tfac = W.taylor_green_tfac


def run_mode(mode, vel, anchor, depth, ccf, steps, warmup):
    import torch

    import tbslas_amd as tb

    conc = tb.Tree(anchor, depth, Q, 1, ccf)
    vals = torch.empty((conc.L, 1, conc.P), dtype=torch.float64, device="cuda")
    for s in range(1, warmup + 1):
        tb.semilag_values(vel, conc, s, DT, 1, tb.BC_PERIODIC, mode, g=tfac, vals=vals)
        conc.set_grid_values(vals)
    conc.set_coeffs(ccf)  # timed run starts from the initial state
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    t0 = time.perf_counter()
    for s in range(1, steps + 1):
        e = ev[s - 1]
        e[0].record()
        tb.semilag_values(vel, conc, s, DT, 1, tb.BC_PERIODIC, mode, g=tfac, vals=vals)
        e[1].record()
        conc.set_grid_values(vals)
        e[2].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    k_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    r_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    # the step's first launch, the velocity tree's expansion at its grid
    # nodes (TBX_OPT_NODE_VELOCITY), timed alone after the timed steps so
    # that the step kernel's own time is k_ms - x_ms
    x_ms = 0.0
    from tbslas_amd import _lib

    if _lib.get_option(_lib.OPT_NODE_VELOCITY):
        # FP32 nodal values in the FP32 modes, as the step writes them
        buf = torch.empty((vel.L, 3, vel.P), device="cuda",
                          dtype=torch.float64 if mode == "fp64" else torch.float32)
        xe = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps)]
        for e in xe:
            e[0].record()
            vel.grid_values(out=buf)
            e[1].record()
        torch.cuda.synchronize()
        x_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in xe]))
        del buf
    return conc, wall, k_ms - x_ms, r_ms, x_ms


def host_cpu() -> dict:
    """CPU model and core counts of this host (lscpu), and the CPUs this
    process may run on (the calling thread's affinity: call it before an
    OpenMP region binds the thread)."""
    import subprocess

    info = {}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                             text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass

    def num(k):
        try:
            return int(info.get(k, ""))
        except ValueError:
            return None

    cps, sockets = num("Core(s) per socket"), num("Socket(s)")
    return {"model": info.get("Model name"),
            "physical_cores": cps * sockets if cps and sockets else None,
            "threads_per_core": num("Thread(s) per core"),
            "cpus_allowed": len(os.sched_getaffinity(0))}


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    return orc


def cpu_baseline(anchor, depth, ccf, vcf, frac: float, reps: int = 5):
    """The FP64 oracle (the reference algorithm restated, OpenMP on the
    host's cores) timed on a sample of the benchmark step: the arrival points
    of a seeded random `frac` of the leaves, traced back by RK2 through the
    whole velocity tree, the concentration evaluated at their departure
    points on the whole tree, and those leaves refit.  One warm-up step,
    then the median of `reps` timed steps."""
    orc = _oracle()
    L = len(depth)
    n_leaves = max(1, int(round(frac * L)))
    sel = np.sort(np.random.default_rng(0).choice(L, n_leaves, replace=False))
    ov = orc.Tree(anchor, depth, Q, 3, vcf)
    oc = orc.Tree(anchor, depth, Q, 1, ccf)
    sub = orc.Tree(anchor[sel], depth[sel], Q, 1, ccf[sel])
    pts = orc.collect_grid_points(sub)
    M = orc.pt2coeff(Q)

    def step():
        vals = orc.semilag_rk2(ov, oc, pts, 1, DT, 1, orc.BC_PERIODIC, g=tfac)
        orc.set_grid_values(sub, M, vals.reshape(n_leaves, 1, -1))

    step()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"interps_per_s": 3 * len(pts) / t, "points": len(pts), "leaves": n_leaves,
            "median_s": t, "times_s": times}


def oracle_step_reference(anchor, depth, ccf, vcf):
    """One full oracle step of the benchmark workload (timestep 1): the
    concentration coefficients every GPU mode is compared with."""
    orc = _oracle()
    ov = orc.Tree(anchor, depth, Q, 3, vcf)
    oc = orc.Tree(anchor, depth, Q, 1, ccf)
    t0 = time.perf_counter()
    orc.semilag_insitu(ov, oc, orc.pt2coeff(Q), 1, DT, 1, orc.BC_PERIODIC, g=tfac)
    return oc.coeff.copy(), time.perf_counter() - t0


def linf_rel(tree, coeff, pts, bc):
    """max |u - u_ref| / max |u_ref| over pts, u the tree's values and u_ref
    those of the same leaves holding coeff, both evaluated in TBX_EXACT."""
    import tbslas_amd as tb

    ref = tb.Tree(tree.anchor, tree.depth, tree.q, tree.dof)
    ref.set_coeffs(coeff)
    a = tree.eval(pts.clone(), bc, "exact")
    b = ref.eval(pts.clone(), bc, "exact")
    del ref
    return float((a - b).abs().max() / b.abs().max())


def uniform_q8_measure(cpu: bool):
    """The uniform depth-4, degree-8 case (4,096 leaves, 2,985,984 points per
    step, rigid rotation, Gaussian blob, free space, dt = 2 pi / 200) run in
    full: 50 in-situ steps per GPU mode, and the same 50 steps of the FP64
    oracle (OpenMP on the host's cores); each GPU mode's relative L-inf
    against the oracle after the 50 steps at every grid point."""
    import torch

    import tbslas_amd as tb

    anchor, depth, ccf, vcf = W.uniform_q8()
    q, dt, bc, T = W.UNIFORM_Q, W.UNIFORM_DT, W.UNIFORM_BC, W.UNIFORM_STEPS
    vel = tb.Tree(anchor, depth, q, 3, vcf)
    N = len(depth) * (q + 1) ** 3
    out = {"leaves": len(depth), "q": q, "departure_points_per_step": N, "steps": T, "dt": dt,
           "bc": "free space"}
    trees = {}
    for mode in ("fp64", "fp32", "fp32h"):
        conc = tb.Tree(anchor, depth, q, 1, ccf)
        tb.semilag_insitu(vel, conc, 1, dt, 1, bc, mode)  # warm-up
        conc.set_coeffs(ccf)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in range(1, T + 1):
            tb.semilag_insitu(vel, conc, k, dt, 1, bc, mode)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / T
        out[mode] = {"ms_per_step": ms, "interps_per_s": 3 * N / (ms * 1e-3),
                     "departure_points_per_s": N / (ms * 1e-3)}
        trees[mode] = conc
    if cpu:
        orc = _oracle()
        ov, oc = orc.Tree(anchor, depth, q, 3, vcf), orc.Tree(anchor, depth, q, 1, ccf)
        M = orc.pt2coeff(q)
        t0 = time.perf_counter()
        for k in range(1, T + 1):
            orc.semilag_insitu(ov, oc, M, k, dt, 1, bc)
        t = time.perf_counter() - t0
        out["cpu_oracle"] = {"s_total": t, "ms_per_step": 1e3 * t / T,
                             "interps_per_s": 3 * N * T / t,
                             "threads": int(os.environ.get("OMP_NUM_THREADS", "0")) or None}
        pts = trees["fp64"].grid_points()
        out["linf_rel_vs_oracle_after_all_steps"] = {
            m: linf_rel(trees[m], oc.coeff, pts, bc) for m in trees}
        out["fp64_over_cpu_oracle"] = out["fp64"]["interps_per_s"] / out["cpu_oracle"][
            "interps_per_s"]
        del pts
    del trees, vel
    return out


def tricubic_measure(steps: int):
    """The tricubic variant at full size: 262,144 leaves (uniform depth 6)
    of 8^3 samples, Lamb-Oseen pair velocity, smoothed slotted sphere,
    dt = 5e-4, FreeSpace.  Returns per-mode step times and the achieved HBM
    fraction on algorithmic bytes (samples read once: 3 velocity + 1
    concentration component, new samples written once as FP64)."""
    import torch

    import tbslas_amd as tb
    from tbslas_amd import fields

    anchor, depth = tb.uniform_leaves(6)
    n = 8
    vs = tb.sample_function(fields.lamb_oseen_pair, anchor, depth, n, 3)
    cs = tb.sample_function(lambda p: fields.smoothed_slotted_sphere(p, 2.0 / (64 * (n - 1))),
                            anchor, depth, n, 1)
    vel = tb.Grid(anchor, depth, n, 3, vs)
    npts = len(depth) * n ** 3
    out = {"leaves": len(depth), "samples_per_leaf": n ** 3, "departure_points": npts,
           "dt": 5e-4, "steps": steps}
    for mode, sb in (("fp64", 8), ("fp32", 4)):
        conc = tb.Grid(anchor, depth, n, 1, cs)
        tb.grid_semilag_insitu(vel, conc, 1, 5e-4, 1, 0, mode)  # warm-up
        conc.set_samples(cs)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in range(1, steps + 1):
            tb.grid_semilag_insitu(vel, conc, k, 5e-4, 1, 0, mode)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        nbytes = npts * (4 * sb + 8)
        out[mode] = {"ms_per_step": ms, "departure_points_per_s": npts / (ms * 1e-3),
                     "interps_per_s": 3 * npts / (ms * 1e-3),
                     "algorithmic_bytes": nbytes,
                     "hbm_gbs": nbytes / (ms * 1e-3) / 1e9,
                     "hbm_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK}
        del conc
    return out


def _time_locate(tree, pts, reps: int):
    import torch

    tree.locate(pts, 1)  # warm-up
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        tree.locate(pts, 1)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    n = pts.shape[0]
    nbytes = 28 * n
    return {"points": n, "ms": ms, "points_per_s": n / (ms * 1e-3),
            "algorithmic_bytes": nbytes, "hbm_gbs": nbytes / (ms * 1e-3) / 1e9,
            "hbm_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK}


def locate_measure(tree, n: int = 1 << 27, reps: int = 5):
    """Bit-exact leaf location alone (tbx_locate: periodic wrap, 21-bit
    keys, directory and search) on the benchmark tree, HBM-bound with
    algorithmic bytes 24 (coordinates) + 4 (leaf id) per point, for two
    point sets: n uniformly random points (every directory read a random
    one), and the departure points of the benchmark's first step (its grid
    points traced back one dt), which arrive in leaf order as they do inside
    a step."""
    import torch

    import tbslas_amd as tb

    gen = torch.Generator(device="cuda").manual_seed(11)
    pts = torch.rand((n, 3), dtype=torch.float64, device="cuda", generator=gen)
    out = {"uniform_random": _time_locate(tree, pts, reps)}
    del pts
    x = tree.grid_points()
    dep = tb.traj_rk2(tree, x, DT, 0.0, 1, tb.BC_PERIODIC, "fp64", g=tfac)
    del x
    out["step_departure_points"] = _time_locate(tree, dep, reps)
    del dep
    torch.cuda.empty_cache()
    return out


def contraction_measure(tree, reps: int = 3):
    """The leaf contraction alone (tbx_contraction_only) on the benchmark's
    velocity tree: as many points as a step has departure points, 53
    wavefronts per leaf (3,375 points, as a step's arrival points), the 3
    components contracted one after another at each point.  Rate in
    algorithmic flops, 2 x 815 per component (680 triangle terms + 135 for
    the nesting at q = 14), against the peak of the mode's arithmetic."""
    import torch

    n = tree.L * tree.P
    wpl = -(-tree.P // 64)
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    res = {"points": n, "waves_per_leaf": wpl, "components": tree.dof}
    per_point = 2 * ((Q + 1) * (Q + 2) * (Q + 3) // 6 + (Q + 1) * (Q + 2) // 2 + (Q + 1)) * tree.dof
    for mode in ("fp32", "fp32h", "fp64"):
        tree.contraction_only(n, wpl, mode, out=out)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            tree.contraction_only(n, wpl, mode, out=out)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        tf = per_point * n / (ms * 1e-3) / 1e12
        peak = PEAK["fp64" if mode == "fp64" else "fp32"]
        res[mode] = {"ms": ms, "tflops": tf, "peak": peak, "frac": tf / peak}
    res["flops_per_point"] = per_point
    del out
    return res


def adjoint_measure():
    """The FP64 adjoint at full size: uniform depth 5 (32,768 leaves), q = 8,
    stationary random Chebyshev velocity (SplitMix64 seed 4, |v| <= 1), a
    Gaussian blob, 20 steps of dt = 0.01, FreeSpace, checkpoints every 5
    steps.  The adjoint run = 20 forward steps + 16 replayed + 20 adjoint
    steps + one velocity scatter.  With the stationary velocity the run keeps
    step 1's departure points and its other 35 forward steps evaluate the
    concentration there and refit; the adjoint step is the run less its
    forward steps (one full step + 35 of those, timed on their own), per
    step.  Also reported: one tbx_semilag_insitu_vjp reading back the
    forward's departure points, without and with the velocity cotangent."""
    import torch

    import tbslas_amd as tb
    from tbslas_amd import fields
    from tbslas_amd.adjoint import adjoint_run

    q, d, T, dt, K = 8, 5, 20, 0.01, 5
    anchor, depth = tb.uniform_leaves(d)
    vc = tb.fit_function(fields.random_poly_velocity(q, seed=4), anchor, depth, q, 3)
    c0 = tb.fit_function(lambda p: fields.gaussian(p, (0.5, 0.5, 0.5), 0.1), anchor, depth, q, 1)
    vel, conc = tb.Tree(anchor, depth, q, 3, vc), tb.Tree(anchor, depth, q, 1, c0)
    c0d = torch.from_numpy(c0).cuda()
    target = 0.9 * c0d

    def forward():
        conc.set_coeffs(c0d)
        for n in range(1, T + 1):
            tb.semilag_insitu(vel, conc, n, dt, 1, 0, "fp64")

    def run():
        conc.set_coeffs(c0d)
        adjoint_run(vel, conc, T, dt, target, 0, "fp64", checkpoint_every=K)

    def timed(f, reps):
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    fwd_ms = timed(forward, 3)
    run_ms = timed(run, 3)
    step_ms = fwd_ms / T
    from tbslas_amd.adjoint import semilag_insitu_vjp

    lam = torch.randn(c0.shape, dtype=torch.float64, device="cuda",
                      generator=torch.Generator(device="cuda").manual_seed(5))
    gv = torch.zeros((vel.L, 3, vel.nc), dtype=torch.float64, device="cuda")
    dep = torch.empty((conc.L * conc.P, 3), dtype=torch.float64, device="cuda")
    conc.set_coeffs(c0d)
    tb.semilag_insitu(vel, conc, 1, dt, 1, 0, "fp64", dep=dep)
    conc.set_coeffs(c0d)
    vjp_c = timed(lambda: semilag_insitu_vjp(vel, conc, 1, dt, lam, 0, "fp64", dep=dep), 5)
    vjp_v = timed(lambda: semilag_insitu_vjp(vel, conc, 1, dt, lam, 0, "fp64", grad_vel=gv,
                                             dep=dep), 3)
    conc.set_coeffs(c0d)
    short_ms = timed(lambda: conc.set_grid_values(conc.eval(dep, 0, "fp64")), 5)
    replays = T - (T + K - 1) // K
    adj_ms = (run_ms - step_ms - (T - 1 + replays) * short_ms) / T
    del vel, conc
    return {"leaves": len(depth), "q": q, "steps": T, "checkpoint_every": K,
            "grid_points": len(depth) * (q + 1) ** 3,
            "forward_step_ms": step_ms, "adjoint_run_ms": run_ms,
            "plain_forward_ms": fwd_ms, "stored_trajectory_step_ms": short_ms,
            "vjp_ms": vjp_c, "vjp_with_velocity_ms": vjp_v, "adjoint_step_ms": adj_ms,
            "adjoint_step_over_forward_step": adj_ms / step_ms,
            "adjoint_run_over_plain_forward": run_ms / fwd_ms}


def _lib_option(name: str) -> int:
    from tbslas_amd import _lib

    return _lib.get_option(getattr(_lib, name))


def tracers_measure(steps: int = 2):
    """Batched tracers at full size: 64 seeded Gaussian tracers (SplitMix64
    seed 3) on a uniform depth-6 tree (262,144 leaves), q = 8, carried by one
    trajectory set of the rigid rotation v = 2 pi (0.5 - y, x - 0.5, 0),
    dt = 1/200, FreeSpace; a 64-component concentration tree, one in-situ
    step = trajectories once + 64 contractions per departure point + refit.
    interps/step = N (2 + 64).  Accuracy: FP32 against FP64 after one step
    at 2^20 random points."""
    import torch

    import tbslas_amd as tb
    from tbslas_amd import fields

    q, T, depth_u = 8, 64, 6
    anchor, depth = tb.uniform_leaves(depth_u)
    L, P = len(depth), (q + 1) ** 3
    N = L * P
    omega, dt = 2 * math.pi, 1.0 / 200
    vel = tb.Tree(anchor, depth, q, 3, tb.fit_function(lambda p: fields.rotation(p, omega),
                                                       anchor, depth, q, 3))
    centres, sigma = fields.tracer_blobs(T, 3)
    # initial coefficients on the device: per tracer, values at the grid
    # points and the refit of a one-component tree
    one = tb.Tree(anchor, depth, q, 1)
    gp = one.grid_points()
    C = torch.empty((L, T, tb.ncoef(q)), dtype=torch.float64, device="cuda")
    ct = torch.from_numpy(centres).cuda()
    st = torch.from_numpy(sigma).cuda()
    for k in range(T):
        v = torch.exp(-((gp - ct[k]) ** 2).sum(1) / (2 * st[k] ** 2))
        one.set_grid_values(v)
        C[:, k, :] = one.get_coeffs_device()[:, 0, :]
        del v
    del gp, one
    torch.cuda.empty_cache()
    conc = tb.Tree(anchor, depth, q, T)
    probe = torch.rand((1 << 20, 3), dtype=torch.float64, device="cuda",
                       generator=torch.Generator(device="cuda").manual_seed(3))
    out = {"tracers": T, "leaves": L, "q": q, "departure_points": N,
           "interps_per_step": N * (2 + T), "dt": dt, "steps": steps}
    res = {}
    for mode in ("fp64", "fp32"):
        conc.set_coeffs(C)
        tb.semilag_insitu(vel, conc, 1, dt, 1, 0, mode)  # warm-up, also the accuracy state
        res[mode] = conc.eval(probe.clone(), 0, "fp64")
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in range(2, steps + 2):
            tb.semilag_insitu(vel, conc, k, dt, 1, 0, mode)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        # the tracer contraction runs on the matrix cores: its algorithmic
        # flops (2 T Nc per departure point) over the whole step time, a
        # lower bound of the kernel's own fraction (the step also holds the
        # trajectories and the FP64 refit GEMM)
        flops = 2.0 * T * tb.ncoef(q) * N
        peak = MFMA_PEAK_F64 if mode == "fp64" else MFMA_PEAK_F32
        out[mode] = {"ms_per_step": ms, "interps_per_s": N * (2 + T) / (ms * 1e-3),
                     "departure_points_per_s": N / (ms * 1e-3),
                     "tracer_contraction_tflops_over_step": flops / (ms * 1e-3) / 1e12,
                     "mfma_peak_tflops": peak,
                     "mfma_frac_over_step": flops / (ms * 1e-3) / 1e12 / peak}
    out["tracer_contraction"] = ("matrix cores" if _lib_option("OPT_TRACER_MFMA") else
                                 "vector ALU")
    out["linf_rel_fp32_vs_fp64_after_1_step"] = float(
        (res["fp32"] - res["fp64"]).abs().max() / res["fp64"].abs().max())
    out["fp32_over_fp64"] = out["fp32"]["interps_per_s"] / out["fp64"]["interps_per_s"]
    del conc, C, vel, probe, res
    torch.cuda.empty_cache()
    return out


def batch_measure(steps: int = 2, nprob: int = 256):
    """The first nprob of the 256 small problems (workloads.batch_problems:
    depth in {3..6}, degree in {6, 8, 10}, rigid rotation about a random
    axis, a Gaussian blob, dt = 1/200) in one persistent graph-captured
    launch, FP32 contraction; the launch configuration autotuned.  Also the
    FP32 state after the steps against per-problem FP64 steps on the device
    (relative L-inf of the coefficients, worst problem)."""
    import torch

    import tbslas_amd as tb
    from tbslas_amd import fields
    from tbslas_amd.batch import CONFIGS, Batch

    t0 = time.perf_counter()
    vel, conc, dts, init = [], [], [], []
    npts = 0
    for depth_u, q, axis in W.batch_problems()[:nprob]:
        anchor, depth = tb.uniform_leaves(depth_u)
        vel.append(tb.Tree(anchor, depth, q, 3,
                           fields.rotation_coefficients(anchor, depth, q, axis, 2 * math.pi)))
        c = tb.Tree(anchor, depth, q, 1)
        gp = c.grid_points()
        c.set_grid_values(W.gaussian_grid_values(gp))
        del gp
        conc.append(c)
        init.append(c.get_coeffs_device())
        dts.append(1.0 / 200)
        npts += len(depth) * (q + 1) ** 3
    torch.cuda.empty_cache()
    b = Batch(vel, conc, dts, 0)
    setup = time.perf_counter() - t0
    ms = b.tune()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    b.step(steps)
    e1.record()
    torch.cuda.synchronize()
    step_ms = e0.elapsed_time(e1) / steps
    # tuning repeats one step of the current parity (same input buffer), so
    # the state has advanced by the timed steps only
    taken = steps
    worst = 0.0
    for p in range(0, len(vel), max(1, len(vel) // 16)):
        ref = tb.Tree(conc[p].anchor, conc[p].depth, conc[p].q, 1)
        ref.set_coeffs(init[p])
        for _ in range(taken):
            tb.semilag_insitu(vel[p], ref, 1, dts[p], 1, 0, "fp64")
        r = ref.get_coeffs()
        worst = max(worst, float(np.abs(b.coeffs(p) - r).max() / np.abs(r).max()))
        del ref
    out = {"problems": len(vel), "of": 256, "departure_points_per_step": npts,
           "interps_per_step": 3 * npts, "setup_s": setup,
           "tuned_config": {"threads": CONFIGS[b.config][0],
                            "leaves_per_item": CONFIGS[b.config][1]},
           "tune_ms": {f"{t}x{lpi}": float(m) for (t, lpi), m in zip(CONFIGS, ms)},
           "ms_per_step": step_ms, "interps_per_s": 3 * npts / (step_ms * 1e-3),
           "departure_points_per_s": npts / (step_ms * 1e-3), "dtype": "f32",
           "coeff_linf_rel_vs_fp64_steps": worst, "steps_compared": taken,
           "first_stage": "nodal values" if b.node_velocity else "evaluated"}
    b.close()
    del vel, conc, init
    torch.cuda.empty_cache()
    return out


def combined_measure(steps: int = 2, nprob: int = 32):
    """The many-tracer step and the persistent batch in one graph launch:
    the 64 seeded tracers on the uniform depth-6, q = 8 tree (rotation
    velocity, dt = 1/200, free space) attached to the batch of the first
    nprob problems (workloads.batch_problems), FP32, so that each replay of
    the batch's step graph advances both.  interps/step = N_tracers (2 + 64)
    + 3 N_batch."""
    import torch

    import tbslas_amd as tb
    from tbslas_amd import fields
    from tbslas_amd.batch import Batch

    q, T = 8, 64
    anchor, depth = tb.uniform_leaves(6)
    L, P = len(depth), (q + 1) ** 3
    vel_t = tb.Tree(anchor, depth, q, 3, tb.fit_function(lambda p: fields.rotation(p, 2 * math.pi),
                                                         anchor, depth, q, 3))
    centres, sigma = fields.tracer_blobs(T, 3)
    one = tb.Tree(anchor, depth, q, 1)
    gp = one.grid_points()
    C = torch.empty((L, T, tb.ncoef(q)), dtype=torch.float64, device="cuda")
    ct, st = torch.from_numpy(centres).cuda(), torch.from_numpy(sigma).cuda()
    for k in range(T):
        one.set_grid_values(torch.exp(-((gp - ct[k]) ** 2).sum(1) / (2 * st[k] ** 2)))
        C[:, k, :] = one.get_coeffs_device()[:, 0, :]
    del gp, one
    conc_t = tb.Tree(anchor, depth, q, T)
    conc_t.set_coeffs(C)
    del C
    torch.cuda.empty_cache()
    vel, conc, dts = [], [], []
    npts = 0
    for depth_u, qq, axis in W.batch_problems()[:nprob]:
        a, d = tb.uniform_leaves(depth_u)
        vel.append(tb.Tree(a, d, qq, 3, fields.rotation_coefficients(a, d, qq, axis, 2 * math.pi)))
        c = tb.Tree(a, d, qq, 1)
        g = c.grid_points()
        c.set_grid_values(W.gaussian_grid_values(g))
        del g
        conc.append(c)
        dts.append(1.0 / 200)
        npts += len(d) * (qq + 1) ** 3
    b = Batch(vel, conc, dts, 0)
    b.tune()
    b.attach_step(vel_t, conc_t, 1.0 / 200, 1, "fp32")
    b.step(1)  # warm-up replay
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    b.step(steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    interps = L * P * (2 + T) + 3 * npts
    out = {"tracers": T, "tracer_leaves": L, "tracer_points": L * P, "problems": nprob,
           "problem_points": npts, "interps_per_step": interps, "ms_per_step": ms,
           "interps_per_s": interps / (ms * 1e-3), "dtype": "f32",
           "graph": "one hipGraphLaunch per step: the tracer step's kernels, then the batch's "
                    "counter reset and persistent kernel"}
    b.close()
    del b, vel, conc, vel_t, conc_t
    torch.cuda.empty_cache()
    return out


def traffic_from_profiles(kernel: str, shape: int):
    """HBM bytes per launch of a kernel at a launch shape from the committed
    rocprofv3 counter summary (profiles/step_traffic.txt: "kernel shape
    bytes" lines), if measured."""
    p = os.path.join(ROOT, "profiles", "step_traffic.txt")
    try:
        for line in open(p):
            f = line.split()
            if len(f) == 3 and not line.startswith("#") and f[0] == kernel and int(f[1]) == shape:
                return float(f[2])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-frac", type=float, default=float(os.environ.get("TBX_CPU_FRAC", 0.125)),
                    help="fraction of the leaves whose arrival points the CPU baseline step "
                         "traces (a seeded random sample)")
    ap.add_argument("--shape", type=int, default=None,
                    help="step kernel launch shape (default -1: per arithmetic mode)")
    ap.add_argument("--node-velocity", type=int, choices=(0, 1), default=None,
                    help="TBX_OPT_NODE_VELOCITY for every step (default: the library's, 1)")
    ap.add_argument("--grid-lds", type=int, choices=(0, 1, 2), default=None,
                    help="TBX_OPT_GRID_LDS for the tricubic steps (default: the library's, 1)")
    ap.add_argument("--profile", default=None, metavar="MODE",
                    help="counter-collection run: a 1 GiB device copy (the reference "
                         "for byte counters), then --steps step kernels of one "
                         "arithmetic mode (fp64, fp32, fp32h, exact); no JSON line")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip every host oracle run (CPU baseline and oracle L-inf)")
    ap.add_argument("--no-tricubic", action="store_true", help="skip the tricubic-variant measurement")
    ap.add_argument("--no-tracers", action="store_true", help="skip the 64-tracer measurement")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurement")
    ap.add_argument("--no-uniform", action="store_true", help="skip the uniform degree-8 case")
    ap.add_argument("--batch-problems", type=int, default=32,
                    help="problems of the persistent batch measured in the default run "
                         "(0: none; the first n of the 256)")
    ap.add_argument("--batch-only", action="store_true", help="only the 256-problem batch")
    ap.add_argument("--tracers-only", action="store_true", help="only the 64-tracer measurement")
    ap.add_argument("--combined-only", action="store_true",
                    help="only the 64 tracers and the first --batch-problems problems in one "
                         "graph launch")
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint measurement")
    ap.add_argument("--tricubic-only", action="store_true", help="only the tricubic measurement")
    ap.add_argument("--uniform-only", action="store_true", help="only the uniform degree-8 case")
    ap.add_argument("--locate-only", action="store_true",
                    help="only the stand-alone locate measurement on the benchmark tree")
    ap.add_argument("--contraction-only", action="store_true",
                    help="only the contraction-alone measurement on the benchmark tree")
    args = ap.parse_args()
    # before any OpenMP region: with OMP_PROC_BIND set, the oracle's first
    # parallel region binds this (the initial) thread to one place
    cpu_info = host_cpu()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU; --gpus above 1 is not supported",
              file=sys.stderr)
        sys.exit(2)
    # the CPU baseline's thread placement (set before the oracle's OpenMP
    # runtime starts)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")

    import torch

    import tbslas_amd as tb
    from tbslas_amd import _lib

    if args.node_velocity is not None:
        _lib.set_option(_lib.OPT_NODE_VELOCITY, args.node_velocity)
    if args.grid_lds is not None:
        _lib.set_option(_lib.OPT_GRID_LDS, args.grid_lds)

    if args.batch_only:
        print(json.dumps({"batch_256_problems": batch_measure(2, 256)}))
        return
    if args.tracers_only:
        print(json.dumps({"tracers_64": tracers_measure(2)}))
        return
    if args.combined_only:
        print(json.dumps({"tracers_and_batch_one_graph": combined_measure(
            2, max(1, args.batch_problems))}))
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint_20_steps": adjoint_measure()}))
        return
    if args.tricubic_only:
        print(json.dumps({"tricubic_262144_leaves": tricubic_measure(10)}))
        return
    if args.uniform_only:
        print(json.dumps({"uniform_4096_leaves_q8": uniform_q8_measure(not args.no_cpu)}))
        return
    t_setup = time.perf_counter()
    anchor, depth, ccf, vcf = build_workload()
    L = len(depth)
    P = (Q + 1) ** 3
    N = L * P
    vel = tb.Tree(anchor, depth, Q, 3, vcf)
    print(f"# workload: {L} leaves, {N} departure points/step, setup "
          f"{time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)
    if args.locate_only:
        print(json.dumps({"locate": locate_measure(vel)}))
        return
    if args.contraction_only:
        print(json.dumps({"contraction_only": contraction_measure(vel)}))
        return

    from tbslas_amd import _lib

    if args.shape is not None:
        _lib.set_option(_lib.OPT_STEP_SHAPE, args.shape)

    if args.profile:
        src = torch.full((1 << 27,), 1.0, dtype=torch.float64, device="cuda")
        torch.empty_like(src).copy_(src)
        conc = tb.Tree(anchor, depth, Q, 1, ccf)
        vals = torch.empty((conc.L, 1, conc.P), dtype=torch.float64, device="cuda")
        for s in range(1, args.steps + 1):
            tb.semilag_values(vel, conc, s, DT, 1, tb.BC_PERIODIC, args.profile, g=tfac,
                              vals=vals)
            torch.cuda.synchronize()
            print(f"# {args.profile} step {s} done", file=sys.stderr, flush=True)
        return

    # every mode after one step, against FP64 on the device and (unless
    # --no-cpu) against one full oracle step, at every grid point
    if not args.no_cpu:
        ref1, t_ref = oracle_step_reference(anchor, depth, ccf, vcf)
        print(f"# oracle step: {t_ref:.1f} s", file=sys.stderr, flush=True)
    first = {}
    g = None
    for mode in ("fp64", "fp32", "fp32h"):
        c1 = tb.Tree(anchor, depth, Q, 1, ccf)
        tb.semilag_insitu(vel, c1, 1, DT, 1, tb.BC_PERIODIC, mode, g=tfac)
        if g is None:
            g = c1.grid_points()
        first[mode] = c1
    linf1 = {m: linf_rel(first[m], first["fp64"].get_coeffs_device(), g, tb.BC_PERIODIC)
             for m in ("fp32", "fp32h")}
    linf1_orc = None
    if not args.no_cpu:
        linf1_orc = {m: linf_rel(first[m], ref1, g, tb.BC_PERIODIC) for m in first}
        del ref1
    del first

    res = {}
    trees = {}
    for mode in ("fp64", "fp32", "fp32h"):
        conc, wall, k_ms, r_ms, x_ms = run_mode(mode, vel, anchor, depth, ccf, args.steps,
                                                args.warmup)
        trees[mode] = conc
        ms = 1e3 * wall / args.steps
        fl = flops_per_point(Q) * N
        res[mode] = {
            "value": 3 * N * args.steps / wall,
            "ms_per_step": ms,
            "departure_points_per_s": N * args.steps / wall,
            "steps_per_s": args.steps / wall,
            "step_kernel_ms": k_ms,
            "node_expansion_ms": x_ms,
            "refit_ms": r_ms,
            "step_kernel_tflops": fl / (k_ms * 1e-3) / 1e12,
        }
        print(f"# {mode}: {ms:.2f} ms/step (step kernel {k_ms:.2f} ms, node expansion "
              f"{x_ms:.2f} ms, refit {r_ms:.2f} ms)",
              file=sys.stderr, flush=True)

    # each FP32 mode against FP64 after the timed steps (the FP64 path is
    # pinned to the oracle by the one-step figures above)
    c64 = trees["fp64"].get_coeffs_device()
    linf = {m: linf_rel(trees[m], c64, g, tb.BC_PERIODIC) for m in ("fp32", "fp32h")}
    del g, c64

    kmode = {"fp64": "kFma64", "fp32": "kFp32", "fp32h": "kFp32h"}

    def roof(mode):
        k_s = res[mode]["step_kernel_ms"] * 1e-3
        ach = flops_per_point(Q) * N / k_s / 1e12
        peak = PEAK["fp64" if mode == "fp64" else "fp32"]
        shape = _lib.step_shape(mode, Q)
        kname = "k_step" if shape in (0, 1, 2, 3) else "k_step_rg"
        name = f"{kname}<{kmode[mode]}>"
        r = {"bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
             "frac": ach / peak, "traffic": traffic_from_profiles(name, shape),
             "traffic_source": "profiles/step_traffic.txt (rocprofv3 --pmc FETCH_SIZE and "
                               "WRITE_SIZE of this kernel at this shape, scaled by a 1 GiB copy)",
             "kernel": name, "launch_shape": shape, "flops_per_point": flops_per_point(Q),
             "points_per_launch": N, "timing": "HIP events on the launch stream"}
        if mode != "fp64":
            r["peak_note"] = ("published FP32 vector peak, reached only by packed FP32, which "
                              "these kernels do not issue")
            r["frac_of_unpacked_fp32"] = ach / PEAK_FP32_UNPACKED
        if _lib.get_option(_lib.OPT_NODE_VELOCITY):
            # achieved counts the reference's flops; the kernel executes fewer
            ex = executed_flops_per_point(Q) * N / k_s / 1e12
            r["executed_flops_per_point"] = executed_flops_per_point(Q)
            r["executed_tflops"] = ex
            r["executed_frac"] = ex / peak
            x_s = res[mode]["node_expansion_ms"] * 1e-3
            r["node_expansion"] = {
                "kernel": "k_refit_expand", "ms": res[mode]["node_expansion_ms"],
                "flops_per_point": expansion_flops_per_point(Q),
                "tflops": expansion_flops_per_point(Q) * N / x_s / 1e12 if x_s > 0 else None,
                "bytes": L * 3 * (8 * ((Q + 1) * (Q + 2) * (Q + 3) // 6)
                                  + (8 if mode == "fp64" else 4) * P),
            }
            if x_s > 0:
                r["node_expansion"]["hbm_gbs"] = r["node_expansion"]["bytes"] / x_s / 1e9
        return r

    out = {
        "metric": METRIC,
        "value": res["fp32h"]["value"],
        "unit": "interps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["fp32h"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "adaptive octree q=14 (seed depth 5, Gaussian sigma 0.04 tail > 1e-13, "
                        "max depth 7, 2:1 balanced), Taylor-Green velocity x cos(pi t), "
                        "periodic, dt=1e-3, 1 RK2 substep",
            "leaves": L, "q": Q, "departure_points_per_step": N, "interps_per_step": 3 * N,
            "dt": DT, "bc": "periodic",
        },
        "headline_mode": "fp32h (FP32 contraction, degree <= 1 head in FP64)",
        "first_stage": ("velocity expanded at the grid nodes each step (TBX_OPT_NODE_VELOCITY)"
                        if _lib.get_option(_lib.OPT_NODE_VELOCITY) else "evaluated"),
        "fp64": res["fp64"],
        "fp32": res["fp32"],
        "fp32h": res["fp32h"],
        "fp32h_over_fp64": res["fp32h"]["value"] / res["fp64"]["value"],
        "fp32_over_fp64": res["fp32"]["value"] / res["fp64"]["value"],
        "linf_rel_vs_oracle_after_1_step": linf1_orc,
        "linf_rel_vs_fp64_after_1_step": linf1,
        "linf_rel_vs_fp64_after_all_steps": linf,
        "roofline": roof("fp32h"),
        "roofline_fp32": roof("fp32"),
        "roofline_fp64": roof("fp64"),
    }
    del trees
    out["locate"] = locate_measure(vel)
    out["contraction_only"] = contraction_measure(vel)
    del vel
    torch.cuda.empty_cache()
    if not args.no_uniform:
        out["uniform_4096_leaves_q8"] = uniform_q8_measure(not args.no_cpu)
        torch.cuda.empty_cache()
    if not args.no_tricubic:
        out["tricubic_262144_leaves"] = tricubic_measure(10)
        torch.cuda.empty_cache()
    if not args.no_tracers:
        out["tracers_64"] = tracers_measure(2)
        torch.cuda.empty_cache()
    if not args.no_adjoint:
        out["adjoint_20_steps"] = adjoint_measure()
        torch.cuda.empty_cache()
    if args.batch_problems > 0:
        out["batch_problems"] = batch_measure(2, min(256, args.batch_problems))
        torch.cuda.empty_cache()
    if not args.no_cpu:
        cb = cpu_baseline(anchor, depth, ccf, vcf, args.cpu_frac)
        cpu = cpu_info
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or cpu["cpus_allowed"]
        out["cpu_baseline"] = {
            "value": cb["interps_per_s"], "unit": "interps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle (OpenMP, {threads} threads, OMP_PROC_BIND="
                      f"{os.environ.get('OMP_PROC_BIND')}, OMP_PLACES="
                      f"{os.environ.get('OMP_PLACES')}): one in-situ step of the benchmark "
                      f"workload restricted to the {cb['points']} arrival points of "
                      f"{cb['leaves']} seeded random leaves (RK2 and concentration on the whole "
                      f"tree, refit of those leaves); 1 warm-up, median of "
                      f"{len(cb['times_s'])} timed steps ({cb['median_s']:.2f} s)",
            "cpu": cpu, "times_s": cb["times_s"],
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
