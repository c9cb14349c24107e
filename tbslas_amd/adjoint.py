"""Adjoint of the semi-Lagrangian step over the libtbx C ABI.

The reference has no adjoint.  The step c' = R E_D(V) c (tbx_semilag_insitu:
departure points D of the grid points by RK2 substeps through the velocity
tree V, interpolation E_D, per-leaf refit R) is differentiated by
hand in the library (tbx_semilag_insitu_vjp, tbx_adjoint_run).  The
torch.autograd.Function here wraps the same C entry points; it exists to
check the hand-written adjoint against finite differences of the forward.
"""
from __future__ import annotations

import ctypes as C

from ._lib import BC_FREESPACE, check, lib, tfac
from .tree import Tree, _mode, _ptr, _require_dev, _stream


def semilag_insitu_vjp(vel: Tree, conc: Tree, timestep: int, dt: float, lam_next,
                       bc: int = BC_FREESPACE, mode: str = "fp64", g=None, grad_vel=None,
                       dep=None, num_rk_step: int = 1):
    """Cotangent of one in-situ step taken from conc's current coefficients:
    returns dL/dc for lam_next = dL/dc' (cuda [L, dof, Nc]); grad_vel (cuda
    [L_vel, 3, Nc]) is added to when given.  dep: the departure points the
    forward step wrote (semilag_insitu(..., dep=)), read back instead of
    integrating the trajectories again.  num_rk_step: the forward step's RK
    substeps."""
    import torch

    _require_dev(lam_next, torch.float64, "lam_next")
    if lam_next.numel() != conc.L * conc.dof * conc.nc:
        raise ValueError("lam_next must hold L*dof*Nc values")
    if grad_vel is not None:
        _require_dev(grad_vel, torch.float64, "grad_vel")
        if grad_vel.numel() != vel.L * 3 * vel.nc:
            raise ValueError("grad_vel must hold L_vel*3*Nc values")
    if dep is not None:
        _require_dev(dep, torch.float64, "dep", (3,))
        if dep.shape[0] != conc.L * conc.P:
            raise ValueError("dep must hold the L*P departure points of the step")
    lam_out = torch.empty_like(lam_next)
    cb = tfac(g)
    check(lib().tbx_semilag_insitu_vjp(vel.handle, cb, None, conc.handle, timestep, dt,
                                       num_rk_step, bc,
                                       _mode(mode), _ptr(dep), _ptr(lam_next), _ptr(lam_out),
                                       _ptr(grad_vel), _stream()))
    return lam_out


def adjoint_run(vel: Tree, conc: Tree, nsteps: int, dt: float, c_target, bc: int = BC_FREESPACE,
                mode: str = "fp64", g=None, checkpoint_every: int = 5, with_vel: bool = True):
    """nsteps in-situ steps from conc's coefficients, loss = 0.5 |c_T -
    c_target|^2 and its gradients (grad_c0, grad_vel or None); conc holds
    c_T afterwards."""
    import torch

    _require_dev(c_target, torch.float64, "c_target")
    if c_target.numel() != conc.L * conc.dof * conc.nc:
        raise ValueError("c_target must hold L*dof*Nc values")
    grad_c0 = torch.empty((conc.L, conc.dof, conc.nc), dtype=torch.float64, device="cuda")
    grad_vel = (torch.empty((vel.L, 3, vel.nc), dtype=torch.float64, device="cuda")
                if with_vel else None)
    loss = C.c_double()
    cb = tfac(g)
    check(lib().tbx_adjoint_run(vel.handle, cb, None, conc.handle, nsteps, dt, bc, _mode(mode),
                                _ptr(c_target), checkpoint_every, _ptr(grad_c0),
                                _ptr(grad_vel), C.byref(loss), _stream()))
    return loss.value, grad_c0, grad_vel


def _step_function():
    import torch

    class SemilagInSituStep(torch.autograd.Function):
        """c_next = SolveSemilagInSitu(c; V) as a differentiable function of
        the concentration coefficients c and the velocity coefficients v
        (cuda float64 [L, dof, Nc] and [L_vel, 3, Nc]); the trees give the
        leaf structure and are used as workspaces."""

        @staticmethod
        def forward(ctx, c, v, vel: Tree, conc: Tree, timestep: int, dt: float, bc: int,
                    mode: str):
            from .tree import semilag_insitu

            vel.set_coeffs(v.contiguous())
            conc.set_coeffs(c.contiguous())
            # the trajectories are kept for the backward pass
            dep = torch.empty((conc.L * conc.P, 3), dtype=torch.float64, device="cuda")
            semilag_insitu(vel, conc, timestep, dt, 1, bc, mode, dep=dep)
            out = conc.get_coeffs_device()
            ctx.save_for_backward(c, v, dep)
            ctx.args = (vel, conc, timestep, dt, bc, mode)
            return out

        @staticmethod
        def backward(ctx, lam_next):
            c, v, dep = ctx.saved_tensors
            vel, conc, timestep, dt, bc, mode = ctx.args
            vel.set_coeffs(v.contiguous())
            conc.set_coeffs(c.contiguous())
            grad_v = torch.zeros_like(v) if ctx.needs_input_grad[1] else None
            lam = semilag_insitu_vjp(vel, conc, timestep, dt, lam_next.contiguous(), bc, mode,
                                     grad_vel=grad_v, dep=dep)
            return lam, grad_v, None, None, None, None, None, None

    return SemilagInSituStep


_STEP = None


def semilag_step(c, v, vel: Tree, conc: Tree, timestep: int, dt: float, bc: int = BC_FREESPACE,
                 mode: str = "fp64"):
    """Differentiable in-situ step (torch.autograd) through the hand-written
    adjoint; see SemilagInSituStep."""
    global _STEP
    if _STEP is None:
        _STEP = _step_function()
    return _STEP.apply(c, v, vel, conc, timestep, dt, bc, mode)
