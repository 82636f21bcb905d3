"""Implicit Newton and backward-Euler drivers with the iterate on the device.

The reference's static job runs one NonlinearSolver::solve
(RunPDProblem.cpp:63-92) and its transient job the fixed-step path of
TimeStepping::solve (TimeStepping.cpp:139-230); each Newton iteration
assembles K and RHS, applies the BC rows, solves K dU = RHS with the linear
solver, refuses a non-finite update, applies U += dU and the BC presets, and
stops once |RHS| < abs_tol or |RHS| <= rel_tol |RHS_0|
(NonlinearSolver.cpp:73-140).  Here all of that runs on the context stream:
the element assembler, the BC rows (pxg_csr_apply_bcs_dev), the direct solve
(pxg_solve_dev) and the update with the presets (pxg_newton_update_dev); the
host reads two norms per iteration.

Device vectors are torch tensors on the context's GPU.  The context works on
its own stream, so the drivers synchronise torch's stream before handing it
tensors that torch wrote and the context's stream before torch reads them.
"""
import math
from dataclasses import dataclass

import torch

from . import pxg


@dataclass
class NewtonResult:
    converged: bool
    iters: int
    rnorm: float
    rnorm0: float
    dunorm: float
    linear_ok: bool = True


class Workspace:
    """K (nnz), RHS and dU (rows) on the context's device."""

    def __init__(self, ctx):
        dev = torch.device("cuda", ctx.device)
        rows = ctx.n * ctx.ndof
        self.K = torch.empty(ctx.nnz, dtype=torch.float64, device=dev)
        self.R = torch.empty(rows, dtype=torch.float64, device=dev)
        self.dU = torch.empty(rows, dtype=torch.float64, device=dev)


def newton(ctx, assemble, u, uold, t, max_iters=50, abs_tol=1e-9, rel_tol=1e-12, ws=None):
    """NonlinearSolver::solve on the device.  assemble(u_ptr, uold_ptr, k_ptr,
    r_ptr) enqueues the element assembly on the context stream; u is updated
    in place.  A zero pivot or a non-finite update ends the solve unconverged
    with u at its last good value, as in the reference."""
    ws = ws or Workspace(ctx)
    it, rn, rn0, dn = 0, 0.0, 1.0, 0.0
    while it < max_iters:
        assemble(u.data_ptr(), uold.data_ptr(), ws.K.data_ptr(), ws.R.data_ptr())
        ctx.csr_apply_bcs_dev(t, u.data_ptr(), ws.K.data_ptr(), ws.R.data_ptr())
        try:
            ctx.solve_dev(ws.K.data_ptr(), ws.R.data_ptr(), ws.dU.data_ptr())
        except pxg.PxgError as e:
            if e.code != pxg.ERR_SINGULAR:
                raise
            return NewtonResult(False, it, rn, rn0, dn, linear_ok=False)
        rn_new, dn_new = ctx.newton_update_dev(t, u.data_ptr(), ws.dU.data_ptr(), ws.R.data_ptr())
        if not math.isfinite(dn_new):
            return NewtonResult(False, it, rn, rn0, dn, linear_ok=False)
        rn, dn = rn_new, dn_new
        if it == 0:
            rn0 = rn
        it += 1
        if rn < abs_tol or (rn0 > 0.0 and rn <= rel_tol * rn0):
            return NewtonResult(True, it, rn, rn0, dn)
    return NewtonResult(False, it, rn, rn0, dn)


def static_solve(ctx, assemble, u, **newton_kw):
    """The static job: one Newton solve at t = 0 with Uold = U (info.Dt = 0)."""
    torch.cuda.synchronize()
    uold = u.clone()
    torch.cuda.synchronize()
    res = newton(ctx, assemble, u, uold, 0.0, **newton_kw)
    ctx.sync()
    return res


def transient(ctx, assemble, u, dt, total, max_iters=50, abs_tol=1e-9, rel_tol=1e-12, on_step=None):
    """Backward-Euler steps of fixed dt (the reference's default, non-adaptive
    TimeStepping): per step the presets at t_new, a Newton solve with
    assemble(dt, u_ptr, uold_ptr, k_ptr, r_ptr), then Uold = U.  Raises when a
    step does not converge (the reference ends the run there).  Returns the
    Newton iteration count of every step."""
    ws = Workspace(ctx)
    torch.cuda.synchronize()
    uold = u.clone()
    torch.cuda.synchronize()
    t, step, iters = 0.0, 0, []
    eps = 1e-12 * total
    while t < total - eps:
        step += 1
        dta = min(dt, total - t)
        tnew = t + dta
        ctx.csr_preset_dev(tnew, u.data_ptr())
        res = newton(ctx, lambda up, uop, kp, rp: assemble(dta, up, uop, kp, rp), u, uold, tnew,
                     max_iters=max_iters, abs_tol=abs_tol, rel_tol=rel_tol, ws=ws)
        if not res.converged:
            raise RuntimeError("Newton did not converge at step %d (t=%.5e, dt=%.3e), |R|=%.5e"
                               % (step, tnew, dta, res.rnorm))
        iters.append(res.iters)
        t = tnew
        ctx.sync()
        uold.copy_(u)
        torch.cuda.synchronize()
        if on_step is not None:
            on_step(step, t, u)
    return iters
