"""torch.autograd.Function wrapper of the explicit-fracture adjoint (test
harness over the C ABI).

FracAdjointLoss.apply(E, setup, sensors, comp, ckpt) returns

    J(E) = sum_{s in sensors} u_comp(s)^2 at the last step,

the loss of the explicit PDDO fracture run of `setup` (a
problems.FracSetup built with node_E) with per-node Young's modulus E, and
its backward is the hand-written reverse sweep of the device
(pxg_frac_adjoint): bond health frozen from the recorded forward run,
displacement checkpoints every `ckpt` steps.  The forward pass reconfigures
the context for E, runs the recorded forward loop and evaluates J with the
frozen-health forward map; the reverse sweep runs only in backward, so a
forward without backward costs no adjoint.  If the context was used for
another run between forward and backward, backward records the forward run
for E again first.  E may live on any device; the computation runs on the
context's GPU through the C ABI.
"""
import numpy as np
import torch

from . import pxg
from .problems import NU


class FracAdjointLoss(torch.autograd.Function):
    @staticmethod
    def forward(fctx, E, setup, sensors, comp, ckpt):
        s = setup
        e = np.ascontiguousarray(E.detach().to("cpu", torch.float64).numpy())
        if e.shape != (s.cloud.n,):
            raise ValueError("E must hold one Young's modulus per PD node")
        _record(s, e)
        J = s.ctx.frac_loss(s.u0, s.dt, s.nsteps, sensors, comp)
        fctx.setup, fctx.e, fctx.sensors, fctx.comp, fctx.ckpt = s, e, sensors, comp, ckpt
        fctx.token = s.ctx.frac_serial
        fctx.device, fctx.dtype = E.device, E.dtype
        return E.new_tensor(J)

    @staticmethod
    def backward(fctx, grad_J):
        s = fctx.setup
        if s.ctx.frac_serial != fctx.token:
            _record(s, fctx.e)       # the context ran something else since forward
        _, g = s.ctx.frac_adjoint(s.u0, s.dt, s.nsteps, fctx.sensors, fctx.comp, ckpt=fctx.ckpt)
        return grad_J * torch.from_numpy(g).to(fctx.device, fctx.dtype), None, None, None, None


def _record(s, e):
    """Sets up the context for the per-node modulus e and runs the forward loop,
    which records the break step of every bond (the frozen health of the
    adjoint)."""
    s.ctx.frac_setup(s.params, pxg.FP64, node_E=e, nu=NU, plane_stress=False)
    s.ctx.frac_set_bcs(s.bcs)
    s.ctx.frac_init(s.u0, s.dt)
    s.ctx.frac_run(s.dt, s.nsteps * s.dt, nsteps_max=s.nsteps)


def frac_loss(E, setup, sensors, comp=1, ckpt=50):
    """J(E) with autograd support (see FracAdjointLoss)."""
    return FracAdjointLoss.apply(E, setup, np.ascontiguousarray(sensors, np.int32), comp, ckpt)
