"""torch.autograd front end of the hand-written adjoint.

``torque_loss(alpha, q, prim0, params, nsteps)`` is the time-averaged
gravitational torque on the gas after ``nsteps`` RK2 steps of one 2D disk,

    L = (1/T) sum_steps sum_substeps sum_interior dt_s dV rho F_phi r,

as a differentiable function of the viscosity alpha (``Explicit_Viscosity``),
the binary's mass ratio q (``MassRatio``) and the initial primitives
(NQ x Nzones in the reference's zone order).  The forward runs the steps and
the adjoint sweep on the GPU in one call (``disco_batch_adjoint``: the
checkpointed forward, then the reverse sweep); backward only scales the
gradients it kept by the incoming gradient.  When no input needs a gradient
the forward runs the steps alone (``disco_batch_torque_loss``).

Tensors are plumbing here: the C ABI takes host float64 buffers, so prim0 is
copied to the host and the gradients come back on prim0's device.
"""
from __future__ import annotations

import numpy as np
import torch

from .api import Batch, Setup


class TorqueLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alpha, q, prim0, params, nsteps, checkpoint_every):
        p = dict(params)
        p["Explicit_Viscosity"] = float(alpha)
        p["MassRatio"] = float(q)
        setup = Setup(p)
        batch = Batch([setup])
        try:
            batch.upload("prim", prim0.detach().to("cpu", torch.float64).numpy())
            batch.init()
            if any(ctx.needs_input_grad[:3]):
                loss, ga, gq, gp = batch.adjoint(int(nsteps), checkpoint_every=int(checkpoint_every))
            else:
                loss, ga, gq, gp = batch.torque_loss(int(nsteps)), 0.0, 0.0, np.zeros(tuple(prim0.shape))
        finally:
            batch.close()
            setup.close()
        ctx.save_for_backward(torch.tensor(ga, dtype=torch.float64, device=alpha.device),
                              torch.tensor(gq, dtype=torch.float64, device=q.device),
                              torch.from_numpy(np.ascontiguousarray(gp)).to(prim0.device))
        return torch.tensor(loss, dtype=torch.float64, device=prim0.device)

    @staticmethod
    def backward(ctx, g):
        ga, gq, gp = ctx.saved_tensors
        return (g * ga).to(ga.dtype), (g * gq).to(gq.dtype), (g * gp).to(gp.dtype), None, None, None


def torque_loss(alpha, q, prim0, params: dict, nsteps: int, checkpoint_every: int = 20):
    """The loss above; alpha and q are 0-dim float64 tensors, prim0 an
    (NQ, Nzones) float64 tensor, params the disk's parameter dictionary
    (its Explicit_Viscosity and MassRatio are replaced by alpha and q)."""
    return TorqueLoss.apply(alpha, q, prim0, params, nsteps, checkpoint_every)
