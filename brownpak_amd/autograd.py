"""torch.autograd.Function over the hand-written trajectory adjoint.

``rgsq_trajectory(theta, system, crd0, disp)`` returns the final Rg^2 of a
trajectory replayed with fixed Brownian displacements ``disp`` (recorded by
``BdStep.run_record``) as a differentiable function of
theta = (LJ eps, LJ sigma, bond k).  The forward pass runs the EM steps on
the GPU (bd_run_replay); the backward pass runs the adjoint sweep with
checkpoints (bd_adjoint_rgsq).  There is no autograd tape through torch ops:
the gradient comes from the hand-written adjoint kernels.
"""
from __future__ import annotations

import numpy as np
import torch


def _prepare(system, crd0, theta, tie_rc):
    system.set_coordinates(crd0)
    system.reset_run()
    t = [float(v) for v in theta.detach().cpu().double().tolist()]
    system.set_theta(t[0], t[1], t[2], tie_rc=tie_rc)


class RgsqTrajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, system, crd0, disp, ckpt, tie_rc):
        _prepare(system, crd0, theta, tie_rc)
        system.run_replay(disp)
        loss = system.rgsq()
        ctx.system, ctx.crd0, ctx.disp, ctx.ckpt, ctx.tie_rc = system, crd0, disp, ckpt, tie_rc
        ctx.save_for_backward(theta)
        return torch.tensor(loss, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, gout):
        (theta,) = ctx.saved_tensors
        _prepare(ctx.system, ctx.crd0, theta, ctx.tie_rc)
        _, g = ctx.system.adjoint_rgsq(ctx.disp, ckpt=ctx.ckpt, tie_rc=ctx.tie_rc)
        grad = torch.as_tensor(np.asarray(g), dtype=theta.dtype, device=theta.device)
        return gout * grad, None, None, None, None, None


def rgsq_trajectory(theta, system, crd0, disp, ckpt=20, tie_rc=True):
    return RgsqTrajectory.apply(theta, system, crd0, disp, ckpt, tie_rc)
