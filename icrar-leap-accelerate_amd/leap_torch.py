"""torch.autograd bridge for the calibration-residual loss.

`ResidualLoss.apply(offsets, antenna_delta, problem)` returns the loss of
leap_hip_residual_loss_grad as a float64 scalar tensor; its backward hands
back the gradients the device adjoint computed (direction offsets [D, 2],
antenna position corrections [N, 3]).  `problem` bundles the context-side
objects: a `leap_amd.Metadata`, the `MsView`, the base directions and the
per-timestep ENU-to-UVW matrices.
"""
from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class CalibrationProblem:
    md: object           # leap_amd.Metadata
    view: object         # leap_amd.MsView
    directions: np.ndarray  # [D, 2] base directions (ra, dec)
    enu_to_uvw: np.ndarray  # [T, 3, 3]


class ResidualLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, offsets, antenna_delta, problem):
        off = offsets.detach().to("cpu", torch.float64).numpy()
        dp = antenna_delta.detach().to("cpu", torch.float64).numpy()
        loss, g_off, g_ant = problem.md.loss_grad(problem.view, problem.directions, problem.enu_to_uvw,
                                                  offsets=off, antenna_delta=dp,
                                                  want_grad=ctx.needs_input_grad[0] or ctx.needs_input_grad[1])
        ctx.save_for_backward(torch.from_numpy(g_off).to(offsets.device), torch.from_numpy(g_ant).to(offsets.device))
        return torch.tensor(loss, dtype=torch.float64, device=offsets.device)

    @staticmethod
    def backward(ctx, grad_out):
        g_off, g_ant = ctx.saved_tensors
        return grad_out * g_off, grad_out * g_ant, None


def residual_loss(offsets, antenna_delta, problem):
    return ResidualLoss.apply(offsets, antenna_delta, problem)
