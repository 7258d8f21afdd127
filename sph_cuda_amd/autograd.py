"""torch.autograd binding of the adjoint (include/asph.h, asph_adjoint_rain).

`rain_objective(theta0, qc0, base, cfg, nsteps, dt, checkpoint)` returns the rain water J [kg] at
the end of an nsteps window started from `base` with its potential temperature replaced by
theta0 and the autoconversion threshold set to qc0. Its backward gives dJ/dtheta0 and dJ/dqc0
from the hand-written reverse sweep on the GPU (FP32 state, FP64 reductions). The forward call
runs the whole window and the reverse sweep at once (checkpointed recomputation needs the
forward anyway) and keeps the gradients for backward.
"""
from __future__ import annotations

import numpy as np
import torch

from .sim import Sim, State


class RainObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta0, qc0, base: State, cfg, nsteps: int, dt: float, checkpoint: int):
        th = theta0.detach().to("cpu", torch.float32).numpy()
        c = cfg.copy()
        c.qc_auto = float(qc0.detach().cpu())
        thq = base.thq.copy()
        thq[:, 0] = th
        with Sim(c) as sim:
            sim.set_state(State(base.pos, base.vel, thq, base.rho))
            r = sim.adjoint_rain(nsteps, dt, checkpoint)
        ctx.save_for_backward(torch.from_numpy(r["grad_theta"].copy()).to(theta0.device),
                              torch.tensor(r["dJ_dqc0"], dtype=torch.float64))
        ctx.qc0_dtype = qc0.dtype
        ctx.stats = r
        return torch.tensor(r["J"], dtype=torch.float64, device=theta0.device)

    @staticmethod
    def backward(ctx, grad_out):
        g_theta, g_qc0 = ctx.saved_tensors
        go = grad_out.to(torch.float64)
        gt = (go.to(g_theta.device) * g_theta.to(torch.float64)).to(torch.float32) if ctx.needs_input_grad[0] else None
        gq = (go.cpu() * g_qc0).to(ctx.qc0_dtype) if ctx.needs_input_grad[1] else None
        return gt, gq, None, None, None, None, None


def rain_objective(theta0: torch.Tensor, qc0: torch.Tensor, base: State, cfg, nsteps: int, dt: float,
                   checkpoint: int = 10) -> torch.Tensor:
    """J(theta0, qc0) as a differentiable torch scalar (float64)."""
    if theta0.dim() != 1 or theta0.numel() != base.rho.shape[0]:
        raise ValueError("theta0 must be a vector with one entry per particle")
    if not isinstance(qc0, torch.Tensor):
        qc0 = torch.tensor(float(qc0), dtype=torch.float64)
    return RainObjective.apply(theta0, qc0, base, cfg, nsteps, dt, checkpoint)


def directional_check(theta_grad: np.ndarray, directions: np.ndarray) -> np.ndarray:
    """<dJ/dtheta0, d> for each row d of `directions`, reduced in FP64."""
    return directions.astype(np.float64) @ theta_grad.astype(np.float64)
