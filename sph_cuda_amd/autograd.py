"""torch.autograd binding of the adjoint (include/asph.h, asph_adjoint_rain).

`rain_objective(theta0, qc0, base, cfg, nsteps, dt, checkpoint)` returns the rain water J [kg] at
the end of an nsteps window started from `base` with its potential temperature replaced by
theta0 and the autoconversion threshold set to qc0. Its backward gives dJ/dtheta0 and dJ/dqc0
from the hand-written reverse sweep on the GPU (FP32 state, FP64 reductions).

When an input needs a gradient, the forward call runs the window and the reverse sweep at once
(checkpointed recomputation needs the forward anyway) and keeps the gradients for backward;
otherwise it runs the forward window alone. A theta0 on the GPU is uploaded device to device
(asph_set_state_device); J and the gradients are returned on the devices of the inputs.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from .sim import Sim, State


def _particle_mass(cfg) -> float:
    """Particle mass [kg] as the library evaluates it in float: P0 / (R_D th0) x ((dx dy) dz)."""
    f = np.float32
    rho0 = f(f(1.0e5) / (f(287.05) * f(cfg.th0)))
    return float(f(rho0 * f(f(f(cfg.dx) * f(cfg.dy)) * f(cfg.dz))))


def _forward_only(sim: Sim, cfg, nsteps: int, dt: float, checkpoint: int) -> float:
    """The objective of asph_adjoint_rain without the reverse sweep: the same schedule (sort,
    force evaluation, steps with the neighbour structure rebuilt every `checkpoint` steps) and the
    rain grid summed exactly."""
    sim.sort()
    sim.force(dt)
    t, dtf = 0.0, float(np.float32(dt))
    for k in range(nsteps):
        sim.step(dt, t, do_sort=(k > 0 and k % checkpoint == 0))
        t += dtf
    rain, _, _ = sim.get_rain_fx()
    return float(int(rain.astype(np.int64).sum())) * 2.0 ** -40 * _particle_mass(cfg)


class RainObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta0, qc0, base: State, cfg, nsteps: int, dt: float, checkpoint: int):
        c = cfg.copy()
        c.qc_auto = float(qc0.detach().cpu())
        need_grad = bool(ctx.needs_input_grad[0] or ctx.needs_input_grad[1])
        with Sim(c) as sim:
            if theta0.is_cuda:
                dev = theta0.device
                thq = torch.from_numpy(np.ascontiguousarray(base.thq, np.float32)).to(dev)
                thq[:, 0] = theta0.detach().to(torch.float32)
                arrs = [torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
                        for a in (base.pos, base.vel, base.rho)]
                torch.cuda.synchronize(dev)   # the library reads them on its own stream
                sim.set_state_device(arrs[0], arrs[1], thq, arrs[2])
            else:
                thq = base.thq.copy()
                thq[:, 0] = theta0.detach().to(torch.float32).numpy()
                sim.set_state(State(base.pos, base.vel, thq, base.rho))
            if need_grad:
                r = sim.adjoint_rain(nsteps, dt, checkpoint)
                J = r["J"]
                ctx.save_for_backward(torch.from_numpy(r["grad_theta"].copy()).to(theta0.device),
                                      torch.tensor(r["dJ_dqc0"], dtype=torch.float64, device=qc0.device))
                ctx.stats = r
            else:
                J = _forward_only(sim, c, nsteps, dt, checkpoint)
        ctx.qc0_dtype = qc0.dtype
        return torch.tensor(J, dtype=torch.float64, device=theta0.device)

    @staticmethod
    def backward(ctx, grad_out):
        g_theta, g_qc0 = ctx.saved_tensors
        go = grad_out.to(torch.float64)
        gt = (go.to(g_theta.device) * g_theta.to(torch.float64)).to(torch.float32) if ctx.needs_input_grad[0] else None
        gq = (go.to(g_qc0.device) * g_qc0).to(ctx.qc0_dtype) if ctx.needs_input_grad[1] else None
        return gt, gq, None, None, None, None, None


def rain_objective(theta0: torch.Tensor, qc0: torch.Tensor, base: State, cfg, nsteps: int, dt: float,
                   checkpoint: int = 10) -> torch.Tensor:
    """J(theta0, qc0) as a differentiable torch scalar (float64)."""
    if theta0.dim() != 1 or theta0.numel() != base.rho.shape[0]:
        raise ValueError("theta0 must be a vector with one entry per particle")
    if not isinstance(qc0, torch.Tensor):
        qc0 = torch.tensor(float(qc0), dtype=torch.float64)
    checkpoint = checkpoint if checkpoint > 0 else 10
    if nsteps < 1:
        raise L.AsphError("rain_objective: nsteps must be >= 1")
    return RainObjective.apply(theta0, qc0, base, cfg, nsteps, dt, checkpoint)


def directional_check(theta_grad: np.ndarray, directions: np.ndarray) -> np.ndarray:
    """<dJ/dtheta0, d> for each row d of `directions`, reduced in FP64."""
    return directions.astype(np.float64) @ theta_grad.astype(np.float64)
