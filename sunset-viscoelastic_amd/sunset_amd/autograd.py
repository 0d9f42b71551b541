"""torch.autograd bindings of the device reverse mode (DI form, fixed dt).

`rollout_loss(solver, U0, nsteps, dt)` returns the trace loss
L = (1/T) sum_{t=1..T} mean_i (cxx_i + cyy_i - 2) of a T-step rollout started
from U0, a (6, npfb) float64 tensor of (ro, rou, rov, cxx, cxy, cyy); its
backward is the library's checkpointed adjoint (sv_rollout_vjp), not torch's
tape.  Tensors cross the boundary through host memory (CPU tensors): the
state lives in the library's own device buffers.
"""
from __future__ import annotations

import numpy as np
import torch

KEYS = ("ro", "rou", "rov", "cxx", "cxy", "cyy")


class RolloutLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, solver, nsteps, dt, checkpoint_every):
        u = U0.detach().to("cpu", torch.float64).numpy()
        solver.set_state({k: u[q] for q, k in enumerate(KEYS)}, dt=dt)
        solver.adj_setup(checkpoint_every)
        L = solver.rollout_forward(nsteps)
        ctx.solver = solver
        ctx.device = U0.device
        return torch.tensor(L, dtype=torch.float64, device=U0.device)

    @staticmethod
    def backward(ctx, gL):
        g = ctx.solver.rollout_vjp(loss_bar=float(gL))
        G = torch.from_numpy(np.stack([g[k] for k in KEYS])).to(ctx.device)
        return G, None, None, None, None


def rollout_loss(solver, U0, nsteps: int, dt: float, checkpoint_every: int = 20):
    return RolloutLoss.apply(U0, solver, nsteps, dt, checkpoint_every)
