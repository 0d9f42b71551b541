"""torch.autograd bindings of the device reverse mode (DI form, fixed dt).

`rollout_loss(solver, U0, nsteps, dt, Wi=None, beta=None)` returns the trace
loss L = (1/T) sum_{t=1..T} mean_i (cxx_i + cyy_i - 2) of a T-step rollout
started from U0, a (6, npfb) float64 tensor of (ro, rou, rov, cxx, cxy, cyy);
Wi and beta, when given as 0-d float64 tensors, replace the solver's values
and receive gradients too.  The backward is the library's checkpointed adjoint
(sv_rollout_vjp, sv_adj_param_grad), not torch's tape.  Tensors cross the
boundary through host memory (CPU tensors): the state lives in the library's
own device buffers.
"""
from __future__ import annotations

import numpy as np
import torch

KEYS = ("ro", "rou", "rov", "cxx", "cxy", "cyy")


class RolloutLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, Wi, beta, solver, nsteps, dt, checkpoint_every):
        upd = {}
        if Wi is not None:
            upd["Wi"] = float(Wi)
        if beta is not None:
            upd["beta"] = float(beta)
        if upd:
            solver.set_phys(**dict(solver.phys, **upd))
        u = U0.detach().to("cpu", torch.float64).numpy()
        solver.set_state({k: u[q] for q, k in enumerate(KEYS)}, dt=dt)
        solver.adj_setup(checkpoint_every)
        L = solver.rollout_forward(nsteps)
        ctx.solver = solver
        ctx.device = U0.device
        ctx.has = (Wi is not None, beta is not None)
        return torch.tensor(L, dtype=torch.float64, device=U0.device)

    @staticmethod
    def backward(ctx, gL):
        g = ctx.solver.rollout_vjp(loss_bar=float(gL))
        G = torch.from_numpy(np.stack([g[k] for k in KEYS])).to(ctx.device)
        pg = ctx.solver.adj_param_grad()
        gW = torch.tensor(pg["Wi"], dtype=torch.float64) if ctx.has[0] else None
        gb = torch.tensor(pg["beta"], dtype=torch.float64) if ctx.has[1] else None
        return G, gW, gb, None, None, None, None


def rollout_loss(solver, U0, nsteps: int, dt: float, checkpoint_every: int = 20, Wi=None,
                 beta=None):
    return RolloutLoss.apply(U0, Wi, beta, solver, nsteps, dt, checkpoint_every)
