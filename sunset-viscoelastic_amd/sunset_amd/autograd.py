"""torch.autograd bindings of the device reverse mode (DI form, fixed dt).

`rollout_loss(solver, U0, nsteps, dt, Wi=None, beta=None, eps=None)` returns
the trace loss L = (1/T) sum_{t=1..T} mean_i (cxx_i + cyy_i - 2) of a T-step
rollout started from U0, a (6, npfb) float64 tensor of (ro, rou, rov, cxx,
cxy, cyy); Wi, beta and eps (the sPTT parameter), when given as 0-d float64
tensors, replace the solver's values and receive gradients too.  The backward
is the library's checkpointed adjoint (sv_rollout_vjp, sv_adj_param_grad),
not torch's tape.

A CUDA U0 stays on the device: its data pointers go to sv_set_state_device
and the gradient is written into a CUDA tensor by sv_rollout_vjp_device, on
torch's current stream, with no host copy.  That needs the library and torch
to share one HIP runtime: import torch before the library is first loaded
(this module does, so importing it first is enough); otherwise the device
entry points refuse the pointers with a message.  A CPU U0 crosses through
host memory.
"""
from __future__ import annotations

import numpy as np
import torch

KEYS = ("ro", "rou", "rov", "cxx", "cxy", "cyy")


def _set_params(solver, Wi, beta, eps):
    upd = {}
    if Wi is not None:
        upd["Wi"] = float(Wi)
    if beta is not None:
        upd["beta"] = float(beta)
    if eps is not None:
        upd["eps"] = float(eps)
    if upd:
        solver.set_phys(**dict(solver.phys, **upd))


class RolloutLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, Wi, beta, eps, solver, nsteps, dt, checkpoint_every):
        _set_params(solver, Wi, beta, eps)
        if U0.is_cuda:
            U = U0.detach().to(torch.float64).contiguous()
            if not getattr(solver, "_has_state", False):
                # the first call builds the context's state (boundary and
                # control records) once from a host copy
                u = U.cpu().numpy()
                solver.set_state({k: u[q] for q, k in enumerate(KEYS)}, dt=dt)
                solver._has_state = True
            else:
                solver.set_state_control(dt)
            stream = torch.cuda.current_stream(U.device).cuda_stream
            solver.set_state_device([U[q].data_ptr() for q in range(6)], stream)
            ctx.keep = U
        else:
            u = U0.detach().to("cpu", torch.float64).numpy()
            solver.set_state({k: u[q] for q, k in enumerate(KEYS)}, dt=dt)
            solver._has_state = True
        solver.adj_setup(checkpoint_every)
        L = solver.rollout_forward(nsteps)
        ctx.solver = solver
        ctx.device = U0.device
        ctx.cuda = U0.is_cuda
        ctx.n = U0.shape[1]
        ctx.has = (Wi is not None, beta is not None, eps is not None)
        return torch.tensor(L, dtype=torch.float64, device=U0.device)

    @staticmethod
    def backward(ctx, gL):
        if ctx.cuda:
            G = torch.empty((6, ctx.n), dtype=torch.float64, device=ctx.device)
            stream = torch.cuda.current_stream(ctx.device).cuda_stream
            ctx.solver.rollout_vjp_device([G[q].data_ptr() for q in range(6)],
                                          loss_bar=float(gL), stream=stream)
        else:
            g = ctx.solver.rollout_vjp(loss_bar=float(gL))
            G = torch.from_numpy(np.stack([g[k] for k in KEYS])).to(ctx.device)
        pg = ctx.solver.adj_param_grad()
        out = []
        for has, key in zip(ctx.has, ("Wi", "beta", "eps")):
            out.append(torch.tensor(pg[key], dtype=torch.float64) if has else None)
        return (G, *out, None, None, None, None)


def rollout_loss(solver, U0, nsteps: int, dt: float, checkpoint_every: int = 20, Wi=None,
                 beta=None, eps=None):
    return RolloutLoss.apply(U0, Wi, beta, eps, solver, nsteps, dt, checkpoint_every)
