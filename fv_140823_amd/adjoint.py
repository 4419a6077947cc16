"""torch.autograd.Function over the discrete adjoint of RHS + RK rollouts.

    J = sum_{n=1..nsteps} dt sum_{i in probe} p(S^n_i) Vol_i

Forward runs the rollout on the device (fvg_adj_forward, checkpoints every
`ckpt` steps; 0 keeps the whole rollout in HBM where it fits); backward runs the hand-written discrete adjoint
(fvg_adj_backward) and returns dJ/dS^0.  Tensors stay on the GPU: the C ABI
copies device to device on the solver's stream, and both sides synchronise
torch's current stream first.
"""
from __future__ import annotations

import torch

from .core import Solver


class RolloutObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, S0: torch.Tensor, solver: Solver, nsteps: int, ckpt: int, dt: float):
        if S0.dtype != torch.float64 or not S0.is_cuda:
            raise TypeError("S0 must be a float64 CUDA tensor [neq, ncell]")
        S0c = S0.contiguous()
        torch.cuda.current_stream().synchronize()
        solver.set_state_device(S0c.data_ptr())
        J = solver.adj_forward(nsteps, ckpt, dt)
        ctx.solver = solver
        ctx.shape = tuple(S0.shape)
        ctx.device = S0.device
        return torch.tensor(J, dtype=torch.float64, device=S0.device)

    @staticmethod
    def backward(ctx, gJ: torch.Tensor):
        g = torch.empty(ctx.shape, dtype=torch.float64, device=ctx.device)
        torch.cuda.current_stream().synchronize()
        ctx.solver.adj_backward_device(g.data_ptr())
        return g * gJ, None, None, None, None


def rollout_objective(rho0: torch.Tensor, mom0: torch.Tensor, E0: torch.Tensor, solver: Solver,
                      nsteps: int, ckpt: int, dt: float) -> torch.Tensor:
    """J as a differentiable function of the initial density [N] and momentum [3, N]
    (total energy E0 [N] held fixed)."""
    S0 = torch.cat([rho0[None], mom0, E0[None]], dim=0)
    return RolloutObjective.apply(S0, solver, nsteps, ckpt, dt)
