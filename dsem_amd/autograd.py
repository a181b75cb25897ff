"""PyTorch autograd binding of the GPU path.

``RHS``     -- Qt = R(JQ) as a differentiable function of the state (forward
               dsem_rhs_device, backward dsem_rhs_vjp_device).
``RKSolve`` -- ``nsteps`` RK4(5) steps as a differentiable function of the
               initial state and of nu = 1/Re (forward dsem_rk_steps,
               backward by step adjoints).  Checkpointing is two-level: the
               forward pass keeps the state every ``every`` steps; the
               backward pass re-runs one segment from its checkpoint while
               recording each step's stage states and face intermediates on
               the library's step tape (dsem_rk_step_record), then runs the
               segment's step adjoints in reverse from the tape
               (dsem_rk_step_vjp_tape).  Memory: nsteps/every checkpoints plus
               ``every`` tape slots; work: one extra forward pass.

Tensors are FP64 on the solver's GPU, shaped [K, 5, ncg^3].  The library
works on its own HIP stream and synchronises it before returning; this
module synchronises torch's stream before handing pointers over.
"""
from __future__ import annotations

import numpy as np
import torch

from .solver import DSEM


def _check_state(solver: DSEM, x: torch.Tensor) -> None:
    if x.dtype != torch.float64 or not x.is_cuda or not x.is_contiguous():
        raise TypeError("state tensors must be contiguous float64 CUDA tensors")
    if tuple(x.shape) != (solver.K, 5, solver.np):
        raise ValueError(f"state shape {tuple(x.shape)} != {(solver.K, 5, solver.np)}")


class RHS(torch.autograd.Function):
    @staticmethod
    def forward(ctx, JQ: torch.Tensor, solver: DSEM):
        _check_state(solver, JQ)
        out = torch.empty_like(JQ)
        torch.cuda.synchronize()
        solver._check(solver._L.dsem_rhs_device(solver._h, JQ.data_ptr(), out.data_ptr()))
        solver.synchronize()
        ctx.solver = solver
        ctx.save_for_backward(JQ)
        return out

    @staticmethod
    def backward(ctx, gout: torch.Tensor):
        (JQ,) = ctx.saved_tensors
        lam = gout.detach().to(torch.float64).contiguous()
        out = torch.empty_like(JQ)
        torch.cuda.synchronize()
        ctx.solver.rhs_vjp_device(JQ.data_ptr(), lam.data_ptr(), out.data_ptr())
        return out, None


class RKSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, JQ0: torch.Tensor, nu: torch.Tensor, solver: DSEM, nsteps: int, dt: float, every: int = 10):
        _check_state(solver, JQ0)
        re = solver.params["re"]
        if abs(float(nu) * re - 1.0) > 1e-12:
            raise ValueError("nu must equal 1/Re of the solver (the context's Re is fixed at creation)")
        every = max(1, int(every))
        torch.cuda.synchronize()
        solver.set_state_device(JQ0.data_ptr())
        cps = [JQ0.detach().clone()]
        for s in range(1, nsteps + 1):
            solver.steps(1, dt, check=False)
            if s % every == 0 and s < nsteps:
                t = torch.empty_like(JQ0)
                solver.get_state_device(t.data_ptr())
                cps.append(t)
        out = torch.empty_like(JQ0)
        solver.get_state_device(out.data_ptr())
        ctx.solver, ctx.nsteps, ctx.dt, ctx.every, ctx.cps = solver, nsteps, dt, every, cps
        return out

    @staticmethod
    def backward(ctx, gout: torch.Tensor):
        solver, n, dt, every = ctx.solver, ctx.nsteps, ctx.dt, ctx.every
        qbar = gout.detach().to(torch.float64).contiguous().clone()
        torch.cuda.synchronize()
        nubar = 0.0
        starts = list(range(0, n, every))
        solver.tape_reserve(min(every, n))
        for si in reversed(range(len(starts))):
            s0, s1 = starts[si], min(starts[si] + every, n)
            solver.set_state_device(ctx.cps[si].data_ptr())
            for i in range(s1 - s0):
                solver.rk_step_record(dt, i)
            for i in reversed(range(s1 - s0)):
                nubar += solver.rk_step_vjp_tape(dt, i, qbar.data_ptr())
        return qbar, torch.tensor(nubar, dtype=torch.float64, device=qbar.device), None, None, None, None


def kinetic_energy_integral(JQ: torch.Tensor, bmat: torch.Tensor) -> torch.Tensor:
    """J = int 1/2 rho |u|^2 dV = 1/2 sum_e sum_pts bmat (JQ_1^2+JQ_2^2+JQ_3^2)/JQ_0
    (JQ = Q / jacob carries the volume factor)."""
    return 0.5 * (bmat * (JQ[:, 1] ** 2 + JQ[:, 2] ** 2 + JQ[:, 3] ** 2) / JQ[:, 0]).sum()


def solve(solver: DSEM, JQ0: torch.Tensor, nsteps: int, dt: float, every: int = 10,
          nu: torch.Tensor | None = None) -> torch.Tensor:
    if nu is None:
        nu = torch.tensor(1.0 / solver.params["re"], dtype=torch.float64, device=JQ0.device)
    return RKSolve.apply(JQ0, nu, solver, nsteps, dt, every)


def bmat_tensor(solver: DSEM, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(solver.quadrature_weights())).to(device)
