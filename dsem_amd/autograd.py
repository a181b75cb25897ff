"""PyTorch autograd binding of the GPU path.

``RHS``     -- Qt = R(JQ) as a differentiable function of the state (forward
               dsem_rhs_device, backward dsem_rhs_vjp_device).
``RKSolve`` -- ``nsteps`` RK4(5) steps as a differentiable function of the
               initial state and of nu = 1/Re (forward dsem_rk_steps,
               backward by step adjoints).  Checkpointing is two-level and
               scheduled binomially (dsem_amd.checkpoint, revolve's optimum):
               ``snaps`` state snapshots, and the library's step tape of
               ``tape`` slots (dsem_rk_step_record stores a step's stage
               states and face intermediates; dsem_rk_step_vjp_tape runs its
               adjoint).  The forward pass keeps the snapshots the schedule
               places on its way and records the last steps, which the
               reversal takes first, on the tape as it takes them; the
               backward pass restores, advances, snapshots and
               records/reverses as the schedule says.  The
               default (tape = every, snaps = ceil(nsteps/every) - 1) is the
               uniform schedule: a snapshot every ``every`` steps and one
               extra forward pass; fewer slots trade memory for
               recomputation at the binomial optimum.

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
    def forward(ctx, JQ0: torch.Tensor, nu: torch.Tensor, solver: DSEM, nsteps: int, dt: float, every: int = 10,
                snaps: int | None = None, tape: int | None = None):
        from .checkpoint import Plan
        _check_state(solver, JQ0)
        re = solver.params["re"]
        if abs(float(nu) * re - 1.0) > 1e-12:
            raise ValueError("nu must equal 1/Re of the solver (the context's Re is fixed at creation)")
        every = max(1, int(every))
        tape = max(1, min(int(tape if tape is not None else every), nsteps))
        if snaps is None:
            snaps = -(-nsteps // every) - 1
        plan = Plan(nsteps, max(0, int(snaps)), tape)
        keep = set(plan.forward_snapshots())
        # The reversal starts with the solve's last steps [tail, nsteps): the
        # forward pass records them on the tape as it takes them, so their
        # recording pass (and the restore/advance that would precede it) is
        # not repeated.
        tail = plan.tail()
        # entropy viscosity and the Smagorinsky model read the lagged
        # staggered states: snapshots and the adjoint state carry them
        lagged = bool(solver.params["shock"]) or solver.les
        torch.cuda.synchronize()
        solver.tape_reserve(plan.tape)
        solver.set_state_device(JQ0.data_ptr())      # a fresh start: the first step's mu_h is zero
        held = {0: _snapshot(solver, JQ0.detach().clone(), lagged, initial=True)}
        for s in range(1, nsteps + 1):
            if s - 1 >= tail:
                solver.rk_step_record(dt, s - 1 - tail)
            else:
                solver.steps(1, dt, check=False)
            if s in keep:
                held[s] = _snapshot(solver, None, lagged)
        out = torch.empty_like(JQ0)
        solver.get_state_device(out.data_ptr())
        ctx.token = solver.tape_owner = object()     # the tail on the tape is this solve's
        ctx.solver, ctx.dt, ctx.plan, ctx.held, ctx.lagged = solver, dt, plan, held, lagged
        return out

    @staticmethod
    def backward(ctx, gout: torch.Tensor):
        solver, dt, plan, held, lagged = ctx.solver, ctx.dt, ctx.plan, ctx.held, ctx.lagged
        qbar = gout.detach().to(torch.float64).contiguous().clone()
        # with entropy viscosity or the Smagorinsky model the adjoint state
        # also carries the lagged staggered states' adjoint (zero after the
        # last step: J reads the state only)
        slagbar = torch.zeros((solver.K, 3, 5, solver.ns), dtype=torch.float64, device=qbar.device) if lagged else None
        torch.cuda.synchronize()
        nubar = 0.0
        solver.tape_reserve(plan.tape)
        # another recording since the forward pass (a second solve on the
        # same context) replaced the tail: then it is recorded again
        acts = plan.backward_actions() if getattr(solver, "tape_owner", None) is ctx.token else plan.actions()
        for act in acts:
            kind = act[0]
            if kind == "restore":
                _restore(solver, held[act[1]], lagged)
            elif kind == "advance":
                solver.steps(act[1], dt, check=False)
            elif kind == "snapshot":
                held[act[1]] = _snapshot(solver, None, lagged)
            elif kind == "free":
                del held[act[1]]
            else:   # reverse: record k steps from the current state (unless the forward
                    # pass did), then their adjoints backwards
                k = act[2]
                if kind == "reverse":
                    for i in range(k):
                        solver.rk_step_record(dt, i)
                for i in reversed(range(k)):
                    if lagged:
                        nubar += solver.rk_step_vjp_tape_lagged(dt, i, qbar.data_ptr(), slagbar.data_ptr())
                    else:
                        nubar += solver.rk_step_vjp_tape(dt, i, qbar.data_ptr())
        ctx.held = None
        return qbar, torch.tensor(nubar, dtype=torch.float64, device=qbar.device), None, None, None, None, None, None


def _snapshot(solver: DSEM, state, lagged: bool, initial: bool = False):
    """The solver's current state (or the given tensor) and, when a model
    reads them (entropy viscosity, Smagorinsky), its lagged staggered states
    (None on a fresh start)."""
    if state is None:
        state = torch.empty((solver.K, 5, solver.np), dtype=torch.float64, device="cuda")
        solver.get_state_device(state.data_ptr())
    if not lagged:
        return state
    lag = None
    if not initial:
        lag = torch.empty((solver.K, 3, 5, solver.ns), dtype=torch.float64, device=state.device)
        if not solver.get_lagged_states_device(lag.data_ptr()):
            lag = None
    return (state, lag)


def _restore(solver: DSEM, snap, lagged: bool) -> None:
    if not lagged:
        solver.set_state_device(snap.data_ptr())
        return
    state, lag = snap
    solver.set_state_device(state.data_ptr())
    solver.set_lagged_states_device(None if lag is None else lag.data_ptr())


def kinetic_energy_integral(JQ: torch.Tensor, bmat: torch.Tensor) -> torch.Tensor:
    """J = int 1/2 rho |u|^2 dV = 1/2 sum_e sum_pts bmat (JQ_1^2+JQ_2^2+JQ_3^2)/JQ_0
    (JQ = Q / jacob carries the volume factor)."""
    return 0.5 * (bmat * (JQ[:, 1] ** 2 + JQ[:, 2] ** 2 + JQ[:, 3] ** 2) / JQ[:, 0]).sum()


def solve(solver: DSEM, JQ0: torch.Tensor, nsteps: int, dt: float, every: int = 10,
          nu: torch.Tensor | None = None, snaps: int | None = None, tape: int | None = None) -> torch.Tensor:
    """``nsteps`` differentiable RK steps from JQ0.  Checkpointing: ``snaps``
    snapshots and a ``tape``-slot step tape, scheduled binomially; by default
    a snapshot every ``every`` steps (uniform)."""
    if nu is None:
        nu = torch.tensor(1.0 / solver.params["re"], dtype=torch.float64, device=JQ0.device)
    return RKSolve.apply(JQ0, nu, solver, nsteps, dt, every, snaps, tape)


def bmat_tensor(solver: DSEM, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(solver.quadrature_weights())).to(device)
