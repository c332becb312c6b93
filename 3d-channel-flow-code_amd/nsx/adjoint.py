"""torch.autograd harness of the differentiation form.

`WallShear` is a torch.autograd.Function whose forward runs the channel
steps on the device with a per-prism streamwise body force and returns the
time-averaged bottom-wall shear J (nsx_wall_shear_run), and whose backward
returns dJ/d(force) from the hand-written adjoint kernels (the same entry
point with a gradient field: checkpoints every `every` steps, segments
recomputed, reversed RK2 stages and adjoint pressure solves on A^T).  The
force is a float64 tensor over the interior prisms, shape (N_CELL0, NZ-2).
The forward and backward always start from the case's initial field, so a
call is a pure function of the force.
"""
from __future__ import annotations

import numpy as np
import torch


class WallShearProblem:
    """One device context of a case and the objective's settings."""

    def __init__(self, case, nsteps: int, every: int = 5, eps: float | None = None):
        from .device import Context

        self.case = case
        self.nsteps = int(nsteps)
        self.every = int(every)
        self.ctx = Context(case)
        if eps is not None:
            self.ctx.set_params(eps=eps)

    def _start(self, force: np.ndarray) -> None:
        self.ctx.set_body_force(force)
        self.ctx.init_state()

    def objective(self, force: np.ndarray) -> float:
        self._start(force)
        return self.ctx.wall_shear(self.nsteps, self.every)

    def gradient(self, force: np.ndarray):
        self._start(force)
        return self.ctx.wall_shear(self.nsteps, self.every, grad=True)

    def tangent(self, force: np.ndarray, direction: np.ndarray):
        self._start(force)
        return self.ctx.wall_shear_tangent(self.nsteps, direction)

    def close(self) -> None:
        self.ctx.close()


class WallShear(torch.autograd.Function):
    """J(force) with the adjoint gradient as its backward."""

    @staticmethod
    def forward(ctx, force: torch.Tensor, problem: WallShearProblem) -> torch.Tensor:
        f = force.detach().to("cpu", torch.float64).numpy()
        ctx.problem = problem
        ctx.save_for_backward(force)
        return torch.tensor(problem.objective(f), dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ: torch.Tensor):
        (force,) = ctx.saved_tensors
        _, g = ctx.problem.gradient(force.detach().to("cpu", torch.float64).numpy())
        return gJ * torch.from_numpy(g).to(force.device, force.dtype), None


def wall_shear(force: torch.Tensor, problem: WallShearProblem) -> torch.Tensor:
    return WallShear.apply(force, problem)
