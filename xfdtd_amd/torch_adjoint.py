"""PyTorch view of the design gradient.

`DesignProblem(scene, total_steps)` builds the engine once; calling it with
a permittivity tensor of the design box's shape returns J = sum |F|^2 of the
scene's DFT plane after `total_steps` steps, differentiable with respect to
that tensor.  Each call uploads the new design coefficients
(xf_sim_update_design), rewinds the engine to a checkpoint of step 0 and
runs xf_engine_grad_eps (forward run with checkpoints, then the discrete
adjoint), so one forward + backward costs one gradient run and no set-up.
`design_objective(eps_r, scene, total_steps)` is the one-shot form: it
builds and frees an engine per call.
"""
from __future__ import annotations

import copy

import numpy as np
import torch

from . import scene as S

_START = 0  # checkpoint slot of the step-0 state


class _DesignObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps_r, problem, checkpoint_every):
        sim = problem.sim
        sim.update_design(eps_r.detach().to("cpu", torch.float32).contiguous().numpy())
        sim.restore(_START)
        if ctx.needs_input_grad[0]:
            J, g = sim.grad(checkpoint_every)
            ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(g)).to(eps_r.device, eps_r.dtype))
        else:
            sim.step(problem.total_steps)
            J = float(np.sum(np.abs(sim.get_dft()) ** 2))
        return torch.tensor(J, dtype=torch.float64, device=eps_r.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return (grad_out.to(g.dtype) * g).reshape(g.shape), None, None


class DesignProblem:
    """J(eps_r) for `scene` (which carries the design box and DFT plane) on
    one engine kept across calls."""

    def __init__(self, scene, total_steps, precision=64, checkpoint_every=0):
        from .sim import Simulation
        b = scene.design.box
        self.shape = (b[1] - b[0], b[3] - b[2], b[5] - b[4])
        self.total_steps = int(total_steps)
        self.checkpoint_every = int(checkpoint_every)
        self.sim = Simulation(scene, self.total_steps, precision=precision)
        self.sim.checkpoint(_START)

    def __call__(self, eps_r):
        if tuple(eps_r.shape) != self.shape:
            raise ValueError(f"eps_r has shape {tuple(eps_r.shape)}, the design box {self.shape}")
        return _DesignObjective.apply(eps_r, self, self.checkpoint_every)

    def close(self):
        self.sim.close()


def design_objective(eps_r, scene, total_steps, precision=64, checkpoint_every=0):
    """J(eps_r) for `scene` on an engine built and freed for this call."""
    sc = copy.copy(scene)
    sc.design = S.Design(scene.design.box, np.asarray(scene.design.eps_r, dtype=np.float32))
    problem = DesignProblem(sc, total_steps, precision, checkpoint_every)
    try:
        return problem(eps_r)
    finally:
        problem.close()
