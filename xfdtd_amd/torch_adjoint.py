"""PyTorch view of the design gradient.

`design_objective(eps_r, scene, total_steps)` returns J = sum |F|^2 of the
scene's DFT plane after `total_steps` steps with the design box's
permittivity set to `eps_r` (a tensor of the box's shape).  Its backward
multiplies the incoming gradient by dJ/d eps_r, which the forward obtains in
the same engine call (xf_engine_grad_eps: forward run with checkpoints, then
the discrete adjoint), so one forward + backward costs one gradient run.
"""
from __future__ import annotations

import copy

import numpy as np
import torch

from . import scene as S


class _DesignObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps_r, scene, total_steps, precision, checkpoint_every):
        from .sim import Simulation
        sc = copy.copy(scene)
        e = eps_r.detach().to("cpu", torch.float32).contiguous().numpy()
        sc.design = S.Design(scene.design.box, e.ravel())
        sim = Simulation(sc, total_steps, precision=precision)
        try:
            if ctx.needs_input_grad[0]:
                J, g = sim.grad(checkpoint_every)
                ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(g)).to(eps_r.device, eps_r.dtype))
            else:
                sim.step(total_steps)
                J = float(np.sum(np.abs(sim.get_dft()) ** 2))
        finally:
            sim.close()
        return torch.tensor(J, dtype=torch.float64, device=eps_r.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return (grad_out.to(g.dtype) * g).reshape(g.shape), None, None, None, None


def design_objective(eps_r, scene, total_steps, precision=64, checkpoint_every=0):
    """J(eps_r) for `scene` (which carries the design box and DFT plane)."""
    b = scene.design.box
    shape = (b[1] - b[0], b[3] - b[2], b[5] - b[4])
    if tuple(eps_r.shape) != shape:
        raise ValueError(f"eps_r has shape {tuple(eps_r.shape)}, the design box {shape}")
    return _DesignObjective.apply(eps_r, scene, int(total_steps), int(precision), int(checkpoint_every))
