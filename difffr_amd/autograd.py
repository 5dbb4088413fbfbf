"""torch.autograd.Function over the exact adjoint of libdfr_hip.so.

`RigidRollout.apply(v0w0, sim, body, steps, checkpoint_every)` maps the release
velocity (init_v_rb, init_omega_rb) of one dynamic body to its pose (x_rb, q)
after `steps` steps from the scene's initial state; backward runs the
library's reverse pass (dfr_adjoint_backward), which re-runs the rollout from
checkpoints taken every `checkpoint_every` steps.  This is the loss path of
DiffFR's trajectory optimisation drivers (gradient-based-optimize.py:90-358),
with the exact gradient in place of the forward-mode approximation.
"""
from __future__ import annotations

import numpy as np
import torch

from .simulation import Simulation


class RigidRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0w0: torch.Tensor, sim: Simulation, body: int, steps: int, checkpoint_every: int = 1):
        p = v0w0.detach().to("cpu", torch.float64).numpy()
        sim.reset()
        sim.set_init_velocity(body, p[:3], p[3:6])
        sim.adjoint_begin(checkpoint_every)
        sim.step(steps)
        r = sim.rigid(body)
        ctx.sim = sim
        ctx.body = body
        out = np.concatenate([r["x"], r["q"]])
        return torch.as_tensor(out, dtype=v0w0.dtype, device=v0w0.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        sim: Simulation = ctx.sim
        seed = np.zeros((sim.n_bodies, 7))
        seed[ctx.body] = grad_out.detach().to("cpu", torch.float64).numpy()
        g = sim.adjoint_backward(seed)[ctx.body]
        return torch.as_tensor(g, dtype=grad_out.dtype, device=grad_out.device), None, None, None, None


def rigid_rollout(v0w0: torch.Tensor, sim: Simulation, body: int, steps: int, checkpoint_every: int = 1):
    """Final (x_rb[3], q[4] (w,x,y,z)) of `body` as a differentiable function of v0w0[6]."""
    return RigidRollout.apply(v0w0, sim, body, steps, checkpoint_every)
