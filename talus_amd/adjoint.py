"""torch.autograd harness around the differentiable rollout of libtalus_mi.

`rollout(sim, body, E, nu, rest, ...)` runs `steps` converged implicit steps
of one body on the device (talus_mi_rollout, checkpoints every `every`
steps) and returns its final positions and velocities as torch tensors; the
backward pass calls talus_mi_rollout_backward, which recomputes each
checkpoint segment and solves one adjoint system per substep on the GPU.
Per-element Young's modulus and Poisson ratio reach the library as FP64
Lame parameters, so their gradients come from torch's chain rule through
`lame` below; gradients with respect to the rest positions include the rest
inverse, the volumes and the lumped masses.  Starting at rest (the default)
ties the initial positions to `rest`, so their gradient adds to it too.

Tensors are FP64 on the host: the device state lives inside the library.
"""
from __future__ import annotations

import numpy as np
import torch

from . import api


def lame(E: torch.Tensor, nu: torch.Tensor):
    """Lame parameters of the reference (material.cc:71-78): nu clamped to
    [0, 0.49], mu = E / (2 (1 + nu)), lambda = E nu / ((1 + nu)(1 - 2 nu))."""
    nu = nu.clamp(0.0, 0.49)
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return lam, mu


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(fctx, lam, mu, rest, x0, v0, sim, body, params, steps, every, tolerance):
        sim.set_lame(body, lam.detach().cpu().numpy(), mu.detach().cpu().numpy())
        sim.set_rest(body, rest.detach().cpu().numpy())
        n, _ = sim.info(body)
        sim.set_state(body, x0.detach().cpu().numpy(), v0.detach().cpu().numpy(), np.zeros(3 * n))
        fctx.report = sim.rollout(steps, every, params)
        x, v, _ = sim.state(body)
        fctx.sim, fctx.body, fctx.tolerance = sim, body, tolerance
        return torch.from_numpy(x.copy()), torch.from_numpy(v.copy())

    @staticmethod
    def backward(fctx, x_bar, v_bar):
        xb = np.zeros(0) if x_bar is None else x_bar.detach().cpu().numpy()
        vb = None if v_bar is None else v_bar.detach().cpu().numpy()
        n, _ = fctx.sim.info(fctx.body)
        if xb.size == 0:
            xb = np.zeros((n, 3))
        g = fctx.sim.rollout_backward(fctx.body, xb, vb, fctx.tolerance)
        t = torch.from_numpy
        return (t(g["lambda"]), t(g["mu"]), t(g["rest"]), t(g["x0"]), t(g["v0"]),
                None, None, None, None, None, None)


def rollout(sim: api.Context, body: int, E: torch.Tensor, nu: torch.Tensor, rest: torch.Tensor,
            x0: torch.Tensor | None = None, v0: torch.Tensor | None = None, steps: int = 60,
            every: int = 10, params: api.StepParams | None = None, tolerance: float = 1e-12):
    """Final (x, v) of a converged rollout, differentiable in E, nu (per
    tet), rest (per node) and the initial state."""
    lam, mu = lame(E, nu)
    if x0 is None:
        x0 = rest
    if v0 is None:
        v0 = torch.zeros_like(rest)
    if params is None:
        params = api.default_params(preconditioner=api.FACTOR)
    return Rollout.apply(lam, mu, rest, x0, v0, sim, body, params, steps, every, tolerance)
