"""torch.autograd.Function over the checkpointed MLS-MPM rollout.

Replaces the reference's PhysicsSimulationFunction / physics_simulation
(src/pytorch_autograd.h:144-281), whose backward runs a spring-chain adjoint,
and the launch_mpm_timestep_forward/backward slot (src/pytorch_autograd.cu:
371-402).  Here forward runs nsteps MLS-MPM steps on the GPU keeping a
checkpoint every K steps, and backward runs the hand-written adjoint of the
same steps (recompute per segment, reverse G2P / grid / P2G), all through the
C-ABI on torch's current HIP stream.
"""
from __future__ import annotations

import torch

from .solver import MPMContext


class PhysicsSimulationFunction(torch.autograd.Function):
    """(x0, v0, C0, F0, E) -> (x, v, C, F) after nsteps steps of `mpm`."""

    @staticmethod
    def forward(ctx, x0, v0, C0, F0, E, mpm: MPMContext, nsteps: int, checkpoint_every: int):
        n = x0.shape[0]
        mpm.set_particles(x0.detach(), v0.detach(), C0.detach().reshape(n, 9), F0.detach().reshape(n, 9),
                          E=None if E is None else E.detach())
        mpm.forward_checkpointed(nsteps, checkpoint_every)
        out = mpm.get_particles()
        ctx.mpm = mpm
        ctx.rollout = mpm.rollouts
        ctx.has_E = E is not None
        return out["x"], out["v"], out["C"], out["F"]

    @staticmethod
    def backward(ctx, gx, gv, gC, gF):
        # the context keeps the checkpoints of its last rollout only
        if ctx.mpm.rollouts != ctx.rollout:
            raise RuntimeError("physics_simulation: the MPM context ran another rollout since this one; "
                               "call backward before the next forward on the same context, or use one context "
                               "per rollout")
        n = gx.shape[0] if gx is not None else ctx.mpm.num_particles
        grads = ctx.mpm.backward(gx, gv, None if gC is None else gC.reshape(n, 9),
                                 None if gF is None else gF.reshape(n, 9))
        gE = grads["E"] if ctx.has_E else None
        return grads["x"], grads["v"], grads["C"], grads["F"], gE, None, None, None


def physics_simulation(mpm: MPMContext, x0, v0, C0=None, F0=None, E=None, num_steps: int = 100,
                       checkpoint_every: int = 32):
    """Differentiable rollout of `num_steps` MLS-MPM steps (reference name:
    physics_simulation, src/pytorch_autograd.h:266-281).  C0 defaults to 0,
    F0 to I; E (per-particle Young's modulus) is optional."""
    n = x0.shape[0]
    dev = x0.device
    if C0 is None:
        C0 = torch.zeros((n, 3, 3), dtype=torch.float32, device=dev)
    if F0 is None:
        F0 = torch.eye(3, dtype=torch.float32, device=dev).expand(n, 3, 3).contiguous()
    return PhysicsSimulationFunction.apply(x0, v0, C0, F0, E, mpm, int(num_steps), int(checkpoint_every))
