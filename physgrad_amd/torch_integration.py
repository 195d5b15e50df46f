"""PyTorch modules over the differentiable MLS-MPM rollout.

Counterparts of the reference's torch layer, with the MPM transfer loop and
its hand-written adjoint underneath instead of the reference's contact
integrator:

  PhysicsIntegrator     python/physgrad/torch_integration.py:271-357: an
                        nn.Module whose forward advances the particle state
                        (x, v[, C, F]) by `steps_per_call` MLS-MPM steps and is
                        differentiable in x, v, C, F and the per-particle E
  TorchSimulation       :406-502: state holder with set_state / step /
                        simulate(n_steps, return_trajectory)
  tensor_to_physgrad / physgrad_to_tensor   :360-403
  PhysicsBasedLearning  src/pytorch_autograd.h:395-440: a training loop over
                        learnable initial velocity and log Young's moduli, with
                        a real optimizer step (the reference's update is a
                        placeholder constant)

Every module owns one MPM context on the current HIP device; a backward must
run before the next forward on the same module (the context keeps the
checkpoints of its last rollout only; physics_simulation raises otherwise).
"""
from __future__ import annotations

from typing import Optional, Sequence, Tuple

import numpy as np
import torch
from torch import nn

from . import _lib
from .autograd import physics_simulation
from .solver import MPMContext, make_params


def tensor_to_physgrad(tensor: torch.Tensor) -> np.ndarray:
    """Host float64 copy of a tensor (reference :360-371)."""
    return tensor.detach().cpu().numpy().astype(np.float64)


def physgrad_to_tensor(array: np.ndarray, dtype: Optional[torch.dtype] = None, device=None,
                       requires_grad: bool = False) -> torch.Tensor:
    """Tensor from a host array (reference :373-403)."""
    t = torch.from_numpy(np.asarray(array))
    if dtype is not None:
        t = t.to(dtype)
    if device is not None:
        t = t.to(device)
    if requires_grad:
        t.requires_grad_(True)
    return t


class PhysicsIntegrator(nn.Module):
    """Differentiable MLS-MPM integrator: forward(x, v, C=None, F=None, E=None)
    returns (x, v, C, F) after `steps_per_call` steps on the GPU, keeping a
    checkpoint every `checkpoint_every` steps for the adjoint.  The scene
    (grid, material, time step, gravity, walls) is fixed at construction; E
    given per particle overrides youngs_modulus."""

    def __init__(self, dims: Sequence[int] = (64, 64, 64), dx: float = 1.0 / 64, dt: float = 1e-4,
                 gravity: Tuple[float, float, float] = (0.0, -9.8, 0.0), material: int = _lib.MAT_FIXED_COROTATED,
                 youngs_modulus: float = 1e4, poisson_ratio: float = 0.2, density: float = 1000.0,
                 particle_volume: float = 1e-6, boundary_thickness: int = 3, steps_per_call: int = 1,
                 checkpoint_every: int = 32, max_particles: int = 1 << 20, impl: int = _lib.IMPL_TILED,
                 origin: Tuple[float, float, float] = (0.0, 0.0, 0.0)):
        super().__init__()
        self.steps_per_call = int(steps_per_call)
        self.checkpoint_every = int(checkpoint_every)
        self.params = make_params(mode=_lib.MODE_MLS_MPM, impl=impl, dims=dims, dx=dx, origin=origin, dt=dt,
                                  gravity=gravity, material=material, youngs_modulus=youngs_modulus,
                                  poisson_ratio=poisson_ratio, density=density, particle_volume=particle_volume,
                                  boundary_thickness=boundary_thickness, max_particles=max_particles)
        self.register_buffer("gravity", torch.tensor(gravity, dtype=torch.float32))
        self._ctx: Optional[MPMContext] = None

    @property
    def context(self) -> MPMContext:
        if self._ctx is None:
            self._ctx = MPMContext(self.params)
        return self._ctx

    def forward(self, x, v, C=None, F=None, E=None):
        return physics_simulation(self.context, x, v, C, F, E=E, num_steps=self.steps_per_call,
                                  checkpoint_every=min(self.checkpoint_every, self.steps_per_call))

    def close(self):
        if self._ctx is not None:
            self._ctx.close()
            self._ctx = None


class TorchSimulation:
    """Particle state plus an integrator (reference TorchSimulation,
    python/physgrad/torch_integration.py:406-502)."""

    def __init__(self, n_particles: int, integrator: Optional[PhysicsIntegrator] = None, device=None, **kw):
        self.n_particles = int(n_particles)
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.integrator = integrator if integrator is not None else PhysicsIntegrator(max_particles=n_particles, **kw)
        n = self.n_particles
        self.positions = torch.zeros(n, 3, device=self.device)
        self.velocities = torch.zeros(n, 3, device=self.device)
        self.C = torch.zeros(n, 3, 3, device=self.device)
        self.F = torch.eye(3, device=self.device).expand(n, 3, 3).contiguous()
        self.E: Optional[torch.Tensor] = None

    def set_state(self, positions=None, velocities=None, C=None, F=None, E=None):
        if positions is not None:
            self.positions = positions.to(self.device)
        if velocities is not None:
            self.velocities = velocities.to(self.device)
        if C is not None:
            self.C = C.to(self.device)
        if F is not None:
            self.F = F.to(self.device)
        if E is not None:
            self.E = E.to(self.device)

    def step(self):
        self.positions, self.velocities, self.C, self.F = self.integrator(self.positions, self.velocities, self.C,
                                                                          self.F, self.E)
        return self.positions, self.velocities

    def simulate(self, n_steps: int, return_trajectory: bool = True) -> dict:
        """n_steps calls of step(); with return_trajectory the state after each."""
        traj = [] if return_trajectory else None
        for _ in range(int(n_steps)):
            x, v = self.step()
            if return_trajectory:
                traj.append({"positions": x.detach().clone(), "velocities": v.detach().clone()})
        return {"final_positions": self.positions, "final_velocities": self.velocities, "trajectory": traj}


class PhysicsBasedLearning(nn.Module):
    """System identification / control by gradient descent through the MPM
    rollout (reference PhysicsBasedLearning, src/pytorch_autograd.h:395-440).

    Learnable: a shared initial velocity v0 (3) and log Young's moduli (one
    per particle, or one shared).  forward() runs `num_steps` steps from the
    fixed initial positions (plus an optional fixed initial velocity field)
    and returns the final positions; compute_loss() is the mean squared
    distance to target positions; optimization_step() takes one optimizer
    step and returns the loss."""

    def __init__(self, x0: torch.Tensor, integrator: PhysicsIntegrator, E0, per_particle_E: bool = True,
                 v_field: Optional[torch.Tensor] = None, learn_velocity: bool = True, lr: float = 0.05):
        super().__init__()
        n = x0.shape[0]
        self.register_buffer("x0", x0.detach().clone().float())
        self.register_buffer("v_field", torch.zeros_like(self.x0) if v_field is None else v_field.detach().float())
        init = torch.full((n if per_particle_E else 1,), float(np.log(E0)), dtype=torch.float32, device=x0.device)
        self.log_E = nn.Parameter(init)
        self.v0 = nn.Parameter(torch.zeros(3, device=x0.device), requires_grad=learn_velocity)
        self.integrator = integrator
        self.optimizer = torch.optim.Adam([p for p in self.parameters() if p.requires_grad], lr=lr)

    def youngs_modulus(self) -> torch.Tensor:
        return torch.exp(self.log_E).expand(self.x0.shape[0]).contiguous()

    def forward(self) -> torch.Tensor:
        v = self.v0.expand_as(self.x0) + self.v_field
        x, _, _, _ = self.integrator(self.x0, v, None, None, self.youngs_modulus())
        return x

    def compute_loss(self, target_positions: torch.Tensor) -> torch.Tensor:
        d = self.forward().double() - target_positions.double()
        return (d * d).sum(1).mean()

    def optimization_step(self, target_positions: torch.Tensor) -> float:
        self.optimizer.zero_grad()
        loss = self.compute_loss(target_positions)
        loss.backward()
        self.optimizer.step()
        return float(loss.item())
