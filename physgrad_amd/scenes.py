"""Synthetic workloads of the MLS-MPM transfer loop (seeded, reproducible).

    single_box        64^3, dx 1/64, 32,768 particles uniform in
                      [0.3,0.5]^3 (seed 1), fixed-corotated E 1e4, nu 0.2, rho 1,
                      v (0,-1,0), g -9.8, dt 1e-4, sticky walls 3 nodes, 200 steps
    two_boxes         128^3, dx 1/128, 1,048,576 particles, half in
                      [0.2,0.45]^3 and half in [0.55,0.8]x[0.2,0.45]^2 (seed 2),
                      v = +-2 along x (towards each other), neo-Hookean E 5e4,
                      nu 0.3, dt 5e-5, 1000 steps.  The density is not given by
                      the workload; rho = 1000 keeps the elastic CFL number
                      c dt/dx = 0.05 (rho = 1 would give 1.66, an unstable run).
    rollout           64^3, 262,144 particles in a seeded random box
                      (seed 3), fixed-corotated, nu 0.2, per-particle E
                      log-uniform in [5e3,5e4], rho 1000 (unstated; CFL 0.05),
                      dt 1e-4, 512 steps; loss = |mean x - (0.6,0.3,0.5)|^2.
    batch_scenes      64 independent scenes (seed 4): grid side in {32,48,64}
                      (dx = 1/side), log-uniform 4,096..65,536 particles in one
                      seeded random box per scene, fixed-corotated or
                      neo-Hookean per scene, E log-uniform in [5e3,5e4],
                      nu 0.2..0.35, rho 1000, dt 1e-4, 500 steps, all advanced
                      by the same launches.  The workload fixes neither box
                      size nor density: each box holds 8 particles per cell
                      (2 per axis, the usual MPM sampling) with aspect ratios
                      in [0.7, 1.4], placed at random inside the walls.
    batch_fb          16 independent scenes (seed 5) of 128^3 (dx 1/128) with
                      1,048,576 particles each, uniform in one seeded box per
                      scene, initial velocity N(0, 0.5^2) per component,
                      neo-Hookean, nu 0.3, per-particle E log-uniform in
                      [5e3,5e4], rho 1000, dt 5e-5, 256 steps, centre-of-mass
                      loss per scene.  The workload fixes neither box size nor
                      target: each box holds 8 particles per cell with aspect
                      ratios in [0.8, 1.25], at least 0.12 from every wall
                      (about 15 cells; the particles move less than one cell
                      in 256 steps), and the target is the box centre moved by
                      (0.05, -0.03, 0.02).
Particle volume V0 = box volume / N; mass = rho V0.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class Scene:
    name: str
    dims: tuple
    dx: float
    dt: float
    steps: int
    material: int           # 0 fixed-corotated, 1 neo-Hookean
    E: float
    nu: float
    rho: float
    p_vol: float
    x: np.ndarray
    v: np.ndarray
    gravity: tuple = (0.0, -9.8, 0.0)
    bc: int = 3
    origin: tuple = (0.0, 0.0, 0.0)
    E_particles: Optional[np.ndarray] = None
    target: Optional[tuple] = None

    @property
    def n(self) -> int:
        return self.x.shape[0]

    @property
    def p_mass(self) -> float:
        return self.rho * self.p_vol

    def desc(self, max_particles: Optional[int] = None):
        """pg_scene_desc of this scene (one entry of a batch context)."""
        from .solver import scene_desc
        return scene_desc(self.dims, self.dx, self.origin, self.dt, self.gravity, self.material, self.E, self.nu,
                          self.rho, self.p_vol, self.bc, int(max_particles or max(self.n, 1)))

    def params(self, impl: int = 1, max_particles: Optional[int] = None):
        from .solver import make_params
        from . import _lib
        return make_params(mode=_lib.MODE_MLS_MPM, impl=impl, dims=self.dims, dx=self.dx, origin=self.origin,
                           dt=self.dt, gravity=self.gravity, material=self.material, youngs_modulus=self.E,
                           poisson_ratio=self.nu, density=self.rho, particle_volume=self.p_vol,
                           boundary_thickness=self.bc, max_particles=int(max_particles or self.n))


def _uniform(rng, n, lo, hi):
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    return (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)


def single_box(n: int = 32768, seed: int = 1) -> Scene:
    rng = np.random.default_rng(seed)
    x = _uniform(rng, n, (0.3,) * 3, (0.5,) * 3)
    v = np.tile(np.array([0.0, -1.0, 0.0], np.float32), (n, 1))
    return Scene("single_box_64", (64, 64, 64), 1 / 64, 1e-4, 200, 0, 1e4, 0.2, 1.0, 0.2 ** 3 / n, x, v)


def two_boxes(n: int = 1 << 20, seed: int = 2, grid: int = 128) -> Scene:
    rng = np.random.default_rng(seed)
    h = n // 2
    a = _uniform(rng, h, (0.2, 0.2, 0.2), (0.45, 0.45, 0.45))
    b = _uniform(rng, n - h, (0.55, 0.2, 0.2), (0.8, 0.45, 0.45))
    x = np.concatenate([a, b])
    v = np.zeros((n, 3), np.float32)
    v[:h, 0] = 2.0
    v[h:, 0] = -2.0
    vol = 2 * 0.25 ** 3
    return Scene("two_boxes_128", (grid,) * 3, 1 / grid, 5e-5, 1000, 1, 5e4, 0.3, 1000.0, vol / n, x, v)


def rollout(n: int = 262144, seed: int = 3) -> Scene:
    rng = np.random.default_rng(seed)
    size = 0.15 + 0.1 * rng.random(3)
    lo = 0.25 + (0.5 - size) * rng.random(3) * 0.5
    x = _uniform(rng, n, lo, lo + size)
    v = np.zeros((n, 3), np.float32)
    E = np.exp(rng.uniform(np.log(5e3), np.log(5e4), n)).astype(np.float32)
    s = Scene("rollout_64", (64, 64, 64), 1 / 64, 1e-4, 512, 0, 1e4, 0.2, 1000.0, float(np.prod(size)) / n, x, v,
              E_particles=E, target=(0.6, 0.3, 0.5))
    return s


def batch_scenes(nscenes: int = 64, seed: int = 4, min_n: int = 4096, max_n: int = 65536,
                 sides=(32, 48, 64), ppc: float = 8.0) -> list:
    """Independent scenes of a batch (see the module docstring)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(nscenes):
        side = int(sides[rng.integers(len(sides))])
        n = int(np.exp(rng.uniform(np.log(min_n), np.log(max_n))))
        dx = 1.0 / side
        aspect = rng.uniform(0.7, 1.4, 3)
        aspect /= np.prod(aspect) ** (1.0 / 3.0)
        size = np.minimum(aspect * (n * dx ** 3 / ppc) ** (1.0 / 3.0), 0.7)
        lo = 0.15 + (0.7 - size) * rng.random(3)
        x = _uniform(rng, n, lo, lo + size)
        v = rng.normal(0.0, 0.5, (n, 3)).astype(np.float32)
        material = int(rng.integers(2))
        E = float(np.exp(rng.uniform(np.log(5e3), np.log(5e4))))
        nu = float(rng.uniform(0.2, 0.35))
        out.append(Scene(f"batch_{s}_{side}", (side,) * 3, dx, 1e-4, 500, material, E, nu, 1000.0,
                         float(np.prod(size)) / n, x, v))
    return out


def batch_fb(nscenes: int = 16, n: int = 1 << 20, seed: int = 5, grid: int = 128, ppc: float = 8.0) -> list:
    """Scenes of the batched forward+backward workload (see the module docstring)."""
    rng = np.random.default_rng(seed)
    out = []
    dx = 1.0 / grid
    for s in range(nscenes):
        aspect = rng.uniform(0.8, 1.25, 3)
        aspect /= np.prod(aspect) ** (1.0 / 3.0)
        size = np.minimum(aspect * (n * dx ** 3 / ppc) ** (1.0 / 3.0), 0.76)
        lo = 0.12 + (0.76 - size) * rng.random(3)
        x = _uniform(rng, n, lo, lo + size)
        v = rng.normal(0.0, 0.5, (n, 3)).astype(np.float32)
        E = np.exp(rng.uniform(np.log(5e3), np.log(5e4), n)).astype(np.float32)
        target = tuple(float(c) for c in (lo + 0.5 * size + np.array([0.05, -0.03, 0.02])))
        out.append(Scene(f"batch_fb_{s}", (grid,) * 3, dx, 5e-5, 256, 1, 2e4, 0.3, 1000.0,
                         float(np.prod(size)) / n, x, v, E_particles=E, target=target))
    return out
