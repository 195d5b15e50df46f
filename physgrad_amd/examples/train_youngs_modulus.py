"""Training-loop example: identify an elastic body's Young's modulus from its
observed end state by gradient descent through the MPM adjoint.

A 4,096-particle block on a 32^3 grid starts from a deforming velocity field
and is simulated for 160 steps with the true modulus; a PhysicsBasedLearning
module then fits log E (starting from 40% of the true value) to the observed
final positions with Adam.  Needs a GPU and the built library:

    python -m physgrad_amd.examples.train_youngs_modulus [--iters 70] [--vtk out_dir]
"""
from __future__ import annotations

import argparse
import os

import numpy as np
import torch

from physgrad_amd.state_io import write_vtk
from physgrad_amd.torch_integration import PhysicsBasedLearning, PhysicsIntegrator


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=70)
    ap.add_argument("--true-E", type=float, default=1e4)
    ap.add_argument("--vtk", default=None, help="write the observed and fitted end states here")
    args = ap.parse_args()
    rng = np.random.default_rng(72)
    n = 4096
    x = torch.as_tensor((0.35 + 0.25 * rng.random((n, 3))).astype(np.float32), device="cuda")
    A = torch.tensor([[2.0, 1.0, 0.0], [-1.0, -1.5, 0.5], [0.3, 0.0, 1.2]], device="cuda")
    v_field = (x - x.mean(0)) @ A.T
    integ = PhysicsIntegrator(dims=(32, 32, 32), dx=1 / 32, dt=2e-4, youngs_modulus=args.true_E, poisson_ratio=0.3,
                              density=1000.0, particle_volume=0.25 ** 3 / n, max_particles=n, steps_per_call=160,
                              checkpoint_every=32)
    with torch.no_grad():
        observed, _, _, _ = integ(x, v_field, None, None, torch.full((n,), args.true_E, device="cuda"))
        observed = observed.clone()
    model = PhysicsBasedLearning(x, integ, E0=0.4 * args.true_E, per_particle_E=False, v_field=v_field,
                                 learn_velocity=False, lr=0.1)
    for it in range(args.iters):
        if it == 30:
            for grp in model.optimizer.param_groups:
                grp["lr"] = 0.01
        loss = model.optimization_step(observed)
        if it % 10 == 0 or it == args.iters - 1:
            print(f"iter {it:3d}  loss {loss:.3e}  E {float(torch.exp(model.log_E).item()):.1f}")
    if args.vtk:
        os.makedirs(args.vtk, exist_ok=True)
        with torch.no_grad():
            fitted = model.forward()
        write_vtk(os.path.join(args.vtk, "observed.vtk"), observed.cpu().numpy())
        write_vtk(os.path.join(args.vtk, "fitted.vtk"), fitted.cpu().numpy())
    integ.close()


if __name__ == "__main__":
    main()
