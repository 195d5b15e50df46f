"""GPU: the PyTorch modules over the MPM rollout (physgrad_amd.torch_integration).

PhysicsIntegrator against the context it wraps, TorchSimulation's step and
simulate, the guard against a backward after another forward on the same
context, and a training loop that identifies a Young's modulus from a target
end state by gradient descent through the hand-written adjoint.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def blob(rng, n=4096, lo=0.35, size=0.25):
    x = (lo + size * rng.random((n, 3))).astype(np.float32)
    return torch.as_tensor(x, device="cuda")


def rate_field(x, A):
    A = torch.as_tensor(np.asarray(A, np.float32), device="cuda")
    return (x - x.mean(0)) @ A.T


def integrator(n, **kw):
    from physgrad_amd.torch_integration import PhysicsIntegrator
    args = dict(dims=(32, 32, 32), dx=1 / 32, dt=2e-4, youngs_modulus=1e4, poisson_ratio=0.3, density=1000.0,
                particle_volume=0.25 ** 3 / n, max_particles=n)
    args.update(kw)
    return PhysicsIntegrator(**args)


def test_integrator_matches_context_and_differentiates(gpu):
    from physgrad_amd.solver import MPMContext
    rng = np.random.default_rng(70)
    x = blob(rng)
    n = x.shape[0]
    v = rate_field(x, [[1.0, 0.5, 0.0], [-0.5, -0.8, 0.2], [0.1, 0.0, 0.6]])
    integ = integrator(n, steps_per_call=12, checkpoint_every=4)
    ref = MPMContext(integ.params)
    ref.set_particles(x, v)
    ref.step(12)
    want = ref.get_particles()
    xr = x.clone().requires_grad_(True)
    vr = v.clone().requires_grad_(True)
    xo, vo, Co, Fo = integ(xr, vr)
    for got, k in ((xo, "x"), (vo, "v"), (Fo, "F")):
        err = (got - want[k]).abs().max().item() / want[k].abs().max().item()
        assert err <= 1e-5, (k, err)
    (xo.double().mean(0).sum() + 0.1 * (vo.double() ** 2).sum()).backward()
    assert xr.grad is not None and torch.isfinite(xr.grad).all() and vr.grad.abs().max() > 0
    # a second forward before the first backward: the first graph can no longer be reversed
    x1, _, _, _ = integ(x.clone().requires_grad_(True), v)
    integ(x, v)
    with pytest.raises(RuntimeError, match="ran another rollout"):
        x1.sum().backward()
    integ.close()
    ref.close()


def test_torch_simulation_steps(gpu):
    from physgrad_amd.solver import MPMContext
    from physgrad_amd.torch_integration import TorchSimulation
    rng = np.random.default_rng(71)
    x = blob(rng, n=2048)
    n = x.shape[0]
    v = torch.zeros_like(x)
    sim = TorchSimulation(n, integrator=integrator(n, steps_per_call=5))
    sim.set_state(positions=x, velocities=v)
    out = sim.simulate(3, return_trajectory=True)
    assert len(out["trajectory"]) == 3
    ref = MPMContext(sim.integrator.params)
    ref.set_particles(x, v)
    ref.step(15)
    want = ref.get_particles()["x"]
    err = (out["final_positions"] - want).abs().max().item() / want.abs().max().item()
    assert err <= 1e-5
    # free fall away from the walls: every particle moved by the same dt^2 g k (k + 1) / 2
    dy = (out["trajectory"][0]["positions"] - x)[:, 1]
    np.testing.assert_allclose(dy.mean().item(), -9.8 * (2e-4) ** 2 * 15, rtol=1e-2)
    sim.integrator.close()
    ref.close()


def test_training_loop_identifies_youngs_modulus(gpu):
    """Target end state simulated with E* = 1e4 from a deforming initial state;
    starting from E = 4e3, Adam on log E through the adjoint recovers E*."""
    from physgrad_amd.torch_integration import PhysicsBasedLearning
    rng = np.random.default_rng(72)
    x = blob(rng)
    n = x.shape[0]
    vf = rate_field(x, [[2.0, 1.0, 0.0], [-1.0, -1.5, 0.5], [0.3, 0.0, 1.2]])
    integ = integrator(n, steps_per_call=160, checkpoint_every=32)
    with torch.no_grad():
        target, _, _, _ = integ(x, vf, None, None, torch.full((n,), 1e4, device="cuda"))
        target = target.clone()
    model = PhysicsBasedLearning(x, integ, E0=4e3, per_particle_E=False, v_field=vf, learn_velocity=False, lr=0.1)
    losses = []
    for it in range(70):
        if it == 30:   # anneal: smaller Adam steps once near the minimum
            for grp in model.optimizer.param_groups:
                grp["lr"] = 0.01
        losses.append(model.optimization_step(target))
    E = float(torch.exp(model.log_E).item())
    print("losses", losses[0], losses[-1], "E", E, flush=True)
    assert losses[-1] < 1e-3 * losses[0]
    assert abs(E - 1e4) < 0.01 * 1e4
    integ.close()
