"""GPU tests of the benchmark workloads at their full sizes.

  two_boxes_128   1,048,576 particles, 128^3, 1,000 steps through the contact
                  of the two boxes: error against the FP64 oracle after 20
                  steps, then conservation of mass and momentum, finite state,
                  every particle inside the grid, and the collision itself.
  rollout_64      262,144 particles, 512 steps, checkpoint every 32: dL/dv0 of
                  the centre-of-mass loss against its closed form (dL/dE is
                  zero: the body falls rigidly, F stays I), and, from a
                  deforming initial state (v0 + A (x - c)), dL/dv0 and dL/dE
                  of 16 sampled particles of a weighted-position loss against
                  FP64 central finite differences (tests/golden/rollout_fd.npz,
                  written by tests/golden/make_rollout_fd.py).
  batch_64        64 scenes (1.23 M particles): one step of every scene against
                  the FP64 oracle, then 500 steps (autotuned launch, hipGraph
                  replay) with per-scene mass conservation and finite state.
  batch_fb_16     16 scenes of 1,048,576 particles on 128^3, 256 steps,
                  checkpoint every 16 (recompute in the backward), centre-of-
                  mass loss per scene: dL/dx0 and dL/dv0 of every particle
                  against the closed form that momentum conservation gives.

Away from the walls the transfers conserve linear momentum exactly (the
stencil is a partition of unity and sum_i w_ip (x_i - x_p) = 0), so after T
steps mean x = mean x0 + T dt mean v0 + dt^2 g T (T + 1) / 2 for every scene
whose particles never reach the sticky band, whatever the elastic forces do.
"""
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def test_two_boxes_fp64_error_and_contact_run(gpu, oracle_lib):
    from physgrad_amd.scenes import two_boxes
    from physgrad_amd.solver import MPMContext
    O = oracle_lib
    sc = two_boxes()
    n, h = sc.n, sc.n // 2
    ctx = MPMContext(sc.params(impl=1))
    ctx.set_particles(sc.x, sc.v)
    ctx.step(20)
    got = {k: t.cpu().numpy() for k, t in ctx.get_particles().items()}
    print("two_boxes: FP64 oracle, 20 steps", flush=True)
    want = O.mls_run(O.scene_params(sc), sc.x.astype(np.float64), sc.v.astype(np.float64), None, None, sc.p_mass,
                     sc.E, 20)
    errs = {k: rel(got[k], want[k]) for k in ("x", "v", "F")}
    # the boxes translate rigidly until they meet, so C is zero up to round-off:
    # its error is taken relative to the velocity-gradient scale max |v| / dx
    errs["C"] = float(np.abs(got["C"] - want["C"]).max() / (np.abs(want["v"]).max() / sc.dx))
    print("two_boxes: rel error after 20 steps vs FP64", errs, flush=True)
    assert errs["x"] <= 1e-6 and errs["v"] <= 1e-5 and errs["F"] <= 1e-6 and errs["C"] <= 1e-5, errs
    # through the contact (the boxes meet near step 500)
    ctx.step(980)
    first_bad, first_out = ctx.check_state()
    assert first_out == -1
    p = ctx.get_particles()
    v = p["v"].double()
    T, dt = 1000, sc.dt
    vbar = v.mean(0).cpu().numpy()
    np.testing.assert_allclose(vbar, [0.0, -9.8 * T * dt, 0.0], atol=2e-4)
    left = v[:h, 0].mean().item()
    print("two_boxes: mean v_x of the left box after 1000 steps", left, flush=True)
    assert left < 1.8   # the boxes have met and are pushing each other back
    m = ctx.get_grid()["mass"].double().sum().item()
    assert abs(m - n * sc.p_mass) <= 1e-5 * n * sc.p_mass
    ctx.close()


def test_rollout_gradients_vs_fp64_finite_differences(gpu):
    from physgrad_amd import _lib
    from physgrad_amd.autograd import physics_simulation
    from physgrad_amd.scenes import rollout
    from physgrad_amd.solver import MPMContext
    fd_path = os.path.join(GOLDEN, "rollout_fd.npz")
    f = np.load(fd_path)
    sc = rollout()
    n, T, K = sc.n, sc.steps, 32
    ctx = MPMContext(sc.params(impl=_lib.IMPL_TILED))
    x0 = torch.as_tensor(sc.x, device="cuda")
    w = torch.as_tensor(np.random.default_rng(int(f["weights_seed"])).normal(size=(n, 3)) / n, device="cuda")
    target = torch.tensor(sc.target, dtype=torch.float64, device="cuda")
    x64 = sc.x.astype(np.float64)
    vdef = torch.as_tensor(((x64 - x64.mean(0)) @ f["A_rate"].T).astype(np.float32), device="cuda")
    for loss_kind in ("com", "weighted"):
        print("rollout loss", loss_kind, flush=True)
        v0 = torch.zeros(3, dtype=torch.float32, device="cuda", requires_grad=True)
        E = torch.as_tensor(sc.E_particles, device="cuda").clone().requires_grad_(True)
        vinit = v0.expand(n, 3) + (vdef if loss_kind == "weighted" else 0.0)
        xT, _, _, _ = physics_simulation(ctx, x0, vinit, E=E, num_steps=T, checkpoint_every=K)
        if loss_kind == "com":
            d = xT.double().mean(0) - target
            loss = (d * d).sum()
        else:
            loss = (w * xT.double()).sum()
        loss.backward()
        gv = v0.grad.double().cpu().numpy()
        gE = E.grad.double().cpu().numpy()
        if loss_kind == "com":
            g = np.array(sc.gravity, np.float64)
            xbar = sc.x.astype(np.float64).mean(0) + sc.dt ** 2 * g * T * (T + 1) / 2
            want = 2.0 * (xbar - np.array(sc.target)) * T * sc.dt
            assert rel(gv, want) <= 1e-4, (gv, want)
            # no wall contact: the elastic moduli cannot move the centre of mass
            assert np.abs(gE * sc.E_particles).max() <= 1e-6 * np.abs(gv).max() * sc.dx
        else:
            print("dv0", gv, "fd", f["dv0"], flush=True)
            assert rel(gv, f["dv0"]) <= 1e-3
            idx = f["E_index"]
            got = gE[idx]
            print("dE", got, "fd", f["dE"], flush=True)
            # per parameter, relative to the largest of the sampled derivatives
            assert np.abs(got - f["dE"]).max() <= 1e-3 * np.abs(f["dE"]).max()
    ctx.close()


def test_batch_64_step_parity_and_500_steps(gpu, oracle_lib):
    from physgrad_amd.scenes import batch_scenes
    from physgrad_amd.solver import MPMBatch
    O = oracle_lib
    scenes = batch_scenes()
    counts = [s.n for s in scenes]
    off = np.concatenate([[0], np.cumsum(counts)])
    x = np.concatenate([s.x for s in scenes])
    v = np.concatenate([s.v for s in scenes])
    b = MPMBatch([s.desc() for s in scenes])
    b.set_particles(counts, x, v)
    b.step(1)
    got = {k: t.cpu().numpy() for k, t in b.get_particles().items()}
    print("batch_64: FP64 oracle, one step of 64 scenes", flush=True)
    for i, s in enumerate(scenes):
        want = O.mls_run(O.scene_params(s), s.x.astype(np.float64), s.v.astype(np.float64), None, None, s.p_mass,
                         s.E, 1)
        for k in ("x", "v", "C", "F"):
            e = rel(got[k][off[i]:off[i + 1]], want[k])
            assert e <= 1e-5, (i, k, e)
    # the rest of the 500 steps with the autotuned launch, replayed from a graph
    b.set_particles(counts, x, v)
    tuned = b.autotune(trial_steps=3, use_cache=False)
    print("batch_64: tile workgroups", tuned, flush=True)
    b.set_graph(True)
    b.step(250)
    b.step(250)
    assert b.check_state() == (-1, -1)
    gm = b.get_grid()["mass"].double().cpu().numpy()
    goff = 0
    for s in scenes:
        G = int(np.prod(s.dims))
        m = gm[goff:goff + G].sum()
        assert abs(m - s.n * s.p_mass) <= 1e-5 * s.n * s.p_mass, s.name
        goff += G
    b.close()


def test_batch_fb_16_closed_form_gradients(gpu):
    from physgrad_amd.scenes import batch_fb
    from physgrad_amd.solver import MPMBatch
    print("batch_fb: building 16 scenes of 1,048,576 particles", flush=True)
    scenes = batch_fb()
    counts = [s.n for s in scenes]
    b = MPMBatch([s.desc() for s in scenes])
    b.set_particles(counts, np.concatenate([s.x for s in scenes]), np.concatenate([s.v for s in scenes]),
                    E=np.concatenate([s.E_particles for s in scenes]))
    T, K = scenes[0].steps, 16
    b.forward_checkpointed(T, K)
    xT = b.get_particles()["x"].double()
    seed = torch.empty_like(xT)
    resid = []
    off = 0
    for s in scenes:
        sl = slice(off, off + s.n)
        d = xT[sl].mean(0) - torch.tensor(s.target, dtype=torch.float64, device="cuda")
        seed[sl] = (2.0 * d / s.n).expand(s.n, 3)
        resid.append(d.cpu().numpy())
        off += s.n
    print("batch_fb: backward (recompute every 16 steps)", flush=True)
    gr = b.backward(ax=seed.float())
    off = 0
    for s, d in zip(scenes, resid):
        sl = slice(off, off + s.n)
        g = np.array(s.gravity, np.float64)
        xbar = s.x.astype(np.float64).mean(0) + T * s.dt * s.v.astype(np.float64).mean(0) + \
            s.dt ** 2 * g * T * (T + 1) / 2
        np.testing.assert_allclose(xT[sl].mean(0).cpu().numpy(), xbar, atol=2e-6)
        want_x = 2.0 * d / s.n
        want_v = want_x * T * s.dt
        ex = rel(gr["x"][sl].cpu().numpy(), np.broadcast_to(want_x, (s.n, 3)))
        ev = rel(gr["v"][sl].cpu().numpy(), np.broadcast_to(want_v, (s.n, 3)))
        eE = float((gr["E"][sl].double().abs() * torch.as_tensor(s.E_particles, device="cuda")).max().item())
        print(s.name, "rel err dL/dx0", ex, "dL/dv0", ev, "max |E dL/dE|", eE, flush=True)
        assert ex <= 1e-3 and ev <= 1e-3, (s.name, ex, ev)
        assert eE <= 1e-3 * np.abs(want_x).max() * s.dx
        off += s.n
    b.close()


def test_full_size_adjoint_is_linear_in_its_seed(gpu):
    """The reverse pass is linear in the adjoint seed whatever the particles
    do: at full size (batch_fb_0, 1,048,576 particles, 48 steps with a
    checkpoint every 16, so two recomputed segments and two segment
    boundaries), backward(a s1 + b s2) must equal a backward(s1) +
    b backward(s2) for random seeds on every state component and on E.  Lost
    or doubled adjoint contributions in the fused reverse kernel or at the
    segment boundaries break this beyond round-off."""
    from physgrad_amd.scenes import batch_fb
    from physgrad_amd.solver import MPMContext
    sc = batch_fb(nscenes=1)[0]
    n = sc.n
    ctx = MPMContext(sc.params(impl=1))
    ctx.set_particles(sc.x, sc.v, E=sc.E_particles)
    ctx.forward_checkpointed(48, 16)
    g = torch.Generator(device="cuda").manual_seed(53)
    shapes = ((n, 3), (n, 3), (n, 9), (n, 9))
    s1 = [torch.randn(sh, device="cuda", generator=g) / n for sh in shapes]
    s2 = [torch.randn(sh, device="cuda", generator=g) / n for sh in shapes]
    a, b = 0.7, -1.3
    g1 = ctx.backward(*s1)
    g2 = ctx.backward(*s2)
    g12 = ctx.backward(*[a * u + b * w for u, w in zip(s1, s2)])
    for key in ("x", "v", "C", "F", "E"):
        want = (a * g1[key].double() + b * g2[key].double()).cpu().numpy()
        err = rel(g12[key].double().cpu().numpy(), want)
        print("linearity", key, err, flush=True)
        assert err < 1e-4, (key, err)
    ctx.close()
