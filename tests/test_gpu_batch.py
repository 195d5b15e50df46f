"""Batched scenes, hipGraph replay and the launch autotuner of the tiled
MLS-MPM path.

Bars: one step of every scene of a batch against the FP64 oracle, inf-norm
relative 1e-5 per field (the single-scene bar); multi-step runs against
separate single-scene contexts and eager against graph replay, relative 1e-4
(FP32 trajectories whose atomics sum in different orders); the batch adjoint
against the FP64 oracle's reverse steps per scene, relative 1e-3.
"""
import numpy as np
import pytest
import torch

from test_gpu_mls import check_grid

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-5


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def mixed_scenes(seed=11):
    """Five scenes with different grids (16^3, 24^3 with a non-power-of-two
    tile count, 32^3, 20x28x16, an empty one), cell sizes, origins, gravity,
    materials and elastic constants."""
    from physgrad_amd.scenes import Scene
    rng = np.random.default_rng(seed)
    specs = [((16, 16, 16), 1 / 16, (0, 0, 0), (0, -9.8, 0), 0, 2e4, 0.3, 700),
             ((24, 24, 24), 1 / 24, (0.1, -0.2, 0.05), (1.0, -5.0, 0.5), 1, 5e4, 0.25, 1500),
             ((32, 32, 32), 1 / 32, (0, 0, 0), (0, -9.8, 0), 1, 8e3, 0.35, 2048),
             ((20, 28, 16), 1 / 20, (0, 0, 0), (0, 0, -9.8), 0, 1.2e4, 0.2, 900),
             ((16, 16, 16), 1 / 16, (0, 0, 0), (0, -9.8, 0), 0, 1e4, 0.3, 0)]
    out = []
    for dims, dx, org, grav, mat, E, nu, n in specs:
        ext = np.array(dims) * dx
        lo = np.array(org) + 0.3 * ext
        x = (lo + 0.4 * ext * rng.random((n, 3))).astype(np.float32)
        v = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
        vol = float(np.prod(0.4 * ext)) / max(n, 1)
        out.append(Scene("mixed", dims, dx, 2e-4, 1, mat, E, nu, 1000.0, vol, x, v, gravity=grav, origin=org))
    return out


def concat(scenes):
    x = np.concatenate([s.x for s in scenes])
    v = np.concatenate([s.v for s in scenes])
    return x, v


def make_batch(scenes, spare=0):
    from physgrad_amd.solver import MPMBatch
    return MPMBatch([s.desc(max_particles=s.n + spare) for s in scenes])


def test_batch_step_parity_vs_fp64_oracle(gpu, oracle_lib):
    O = oracle_lib
    scenes = mixed_scenes()
    rng = np.random.default_rng(5)
    counts = [s.n for s in scenes]
    N = sum(counts)
    C = rng.normal(0, 0.5, (N, 9)).astype(np.float32)
    F = (np.eye(3).reshape(1, 9) + rng.normal(0, 0.05, (N, 9))).astype(np.float32)
    E = np.exp(rng.uniform(np.log(5e3), np.log(5e4), N)).astype(np.float32)
    x, v = concat(scenes)
    for per_particle_E in (False, True):
        print("batch parity, per-particle E:", per_particle_E, flush=True)
        b = make_batch(scenes, spare=7)
        assert b.num_scenes == len(scenes)
        b.set_particles(counts, x, v, C, F, E=E if per_particle_E else None)
        b.step(1)
        got = {k: t.cpu().numpy().reshape(N, -1) for k, t in b.get_particles().items()}
        grid = b.get_grid()
        gm = grid["mass"].cpu().numpy()
        gv = grid["velocity"].cpu().numpy()
        p0 = g0 = 0
        for s, sc in enumerate(scenes):
            n = sc.n
            G = int(np.prod(sc.dims))
            Es = E[p0:p0 + n].astype(np.float64) if per_particle_E else sc.E
            want = O.mls_run(O.scene_params(sc), x[p0:p0 + n].astype(np.float64), v[p0:p0 + n].astype(np.float64),
                             C[p0:p0 + n].astype(np.float64), F[p0:p0 + n].astype(np.float64), sc.p_mass, Es, 1,
                             want_grid=True)
            if n:
                for key in ("x", "v", "C", "F"):
                    err = rel(got[key][p0:p0 + n].reshape(-1), want[key].reshape(-1))
                    assert err <= TOL_STEP, (s, key, err)
                check_grid(gm[g0:g0 + G], gv[g0:g0 + G], want["grid_m"], want["grid_v"])
            else:
                assert not gm[g0:g0 + G].any()
            p0 += n
            g0 += G
        b.close()


def single_run(sc, steps, graph=False):
    from physgrad_amd.solver import MPMContext
    ctx = MPMContext(sc.params(impl=1))
    ctx.set_particles(sc.x, sc.v)
    if graph:
        ctx.set_graph(True)
    ctx.step(steps)
    out = {k: t.cpu().numpy() for k, t in ctx.get_particles().items()}
    ctx.close()
    return out


def test_batch_matches_single_contexts(gpu):
    scenes = [s for s in mixed_scenes(12) if s.n]
    counts = [s.n for s in scenes]
    x, v = concat(scenes)
    b = make_batch(scenes)
    b.set_particles(counts, x, v)
    b.step(12)
    got = {k: t.cpu().numpy().reshape(sum(counts), -1) for k, t in b.get_particles().items()}
    t, steps = b.time()
    assert steps == 12 and abs(t - 12 * scenes[0].dt) < 1e-9
    b.close()
    p0 = 0
    for s, sc in enumerate(scenes):
        want = single_run(sc, 12)
        for key in ("x", "v", "C", "F"):
            err = rel(got[key][p0:p0 + sc.n].reshape(-1), want[key].reshape(-1))
            assert err < 1e-4, (s, key, err)
        p0 += sc.n


def test_batch_rejects_bad_calls(gpu):
    from physgrad_amd import PgError
    from physgrad_amd.solver import MPMContext
    scenes = [s for s in mixed_scenes(13) if s.n][:2]
    b = make_batch(scenes)
    x, v = concat(scenes)
    with pytest.raises(PgError):
        b.set_particles([scenes[0].n + scenes[1].n + 1, 0], np.zeros((x.shape[0] + 1, 3)),
                        np.zeros((x.shape[0] + 1, 3)))
    with pytest.raises(PgError):  # a batch takes per-scene counts
        MPMContext.set_particles(b, x, v)
    b.close()


def test_batch_backward_vs_oracle_reverse(gpu, oracle_lib):
    from test_gpu_adjoint import oracle_reverse
    O = oracle_lib
    scenes = [s for s in mixed_scenes(14) if s.n][:3]
    rng = np.random.default_rng(15)
    counts = [s.n for s in scenes]
    N = sum(counts)
    x, v = concat(scenes)
    C = rng.normal(0, 0.3, (N, 9)).astype(np.float32)
    F = (np.eye(3).reshape(1, 9) + rng.normal(0, 0.03, (N, 9))).astype(np.float32)
    E = np.exp(rng.uniform(np.log(5e3), np.log(5e4), N)).astype(np.float32)
    seeds = [rng.normal(size=(N, c)) for c in (3, 3, 9, 9)]
    b = make_batch(scenes)
    b.set_particles(counts, x, v, C, F, E=E)
    b.forward_checkpointed(5, 2)
    got = b.backward(*[torch.as_tensor(s, dtype=torch.float32) for s in seeds])
    got = {k: t.cpu().numpy().reshape(N, -1) for k, t in got.items()}
    b.close()
    p0 = 0
    for s, sc in enumerate(scenes):
        sl = slice(p0, p0 + sc.n)
        want, wE = oracle_reverse(O, sc, x[sl].astype(np.float64), v[sl].astype(np.float64),
                                  C[sl].astype(np.float64), F[sl].astype(np.float64), E[sl].astype(np.float64), 5,
                                  [sd[sl] for sd in seeds])
        for key, w in zip(("x", "v", "C", "F"), want):
            err = rel(got[key][sl].reshape(-1), w.reshape(-1))
            assert err < 1e-3, (s, key, err)
        assert rel(got["E"][sl].reshape(-1), wE) < 1e-3, s
        p0 += sc.n


def test_graph_replay_matches_eager(gpu):
    """Step counts 5, 4, 5, 5 (odd counts swap the state buffers, so the
    cached graph of the other parity is captured too); each replay is checked
    against the eager context after eager work and synchronizations between
    the calls."""
    from physgrad_amd.scenes import two_boxes
    from physgrad_amd.solver import MPMContext
    sc = two_boxes(n=65536)
    sc.v = sc.v + np.random.default_rng(21).normal(0, 0.5, sc.v.shape).astype(np.float32)
    ctxs = []
    for graph in (False, True):
        c = MPMContext(sc.params(impl=1))
        c.set_particles(sc.x, sc.v)
        c.set_graph(graph)
        ctxs.append(c)
    for k, steps in enumerate((5, 4, 5, 5)):
        outs = []
        for c in ctxs:
            c.step(steps)
            torch.cuda.synchronize()
            junk = torch.randn(1 << 16, device="cuda").sum()  # unrelated eager work between replays
            outs.append({key: t.cpu().numpy() for key, t in c.get_particles().items()})
            del junk
        for key in ("x", "v", "C", "F"):
            err = rel(outs[1][key].reshape(-1), outs[0][key].reshape(-1))
            assert err < 1e-4, (k, key, err)
        assert ctxs[0].time()[1] == ctxs[1].time()[1]
    for c in ctxs:
        c.close()


def test_graph_on_null_stream_and_batch(gpu):
    """Graph replay of a batch context, launched from torch's default stream
    (the library captures on an internal stream ordered against it)."""
    scenes = [s for s in mixed_scenes(16) if s.n]
    counts = [s.n for s in scenes]
    x, v = concat(scenes)
    outs = []
    for graph in (False, True):
        b = make_batch(scenes)
        b.set_particles(counts, x, v)
        b.set_graph(graph)
        b.step(3)
        b.step(3)
        outs.append({k: t.cpu().numpy() for k, t in b.get_particles().items()})
        b.close()
    for key in ("x", "v", "C", "F"):
        assert rel(outs[1][key].reshape(-1), outs[0][key].reshape(-1)) < 1e-4, key


def test_autotune_restores_state_and_caches(gpu, tmp_path, monkeypatch):
    from physgrad_amd.scenes import two_boxes
    from physgrad_amd.solver import MPMContext
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("PG_AUTOTUNE_CACHE", str(cache))
    sc = two_boxes(n=131072)
    sc.v = sc.v + np.random.default_rng(22).normal(0, 0.5, sc.v.shape).astype(np.float32)
    c = MPMContext(sc.params(impl=1))
    c.set_particles(sc.x, sc.v)
    c.step(2)
    before = {k: t.cpu().numpy() for k, t in c.get_particles().items()}
    choice = c.autotune(trial_steps=3, use_cache=False)
    assert choice in (256, 512, 1024, 2048, 4096, 8192)
    after = {k: t.cpu().numpy() for k, t in c.get_particles().items()}
    for key in before:
        assert np.array_equal(before[key], after[key]), key
    assert c.time()[1] == 2
    tuned = c.tuning()
    assert tuned["tile_workgroups"] == choice and tuned["work_item"] in (640, 1280, 2560)
    assert tuned["block_threads"] in (256, 512)
    last = cache.read_text().strip().splitlines()[-1].split()
    assert last[1:] == [str(choice), str(tuned["work_item"]), str(tuned["block_threads"])]
    # a second context of the same shape takes the cached values without timing
    d = MPMContext(sc.params(impl=1))
    d.set_particles(sc.x, sc.v)
    d.launch = 256 if choice != 256 else 512
    d.work_item = 5120
    d.block_threads = 512 if tuned["block_threads"] == 256 else 256
    assert d.autotune(trial_steps=3, use_cache=True) == choice
    assert d.tuning() == tuned
    # the tuned launch gives the same trajectory
    d.step(2)
    c2 = {k: t.cpu().numpy() for k, t in d.get_particles().items()}
    for key in before:
        assert rel(c2[key].reshape(-1), before[key].reshape(-1)) < 1e-5, key
    c.close()
    d.close()
