"""State I/O: VTK snapshots, checkpoint files and resume.

CPU: the VTK writer's layout read back (points, scalar and vector fields),
the checkpoint arrays round trip exactly through a file, and a checkpoint
taken mid-run of a deterministic CPU integrator (the FP64 oracle step)
resumes to a bit-identical end state.  GPU: save -> load -> continue against
an uninterrupted run for a tiled context, a plain context and a batch (float
rounding: a restored tiled context bins its particles afresh), and run()'s
output cadence.
"""
import os

import numpy as np
import pytest


def test_vtk_round_trip(tmp_path):
    from physgrad_amd.state_io import read_vtk_points, write_vtk
    rng = np.random.default_rng(0)
    x = rng.random((37, 3)).astype(np.float32)
    v = rng.normal(size=(37, 3)).astype(np.float32)
    s = rng.random(37).astype(np.float32)
    p = str(tmp_path / "a.vtk")
    write_vtk(p, x, {"velocity": v, "speed": s})
    head = open(p, "rb").read(200)
    assert head.startswith(b"# vtk DataFile Version 3.0\n") and b"DATASET POLYDATA" in head
    x2, f = read_vtk_points(p)
    assert np.array_equal(x2, x)
    assert np.array_equal(f["velocity"], v) and np.array_equal(f["speed"], s)
    with pytest.raises(ValueError):
        write_vtk(p, x, {"bad": np.zeros((37, 2))})


def test_checkpoint_arrays_round_trip(tmp_path):
    from physgrad_amd.state_io import pack_checkpoint, unpack_checkpoint
    rng = np.random.default_rng(1)
    parts = {"x": rng.random((10, 3)), "v": rng.random((10, 3)), "C": rng.random((10, 3, 3)),
             "F": rng.random((10, 3, 3)), "E": rng.random(10), "mass": None}
    cfg = {"nx": np.array(16), "dx": np.array(0.0625, np.float32)}
    arrays = pack_checkpoint("context", cfg, parts, 0.125, 1250)
    p = str(tmp_path / "c.npz")
    np.savez(p, **arrays)
    with np.load(p, allow_pickle=False) as f:
        ck = unpack_checkpoint(f)
    assert ck["kind"] == "context" and ck["time"] == 0.125 and ck["steps"] == 1250
    for k in ("x", "v", "C", "F", "E"):
        assert np.array_equal(ck["particles"][k], parts[k].astype(np.float32))
    assert "mass" not in ck["particles"]
    assert ck["config"]["nx"] == 16 and ck["config"]["dx"] == np.float32(0.0625)


def test_resume_is_bitwise_on_a_deterministic_cpu_integrator(oracle_lib, tmp_path):
    """20 steps of the sequential FP32 CPU step against 10 steps, checkpoint
    to disk, reload, 10 more steps: identical bits (the sequential CPU
    integrator is deterministic, so a difference could only come from the
    file round trip)."""
    from physgrad_amd.state_io import pack_checkpoint, unpack_checkpoint
    O = oracle_lib
    rng = np.random.default_rng(2)
    n = 300
    x = (0.35 + 0.3 * rng.random((n, 3))).astype(np.float32)
    v = rng.normal(0, 0.3, (n, 3)).astype(np.float32)
    prm = O.mls_params((16, 16, 16), 1 / 16, (0, 0, 0), 1e-4, (0, -9.8, 0), 3, 1, 0.3, 1e-5)
    E = (1e4 * (0.5 + rng.random(n))).astype(np.float32)
    full = O.mls_run(prm, x, v, None, None, 1e-3, E, 20, dtype=np.float32)
    half = O.mls_run(prm, x, v, None, None, 1e-3, E, 10, dtype=np.float32)
    p = str(tmp_path / "ck.npz")
    np.savez(p, **pack_checkpoint("context", {}, {k: half[k] for k in ("x", "v", "C", "F")} | {"E": E}, 10e-4, 10))
    with np.load(p, allow_pickle=False) as f:
        ck = unpack_checkpoint(f)
    pp = ck["particles"]
    rest = O.mls_run(prm, pp["x"], pp["v"], pp["C"], pp["F"], 1e-3, pp["E"], 10, dtype=np.float32)
    for k in ("x", "v", "C", "F"):
        assert np.array_equal(rest[k], full[k]), k
    assert not np.array_equal(half["x"], full["x"])


def _scene(rng, n=3000):
    from physgrad_amd.scenes import Scene
    x = (0.3 + 0.4 * rng.random((n, 3))).astype(np.float32)
    v = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    E = np.exp(rng.uniform(np.log(5e3), np.log(5e4), n)).astype(np.float32)
    return Scene("ck", (32, 32, 32), 1 / 32, 1e-4, 20, 1, 2e4, 0.3, 1000.0, 0.4 ** 3 / n, x, v, E_particles=E)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [0, 1], ids=["plain", "tiled"])
def test_gpu_checkpoint_resume(gpu, tmp_path, impl):
    from physgrad_amd.solver import MPMContext
    from physgrad_amd.state_io import load_checkpoint, save_checkpoint
    rng = np.random.default_rng(3)
    sc = _scene(rng)
    a = MPMContext(sc.params(impl=impl))
    a.set_particles(sc.x, sc.v, E=sc.E_particles)
    a.step(20)
    want = {k: t.cpu().numpy() for k, t in a.get_particles().items()}
    b = MPMContext(sc.params(impl=impl))
    b.set_particles(sc.x, sc.v, E=sc.E_particles)
    b.step(10)
    p = str(tmp_path / "ck.npz")
    save_checkpoint(b, p)
    clock = b.time()
    b.close()
    c = load_checkpoint(p)
    assert c.time() == clock and clock[1] == 10
    c.step(10)
    got = {k: t.cpu().numpy() for k, t in c.get_particles().items()}
    assert c.time()[1] == 20
    for k in ("x", "v", "C", "F"):
        err = np.abs(got[k] - want[k]).max() / np.abs(want[k]).max()
        assert err <= 1e-5, (k, err)
    a.close()
    c.close()


@pytest.mark.gpu
def test_gpu_batch_checkpoint_resume_and_run_cadence(gpu, tmp_path):
    from physgrad_amd.scenes import batch_scenes
    from physgrad_amd.solver import MPMBatch
    from physgrad_amd.state_io import load_checkpoint, read_vtk_points, run, save_checkpoint
    scenes = batch_scenes(nscenes=5, max_n=8000)
    counts = [s.n for s in scenes]
    x = np.concatenate([s.x for s in scenes])
    v = np.concatenate([s.v for s in scenes])
    a = MPMBatch([s.desc() for s in scenes])
    a.set_particles(counts, x, v)
    a.step(12)
    want = a.get_particles()["x"].cpu().numpy()
    b = MPMBatch([s.desc() for s in scenes])
    b.set_particles(counts, x, v)
    b.step(5)
    p = str(tmp_path / "b.npz")
    save_checkpoint(b, p)
    c = load_checkpoint(p)
    assert c.counts == counts and c.num_scenes == 5
    # run() to t = 12 dt with a snapshot every 2 dt: frames at steps 6, 8, 10, 12
    d = str(tmp_path / "frames")
    frames = run(c, 12 * scenes[0].dt, output_frequency=2 * scenes[0].dt, output_dir=d)
    assert frames == 4 and sorted(os.listdir(d)) == [f"frame_{i:06d}.vtk" for i in range(4)]
    got = c.get_particles()["x"].cpu().numpy()
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    xs, fields = read_vtk_points(os.path.join(d, "frame_000003.vtk"))
    assert np.array_equal(xs, got) and set(fields) == {"velocity", "speed", "J"}
    for ctx in (a, b, c):
        ctx.close()
