"""GPU: the tiled path's launch-shape knobs of the autotuner: the work-item
size (pg_mpm_set_work_item: a dense tile split into work items of 640 .. 5120
particles, one to several LDS passes each) and the tile workgroup size
(pg_mpm_set_block: 256 or 512 threads), against the FP64 oracle, forward and
reverse."""
import numpy as np
import pytest
import torch

from test_gpu_adjoint import oracle_reverse, rel, scene

pytestmark = pytest.mark.gpu


def dense(rng, n=6000):
    return scene(rng, n, 1, span=0.2)   # 6,000 particles in about one 4^3-cell tile


@pytest.mark.parametrize("work_item,block", [(640, 256), (2560, 256), (5120, 256), (1280, 512), (2560, 512)])
def test_work_item_forward_vs_fp64(gpu, oracle_lib, work_item, block):
    from physgrad_amd.solver import MPMContext
    O = oracle_lib
    rng = np.random.default_rng(60)
    sc = dense(rng)
    n = sc.n
    ctx = MPMContext(sc.params(impl=1))
    ctx.work_item = work_item
    ctx.block_threads = block
    assert ctx.tuning()["work_item"] == work_item and ctx.tuning()["block_threads"] == block
    ctx.set_particles(sc.x, sc.v, E=sc.E_particles)
    prm = O.scene_params(sc)
    x64, v64 = sc.x.astype(np.float64), sc.v.astype(np.float64)
    C64, F64 = np.zeros((n, 9)), np.tile(np.eye(3).reshape(1, 9), (n, 1))
    for k in range(3):
        print("work item", work_item, "step", k, flush=True)
        ctx.step(1)
        want = O.mls_run(prm, x64, v64, C64, F64, sc.p_mass, sc.E_particles.astype(np.float64), 1)
        got = {kk: t.cpu().numpy() for kk, t in ctx.get_particles().items()}
        for key in ("x", "v", "C", "F"):
            err = rel(got[key].reshape(-1), want[key].reshape(-1))
            assert err <= (1e-5 if k == 0 else 1e-4), (key, k, err)
        x64, v64, C64, F64 = want["x"], want["v"], want["C"].reshape(-1, 9), want["F"].reshape(-1, 9)
    ctx.close()


@pytest.mark.parametrize("work_item,block", [(640, 256), (2560, 256), (1280, 512)])
def test_work_item_backward_vs_oracle_reverse(gpu, oracle_lib, work_item, block):
    from physgrad_amd.solver import MPMContext
    O = oracle_lib
    rng = np.random.default_rng(61)
    sc = dense(rng)
    n, nsteps, K = sc.n, 4, 2
    C = rng.normal(0, 0.3, (n, 9)).astype(np.float32)
    F = (np.eye(3).reshape(1, 9) + rng.normal(0, 0.03, (n, 9))).astype(np.float32)
    seeds = [rng.normal(size=(n, c)) for c in (3, 3, 9, 9)]
    print("work item backward", work_item, flush=True)
    ctx = MPMContext(sc.params(impl=1))
    ctx.work_item = work_item
    ctx.block_threads = block
    ctx.set_particles(sc.x, sc.v, C, F, E=sc.E_particles)
    ctx.forward_checkpointed(nsteps, K)
    got = ctx.backward(*[torch.as_tensor(s, dtype=torch.float32) for s in seeds])
    want, wE = oracle_reverse(O, sc, sc.x.astype(np.float64), sc.v.astype(np.float64), C.astype(np.float64),
                              F.astype(np.float64), sc.E_particles.astype(np.float64), nsteps, seeds)
    for key, w in zip(("x", "v", "C", "F"), want):
        assert rel(got[key].cpu().numpy().reshape(-1), w.reshape(-1)) < 1e-3, key
    assert rel(got["E"].cpu().numpy(), wE) < 1e-3
    ctx.close()


def test_work_item_change_between_calls(gpu, oracle_lib):
    """Changing the size while a step's binning is pending: the next call bins
    with the new size; the trajectory equals the default size's."""
    from physgrad_amd.solver import MPMContext
    rng = np.random.default_rng(62)
    sc = dense(rng)
    out = []
    for sizes in ((1280, 1280, 256), (1280, 640, 512)):
        ctx = MPMContext(sc.params(impl=1))
        ctx.set_graph(True)
        ctx.set_particles(sc.x, sc.v, E=sc.E_particles)
        ctx.work_item = sizes[0]
        ctx.step(2)
        ctx.work_item = sizes[1]
        ctx.block_threads = sizes[2]
        ctx.step(2)
        out.append({k: t.cpu().numpy() for k, t in ctx.get_particles().items()})
        ctx.close()
    for key in ("x", "v", "C", "F"):
        assert rel(out[1][key].reshape(-1), out[0][key].reshape(-1)) < 1e-5, key


def test_work_item_bounds(gpu):
    from physgrad_amd._lib import PgError
    from physgrad_amd.solver import MPMContext
    rng = np.random.default_rng(63)
    ctx = MPMContext(dense(rng, 100).params(impl=1))
    t = ctx.tuning()
    assert t == {"tile_workgroups": 4096, "work_item": 1280, "tile_edge": 4, "block_threads": 256}
    for bad in (639, 5121, 0):
        with pytest.raises(PgError, match="work item"):
            ctx.work_item = bad
    for bad in (128, 384, 1024):
        with pytest.raises(PgError, match="256 or 512"):
            ctx.block_threads = bad
    ctx.close()
    plain = MPMContext(dense(rng, 100).params(impl=0))
    with pytest.raises(PgError):
        plain.work_item = 1280
    plain.close()
