"""Benchmark of the MLS-MPM transfer loop on one MI355X.

Workload (two_boxes, 128^3 grid, dx 1/128, 1,048,576 particles, neo-Hookean
E 5e4 nu 0.3 rho 1000, dt 5e-5): a step is one P2G -> grid update -> G2P pass
over all particles with the particle state resident in HBM.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--shape W,I,T]

Prints one JSON line: particle-steps/s of the tiled path (value, launch shape
autotuned, or pinned with --shape workgroups,work-item,threads), the roofline of the dominant kernel (the fused G2P+P2G tile
kernel, timed with HIP events on its launch stream), forward+backward with
checkpoint recompute (one 1M scene and the 16-scene batch, K = 16), rollout
gradients vs closed form and FP64 finite differences, the 64-scene batch, the
SPH-MPM coupling pair search and forces (1M x 1M), the plain GPU port measured
in the same run, the FP64 error of the forward, and the CPU baselines (FP32
port of the same step on the host cores, the reference's CPU solver, the
all-pairs coupling loop; bounded samples).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_PER_PARTICLE = 192   # x,v,C,F (96 B) read + written once per step
BYTES_PER_NODE = 32        # grid (m,p) written once + (v) read once per active node
COPY_BYTES = 1 << 30       # size of the calibration copy and sum (4x the 256 MB Infinity Cache)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--particles", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain GPU port leg")
    ap.add_argument("--cpu-steps", type=int, default=24)
    ap.add_argument("--no-bwd", action="store_true", help="skip the forward+backward and gradient legs")
    ap.add_argument("--no-grad", action="store_true", help="skip the gradient-accuracy leg")
    # one scene of the batched forward+backward workload: 256 steps with a
    # checkpoint every 16, so the backward recomputes 15 of its 16 segments
    ap.add_argument("--bwd-steps", type=int, default=256)
    ap.add_argument("--bwd-ckpt", type=int, default=16)
    ap.add_argument("--plain-bwd-steps", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batch leg")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--no-batch-fb", action="store_true", help="skip the 16-scene forward+backward leg")
    ap.add_argument("--no-coupling", action="store_true", help="skip the SPH-MPM coupling leg")
    ap.add_argument("--coupling-iters", type=int, default=20)
    ap.add_argument("--shape", type=str, default=None,
                    help="pin the forward leg's launch shape WORKGROUPS,WORK_ITEM,THREADS instead of autotuning "
                         "it (reproducible profiler runs)")
    return ap.parse_args()


def time_steps(ctx, steps, warmup):
    import torch
    ctx.step(warmup)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    a.record(s)
    ctx.step(steps)
    b.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return a.elapsed_time(b) / 1e3, wall


def com_seed(xT, scenes, counts):
    """Adjoint seed dL/dx_T of L = sum over scenes of |mean x_T - target|^2
    (FP64 reduction) and the residuals mean x_T - target."""
    import torch
    seed = torch.empty_like(xT)
    resid = []
    off = 0
    for s, c in zip(scenes, counts):
        d = xT[off:off + c].double().mean(0) - torch.tensor(s.target, dtype=torch.float64, device=xT.device)
        seed[off:off + c] = (2.0 * d / c).float().expand(c, 3)
        resid.append(d.cpu().numpy())
        off += c
    return seed, resid


def closed_form_error(grads, scenes, counts, resid, T):
    """Largest relative error of dL/dx0 and dL/dv0 against the closed form
    2 (mean x_T - target) / N (times T dt for v0), valid while no particle
    reaches a sticky wall: the transfers conserve linear momentum exactly."""
    import numpy as np
    worst = 0.0
    off = 0
    for s, c, d in zip(scenes, counts, resid):
        wx = 2.0 * d / c
        for key, want in (("x", wx), ("v", wx * T * s.dt)):
            g = grads[key][off:off + c].double()
            w = torch_tensor(want, g)
            worst = max(worst, float(((g - w).abs().max() / w.abs().max()).item()))
        off += c
    return worst


def torch_tensor(a, like):
    import torch
    return torch.as_tensor(a, dtype=torch.float64, device=like.device).expand_as(like)


def fwd_bwd_leg(scene, steps, ckpt, impl):
    """Forward with a checkpoint every `ckpt` steps plus the full backward
    (recompute of every segment but the last, then its reverse) on one scene
    of the batched forward+backward workload; the adjoint seed is that of its
    centre-of-mass loss, and the gradients are checked against the closed
    form."""
    import torch
    from physgrad_amd.solver import MPMContext
    ctx = MPMContext(scene.params(impl=impl))
    n = scene.n
    ctx.set_particles(scene.x, scene.v, E=scene.E_particles)
    ctx.forward_checkpointed(2 * ckpt, ckpt)   # warm-up: allocates checkpoints and the tape
    ctx.backward(torch.zeros((n, 3), device="cuda"))
    ctx.set_particles(scene.x, scene.v, E=scene.E_particles)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    a.record(s)
    ctx.forward_checkpointed(steps, ckpt)
    b.record(s)
    seed, resid = com_seed(ctx.get_particles()["x"], [scene], [n])
    b2 = torch.cuda.Event(enable_timing=True)
    b2.record(s)
    grads = ctx.backward(seed)
    c.record(s)
    torch.cuda.synchronize()
    t_f, t_b = a.elapsed_time(b) / 1e3, b2.elapsed_time(c) / 1e3
    err = closed_form_error(grads, [scene], [n], resid, steps)
    active = int((ctx.get_grid()["mass"] > 0).sum().item())
    # per-kernel times (HIP events on the launch stream) of a shorter forward+backward
    # over the same scene: two segments, the later one recomputed
    ctx.set_particles(scene.x, scene.v, E=scene.E_particles)
    ctx.set_profiling(True)
    ctx.forward_checkpointed(2 * ckpt, ckpt)
    ctx.backward(seed)
    torch.cuda.synchronize()
    kt = {k: v[0] / v[1] for k, v in ctx.kernel_times().items()}
    ctx.set_profiling(False)
    ctx.close()
    # algorithmic bytes per step: forward, recompute, tape write, adjoint state read/read/write
    # (96 B each per particle) + per-particle E, and the grid forward (32 B) + adjoint (64 B)
    fb_bytes = n * (768 + 20) + 96 * active
    t = t_f + t_b
    return {"value": n * steps / t, "unit": "particle-steps/s", "steps": steps, "checkpoint_every": ckpt,
            "ms_per_step": 1e3 * t / steps, "forward_ms_per_step": 1e3 * t_f / steps,
            "backward_ms_per_step": 1e3 * t_b / steps, "grad_rel_err_closed_form": err,
            "roofline": {"bound": "hbm", "achieved": fb_bytes * steps / t / 1e9, "peak": HBM_PEAK / 1e9,
                         "unit": "GB/s", "frac": fb_bytes * steps / t / HBM_PEAK,
                         "alg_bytes_per_step": fb_bytes, "active_nodes": active},
            "kernel_ms_per_launch": kt,
            "workload": f"{scene.name}: one scene of batch_fb ({n:,} particles, 128^3, neo-Hookean, per-particle E, "
                        f"dt 5e-5), {steps} steps, checkpoint every {ckpt} (recompute), centre-of-mass loss"}


def batch_fb_leg(ckpt):
    """The batched forward+backward workload: 16 scenes of 1,048,576 particles
    on 128^3 in one batch context, 256 steps with a checkpoint every 16,
    centre-of-mass loss per scene; gradients checked against the closed form."""
    import numpy as np
    import torch
    from physgrad_amd.scenes import batch_fb
    from physgrad_amd.solver import MPMBatch
    scenes = batch_fb()
    counts = [s.n for s in scenes]
    n = sum(counts)
    T = scenes[0].steps
    b = MPMBatch([s.desc() for s in scenes])
    x = np.concatenate([s.x for s in scenes])
    v = np.concatenate([s.v for s in scenes])
    E = np.concatenate([s.E_particles for s in scenes])
    b.set_particles(counts, x, v, E=E)
    b.forward_checkpointed(2 * ckpt, ckpt)   # warm-up: allocates checkpoints and the tape
    b.backward(torch.zeros((n, 3), device="cuda"))
    b.set_particles(counts, x, v, E=E)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record(s)
    b.forward_checkpointed(T, ckpt)
    ev[1].record(s)
    seed, resid = com_seed(b.get_particles()["x"], scenes, counts)
    ev[2].record(s)
    grads = b.backward(seed)
    ev[3].record(s)
    torch.cuda.synchronize()
    t_f, t_b = ev[0].elapsed_time(ev[1]) / 1e3, ev[2].elapsed_time(ev[3]) / 1e3
    err = closed_form_error(grads, scenes, counts, resid, T)
    free, total = torch.cuda.mem_get_info()
    b.close()
    return {"value": n * T / (t_f + t_b), "unit": "particle-steps/s", "scenes": len(scenes), "particles": n,
            "steps": T, "checkpoint_every": ckpt, "ms_per_step": 1e3 * (t_f + t_b) / T,
            "forward_ms_per_step": 1e3 * t_f / T, "backward_ms_per_step": 1e3 * t_b / T,
            "grad_rel_err_closed_form": err, "device_memory_used_gb": (total - free) / 1e9,
            "workload": "batch_fb: 16 scenes x (128^3, 1,048,576 particles, seed 5), v ~ N(0, 0.5^2), neo-Hookean "
                        "with per-particle E, dt 5e-5, 256 steps, checkpoint every 16, centre-of-mass loss per scene"}


def batch_leg(steps):
    """64 independent scenes (grid side 32/48/64, 4,096..65,536 particles each,
    mixed materials, seed 4) advanced by the same launches: autotuned tile
    workgroups, then steps [10 + steps, 10 + 2 steps) timed eager and replayed
    from a captured hipGraph.  A graph holds one call's whole step sequence, so
    the graph leg's untimed steps [10, 10 + steps) capture the graph that the
    timed call replays (capture and instantiation are not in the timed region);
    the eager leg runs the same untimed steps."""
    import numpy as np
    import torch
    from physgrad_amd.scenes import batch_scenes
    from physgrad_amd.solver import MPMBatch
    scenes = batch_scenes()
    counts = [s.n for s in scenes]
    x = np.concatenate([s.x for s in scenes])
    v = np.concatenate([s.v for s in scenes])
    n = int(sum(counts))
    b = MPMBatch([s.desc() for s in scenes])
    b.set_particles(counts, x, v)
    tuned = b.autotune(trial_steps=10, use_cache=False)
    b.set_particles(counts, x, v)
    b.step(10)
    t_eager, _ = time_steps(b, steps, steps)
    # per-kernel times of the tuned launch (HIP events on the launch stream)
    b.set_particles(counts, x, v)
    b.set_profiling(True)
    b.step(min(steps, 50))
    torch.cuda.synchronize()
    kt = b.kernel_times()
    b.set_profiling(False)
    b.set_particles(counts, x, v)
    b.set_graph(True)
    b.step(10)
    t_graph, _ = time_steps(b, steps, steps)
    ok = bool(torch.isfinite(b.get_particles()["x"]).all().item())
    shape = b.tuning()
    b.close()
    return {"value": n * steps / t_graph, "unit": "particle-steps/s", "scenes": len(scenes), "particles": n,
            "steps": steps, "ms_per_step": 1e3 * t_graph / steps, "ms_per_step_eager": 1e3 * t_eager / steps,
            "tile_workgroups": tuned, "tuning": shape, "state_finite": ok,
            "kernel_ms_per_launch": {k: v[0] / v[1] for k, v in kt.items()},
            "workload": "batch_64: 64 scenes, side 32/48/64, 4,096..65,536 particles each (seed 4), "
                        "fixed-corotated / neo-Hookean, dt 1e-4, hipGraph replay of steps [10 + steps, 10 + 2 steps)"}


def grad_leg():
    """Gradient accuracy on the differentiable rollout (64^3, 262,144 particles,
    fixed-corotated with per-particle E, 512 steps, checkpoint every 32, shared
    initial velocity v0, loss = |mean x - (0.6, 0.3, 0.5)|^2).  Away from the
    walls the transfers conserve momentum exactly, so mean x after T steps is
    mean x0 + T dt v0 + dt^2 g T (T + 1) / 2 and dL/dv0 = 2 (mean x_T - target) T dt,
    while dL/dE vanishes: both are checked against the GPU adjoint."""
    import numpy as np
    import torch
    from physgrad_amd import _lib
    from physgrad_amd.autograd import physics_simulation
    from physgrad_amd.scenes import rollout
    from physgrad_amd.solver import MPMContext
    sc = rollout()
    n, T, K = sc.n, sc.steps, 32
    ctx = MPMContext(sc.params(impl=_lib.IMPL_TILED))
    x0 = torch.as_tensor(sc.x, device="cuda")
    v0 = torch.zeros(3, dtype=torch.float32, device="cuda", requires_grad=True)
    E = torch.as_tensor(sc.E_particles, device="cuda").clone().requires_grad_(True)
    target = torch.tensor(sc.target, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xT, _, _, _ = physics_simulation(ctx, x0, v0.expand(n, 3), E=E, num_steps=T, checkpoint_every=K)
    d = xT.double().mean(0) - target
    loss = (d * d).sum()
    loss.backward()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    g = np.array(sc.gravity, np.float64)
    xbar_T = sc.x.astype(np.float64).mean(0) + sc.dt ** 2 * g * T * (T + 1) / 2
    want = 2.0 * (xbar_T - np.array(sc.target)) * T * sc.dt
    got = v0.grad.double().cpu().numpy()
    err = float(np.abs(got - want).max() / np.abs(want).max())
    # d loss / d log E against the v0 sensitivity: zero up to round-off
    gE = float((E.grad.double() * E.double()).abs().max().item())
    out = {"rel_err_dv0": err, "dv0": got.tolist(), "dv0_closed_form": want.tolist(),
           "max_abs_dlogE": gE, "fwd_bwd_wall_s": wall,
           "workload": "rollout_64: 64^3, 262,144 particles, fixed-corotated, per-particle E log-uniform [5e3,5e4], "
                       "512 steps at dt 1e-4, checkpoint every 32"}
    # The workload's body falls rigidly (F stays I), so no loss depends on E.
    # From a deforming initial state v0 + A (x - c), an E-sensitive loss
    # L = sum_p w_p . x_p(T) is checked against FP64 central finite differences
    # of the oracle (eps 1e-4; relative for E) on v0 and 16 sampled E.
    fd_path = os.path.join(ROOT, "tests", "golden", "rollout_fd.npz")
    if os.path.exists(fd_path):
        f = np.load(fd_path)
        w = torch.as_tensor(np.random.default_rng(int(f["weights_seed"])).normal(size=(n, 3)) / n, device="cuda")
        x64 = sc.x.astype(np.float64)
        vdef = torch.as_tensor(((x64 - x64.mean(0)) @ f["A_rate"].T).astype(np.float32), device="cuda")
        v0 = torch.zeros(3, dtype=torch.float32, device="cuda", requires_grad=True)
        E = torch.as_tensor(sc.E_particles, device="cuda").clone().requires_grad_(True)
        xT, _, _, _ = physics_simulation(ctx, x0, v0.expand(n, 3) + vdef, E=E, num_steps=T, checkpoint_every=K)
        (w * xT.double()).sum().backward()
        gv = v0.grad.double().cpu().numpy()
        gE = E.grad.double().cpu().numpy()[f["E_index"]]
        out["weighted_loss_vs_fp64_fd"] = {
            "rel_err_dv0": float(np.abs(gv - f["dv0"]).max() / np.abs(f["dv0"]).max()),
            "rel_err_dE": float(np.abs(gE - f["dE"]).max() / np.abs(f["dE"]).max()),
            "params": 3 + len(f["dE"]), "eps": float(f["eps"])}
    ctx.close()
    return out


def fp64_error_leg(scene, steps):
    """Error of the tiled GPU forward against the FP64 oracle after `steps`
    steps of the full single-scene workload (inf-norm relative per field)."""
    import numpy as np
    from oracle import oracle as O
    from physgrad_amd import _lib
    from physgrad_amd.solver import MPMContext
    ctx = MPMContext(scene.params(impl=_lib.IMPL_TILED))
    ctx.set_particles(scene.x, scene.v)
    ctx.step(steps)
    got = {k: t.cpu().numpy() for k, t in ctx.get_particles().items()}
    ctx.close()
    t0 = time.perf_counter()
    want = O.mls_run(O.scene_params(scene), scene.x.astype(np.float64), scene.v.astype(np.float64), None, None,
                     scene.p_mass, scene.E, steps)
    out = {"steps": steps, "oracle_s": time.perf_counter() - t0}
    for k in ("x", "v", "F"):
        a, b = got[k].astype(np.float64).reshape(-1), want[k].reshape(-1)
        out[f"rel_err_{k}"] = float(np.abs(a - b).max() / np.abs(b).max())
    # C is relative to the velocity-gradient scale max |v| / dx (rigidly moving
    # bodies have C = 0 up to round-off)
    dC = np.abs(got["C"].astype(np.float64) - want["C"]).max()
    out["rel_err_C"] = float(dC / (np.abs(want["v"]).max() / scene.dx))
    return out


def coupling_leg(iters, with_cpu):
    """SPH-MPM coupling (pg_coupling_*): 1,048,576 SPH and 1,048,576 MPM
    particles, uniform in the unit box, coupling radius 0.012 (about 7.6 MPM
    partners per SPH particle).  One iteration = findCouplingPairs (MPM
    binning, count, scan, fill with sorted rows; ends with the host read of
    the pair count) + computeCouplingForces (forces and the two-way reaction).
    The CPU baseline is the oracle's restatement of the reference's all-pairs
    loop on a sample of SPH rows against all MPM particles."""
    import numpy as np
    import torch
    from physgrad_amd.spatial import SPHMPMCoupling
    rng = np.random.default_rng(11)
    n, radius = 1 << 20, 0.012
    xs = torch.as_tensor(rng.random((n, 3)).astype(np.float32)).cuda()
    xm = torch.as_tensor(rng.random((n, 3)).astype(np.float32)).cuda()
    vs = torch.as_tensor(rng.normal(size=(n, 3)).astype(np.float32)).cuda()
    vm = torch.as_tensor(rng.normal(size=(n, 3)).astype(np.float32)).cuda()
    f = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    c = SPHMPMCoupling(coupling_radius=radius)
    for _ in range(3):
        c.findCouplingPairs(xs, xm)
        c.computeCouplingForces(xs, vs, xm, vm, f)
    torch.cuda.synchronize()
    t_find = t_force = 0.0
    for _ in range(iters):
        t0 = time.perf_counter()
        npairs = c.findCouplingPairs(xs, xm)          # synchronises (pair count to the host)
        t1 = time.perf_counter()
        c.computeCouplingForces(xs, vs, xm, vm, f)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_find += t1 - t0
        t_force += t2 - t1
    c.close()
    t_it = (t_find + t_force) / iters
    out = {"workload": "1,048,576 SPH x 1,048,576 MPM particles, uniform in [0,1]^3, radius 0.012",
           "pairs": npairs, "iters": iters, "ms_find_pairs": 1e3 * t_find / iters,
           "ms_forces": 1e3 * t_force / iters, "sph_rows_per_s": n / t_it, "pairs_per_s": npairs / t_it,
           "unit": "SPH particles searched + forced per second"}
    # minimum HBM bytes of one find+forces: positions in (SPH 12 B, MPM 12 B
    # read by the binning + 16 B gathered), row counts and offsets (12 B per
    # SPH row), pairs out (col + penetration 8 B), forces (pairs: col, pen
    # in 8 B, positions/velocities gathered 24 B, 24 B out; rows: 24 B)
    alg = n * (12 + 12 + 16 + 12 + 24) + npairs * (8 + 8 + 24 + 24)
    out["effective_GBps"] = alg / t_it / 1e9
    if with_cpu:
        from oracle import oracle as O
        sample = 1024
        xs_h, xm_h = xs[:sample].cpu().numpy(), xm.cpu().numpy()
        t0 = time.perf_counter()
        O.coupling_pairs(xs_h, xm_h, radius)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": sample / dt, "unit": "SPH particles searched per second", "cores": 1,
                               "kind": "port",
                               "sample": f"{sample} SPH rows x {n} MPM particles with the oracle's restatement "
                                         f"of findCouplingPairs (all-pairs loop, single thread), {dt:.2f} s"}
        out["speedup_vs_cpu_rows"] = out["sph_rows_per_s"] / out["cpu_baseline"]["value"]
    return out


def cpu_baseline(scene, steps):
    """FP32 port of the same step (oracle) on all host cores (OpenMP) for `steps`
    steps of the full workload, the same port on one core for steps // 2 steps,
    and the reference's own CPU solver for one step."""
    import numpy as np
    from oracle import oracle as O
    prm = O.scene_params(scene)
    threads = O.num_threads()
    t0 = time.perf_counter()
    O.mls_run_mt(prm, scene.x, scene.v, None, None, scene.p_mass, scene.E, steps)
    dt = time.perf_counter() - t0
    out = {"value": scene.n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
           "sample": f"{steps} steps of the full workload ({scene.n} particles, 128^3) with the FP32 port of the "
                     f"same MLS-MPM step on {threads} OpenMP threads, {dt:.1f} s"}
    s1 = max(1, steps // 2)
    t0 = time.perf_counter()
    O.mls_run(prm, scene.x.copy(), scene.v, None, None, scene.p_mass, scene.E, s1, dtype=np.float32)
    dt1 = time.perf_counter() - t0
    out["single_core"] = {"value": scene.n * s1 / dt1, "cores": 1,
                          "sample": f"{s1} steps, the sequential FP32 port, {dt1:.1f} s"}
    ref = None
    if O.ref_available():
        # the reference's own CPU solver (PIC semantics, no stress) on the same particles and grid
        prm_r = O.ref_params(scene.dims, (scene.dx,) * 3, scene.origin, scene.dt, scene.gravity, 2)
        m = np.full(scene.n, scene.p_mass, np.float32)
        res = O.reference_run(prm_r, scene.x, scene.v, m, 1, want_grid=False)
        ref = {"value": scene.n / res[-1], "unit": "particle-steps/s", "cores": 1, "kind": "reference",
               "sample": "1 step of MPMSolver<float,64>::step (reference PIC semantics) on the same "
                         f"{scene.n} particles / 128^3 grid, single thread, {res[-1]:.2f} s"}
    return out, ref


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (every box here has one GPU).",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from physgrad_amd import _lib
    from physgrad_amd.scenes import two_boxes
    from physgrad_amd.solver import MPMContext

    _lib.load()
    scene = two_boxes(n=args.particles)
    N, K, W = scene.n, args.steps, args.warmup

    ctx = MPMContext(scene.params(impl=_lib.IMPL_TILED))
    ctx.set_particles(scene.x, scene.v)
    # launch shape (tile workgroups x particles per work item x threads) picked
    # by the autotuner on this particle set (the state is restored after the
    # trials), or pinned by --shape
    if args.shape:
        wg, item, threads = (int(v) for v in args.shape.split(","))
        ctx.launch = wg
        ctx.work_item = item
        ctx.block_threads = threads
    else:
        ctx.autotune(trial_steps=10, use_cache=False)
    tuning = ctx.tuning()
    t_tiled, wall = time_steps(ctx, K, W)
    grid = ctx.get_grid()
    active = int((grid["mass"] > 0).sum().item())
    finite = bool(torch.isfinite(ctx.get_particles()["x"]).all().item())
    # dominant kernel: fused G2P+P2G, timed with events on its launch stream
    ctx.set_profiling(True)
    kp = min(K, 50)
    ctx.step(kp)
    torch.cuda.synchronize()
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    ctx.close()
    # a plain device copy (1 GiB read + 1 GiB written) and a sum (1 GiB read)
    # of known size: the calibration points of the HBM byte counters
    # (FETCH_SIZE, WRITE_SIZE) when this run is profiled
    src = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    for _ in range(3):
        dst.sum()
    torch.cuda.synchronize()
    del src, dst
    g2p2g_ms, g2p2g_n = kt.get("tile_g2p2g", (float("nan"), 1))
    avg_kernel_s = g2p2g_ms / g2p2g_n / 1e3
    alg_bytes = N * BYTES_PER_PARTICLE + BYTES_PER_NODE * active
    achieved = alg_bytes / avg_kernel_s
    step_bytes_rate = alg_bytes * K / t_tiled

    # HBM bytes per launch of the same kernel on the same workload, from the
    # rocprofv3 FETCH_SIZE / WRITE_SIZE runs stored in profiles/ (counters
    # cannot be collected inside this timed run)
    traffic, traffic_source = None, None
    tf = os.path.join(ROOT, "profiles", "traffic_tile_g2p2g_two_boxes_128.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
            traffic_source = "profiles/traffic_tile_g2p2g_two_boxes_128.json (rocprofv3 --pmc, stored)"
        except Exception:
            traffic = None

    plain = None
    if not args.no_plain:
        pctx = MPMContext(scene.params(impl=_lib.IMPL_PLAIN))
        pctx.set_particles(scene.x, scene.v)
        kpl = max(10, K // 4)
        t_plain, _ = time_steps(pctx, kpl, min(W, 5))
        pctx.close()
        plain = {"value": N * kpl / t_plain, "steps": kpl, "ms_per_step": 1e3 * t_plain / kpl}

    fwd_bwd = plain_fwd_bwd = batch_fb = None
    if not args.no_bwd:
        from physgrad_amd.scenes import batch_fb as batch_fb_scenes
        fb_scene = batch_fb_scenes(nscenes=1)[0]
        fwd_bwd = fwd_bwd_leg(fb_scene, args.bwd_steps, args.bwd_ckpt, _lib.IMPL_TILED)
        if not args.no_plain:
            plain_fwd_bwd = fwd_bwd_leg(fb_scene, args.plain_bwd_steps, args.bwd_ckpt, _lib.IMPL_PLAIN)
        del fb_scene
        if not args.no_batch_fb:
            batch_fb = batch_fb_leg(args.bwd_ckpt)
    grad = None if args.no_bwd or args.no_grad else grad_leg()
    coupling = None if args.no_coupling else coupling_leg(args.coupling_iters, not args.no_cpu)
    batch = None if args.no_batch else batch_leg(args.batch_steps)

    cpu, cpu_ref, fp64 = (None, None, None) if args.no_cpu else cpu_baseline(scene, args.cpu_steps) + \
        (fp64_error_leg(scene, 20),)

    out = {
        "metric": "particle-steps/s fwd and fwd+bwd at 1M particles; % HBM peak; grad rel err",
        "value": N * K / t_tiled,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * t_tiled / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded two-box scene, generated in-process)",
        "config": {"workload": "two_boxes_128: 128^3 grid, dx 1/128, 1,048,576 particles in two boxes at +-2 m/s, "
                               "neo-Hookean E 5e4 nu 0.3 rho 1000, dt 5e-5, sticky walls 3 nodes",
                   "particles": N, "grid": list(scene.dims), "impl": "tiled", "active_nodes": active,
                   "tuning": tuning},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_source": traffic_source,
                     "kernel": "k_tile<3> (fused G2P+P2G)", "avg_launch_us": avg_kernel_s * 1e6,
                     "alg_bytes_per_launch": alg_bytes},
        "step_effective_GBps": step_bytes_rate / 1e9,
        "kernel_ms_per_launch": {k: v[0] / v[1] for k, v in kt.items()},
        "fwd_bwd": fwd_bwd,
        "batch_fwd_bwd": batch_fb,
        "grad": grad,
        "batch": batch,
        "coupling": coupling,
        "fp64_error": fp64,
        "plain_port": plain,
        "speedup_vs_plain": (N * K / t_tiled) / plain["value"] if plain else None,
        "plain_port_fwd_bwd": plain_fwd_bwd,
        "speedup_vs_plain_fwd_bwd": fwd_bwd["value"] / plain_fwd_bwd["value"] if fwd_bwd and plain_fwd_bwd else None,
        "cpu_baseline": cpu,
        "cpu_reference": cpu_ref,
        "state_finite": finite,
        "wall_s": wall,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
