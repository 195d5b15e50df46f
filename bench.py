"""Benchmark of the MLS-MPM transfer loop on one MI355X.

Workload (two_boxes, 128^3 grid, dx 1/128, 1,048,576 particles, neo-Hookean
E 5e4 nu 0.3 rho 1000, dt 5e-5): a step is one P2G -> grid update -> G2P pass
over all particles with the particle state resident in HBM.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line: particle-steps/s of the tiled path (value), the plain
GPU port measured in the same run (forward, and forward+backward), the roofline of the dominant kernel (the
fused G2P+P2G tile kernel, timed with HIP events on its launch stream) and the
CPU baseline (FP32 port of the same step on one host core, bounded sample).
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
COPY_BYTES = 256 << 20     # size of the calibration copy


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
    ap.add_argument("--bwd-steps", type=int, default=64)
    # all 64 steps on the segment tape (~11 GB at 1M particles): the backward
    # reverses them without recomputing the forward
    ap.add_argument("--bwd-ckpt", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batch leg")
    ap.add_argument("--batch-steps", type=int, default=500)
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


def fwd_bwd_leg(scene, steps, ckpt, impl):
    """Forward with a checkpoint every `ckpt` steps plus the full backward (recompute
    and reverse of every segment) on the same 1M-particle workload; the adjoint
    seed is that of the centre-of-mass loss."""
    import torch
    from physgrad_amd.solver import MPMContext
    ctx = MPMContext(scene.params(impl=impl))
    ctx.set_particles(scene.x, scene.v)
    n = scene.n
    seed = torch.full((n, 3), 1.0 / n, dtype=torch.float32, device="cuda")
    ctx.forward_checkpointed(ckpt, ckpt)   # warm-up: allocates checkpoints and the tape
    ctx.backward(seed)
    ctx.set_particles(scene.x, scene.v)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    a.record(s)
    ctx.forward_checkpointed(steps, ckpt)
    b.record(s)
    ctx.backward(seed)
    c.record(s)
    torch.cuda.synchronize()
    t_f, t_b = a.elapsed_time(b) / 1e3, b.elapsed_time(c) / 1e3
    ctx.close()
    return {"value": n * steps / (t_f + t_b), "unit": "particle-steps/s", "steps": steps, "checkpoint_every": ckpt,
            "ms_per_step": 1e3 * (t_f + t_b) / steps, "forward_ms_per_step": 1e3 * t_f / steps,
            "backward_ms_per_step": 1e3 * t_b / steps,
            "workload": "two_boxes_128 (1,048,576 particles, 128^3), centre-of-mass adjoint seed"}


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
    b.close()
    return {"value": n * steps / t_graph, "unit": "particle-steps/s", "scenes": len(scenes), "particles": n,
            "steps": steps, "ms_per_step": 1e3 * t_graph / steps, "ms_per_step_eager": 1e3 * t_eager / steps,
            "tile_workgroups": tuned, "state_finite": ok,
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
    ctx.close()
    return {"rel_err_dv0": err, "dv0": got.tolist(), "dv0_closed_form": want.tolist(),
            "max_abs_dlogE": gE, "fwd_bwd_wall_s": wall,
            "workload": "rollout_64: 64^3, 262,144 particles, fixed-corotated, per-particle E log-uniform [5e3,5e4], "
                        "512 steps at dt 1e-4, checkpoint every 32"}


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
    # a plain device copy of known size (256 MiB read + 256 MiB written): the
    # calibration point of the HBM byte counters when this run is profiled
    src = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    g2p2g_ms, g2p2g_n = kt.get("tile_g2p2g", (float("nan"), 1))
    avg_kernel_s = g2p2g_ms / g2p2g_n / 1e3
    alg_bytes = N * BYTES_PER_PARTICLE + BYTES_PER_NODE * active
    achieved = alg_bytes / avg_kernel_s
    step_bytes_rate = alg_bytes * K / t_tiled

    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_tile_g2p2g_two_boxes_128.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
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

    fwd_bwd = None if args.no_bwd else fwd_bwd_leg(scene, args.bwd_steps, args.bwd_ckpt, _lib.IMPL_TILED)
    plain_fwd_bwd = None
    if not args.no_bwd and not args.no_plain:
        k = min(args.bwd_steps, args.bwd_ckpt)
        plain_fwd_bwd = fwd_bwd_leg(scene, k, k, _lib.IMPL_PLAIN)
    grad = None if args.no_bwd or args.no_grad else grad_leg()
    batch = None if args.no_batch else batch_leg(args.batch_steps)

    cpu, cpu_ref = (None, None) if args.no_cpu else cpu_baseline(scene, args.cpu_steps)

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
                   "particles": N, "grid": list(scene.dims), "impl": "tiled", "active_nodes": active},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "k_tile<3> (fused G2P+P2G)", "avg_launch_us": avg_kernel_s * 1e6,
                     "alg_bytes_per_launch": alg_bytes},
        "step_effective_GBps": step_bytes_rate / 1e9,
        "kernel_ms_per_launch": {k: v[0] / v[1] for k, v in kt.items()},
        "fwd_bwd": fwd_bwd,
        "grad": grad,
        "batch": batch,
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
