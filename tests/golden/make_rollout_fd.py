"""Write the FP64 central finite differences of the differentiable-rollout
workload (rollout_64: 64^3, 262,144 particles, fixed-corotated, per-particle
E, 512 steps at dt 1e-4) for an E-sensitive loss.

With the workload's own initial state (all particles at rest, then free fall
far from the walls) the body moves rigidly, F stays I and no loss depends on
E.  The differences are therefore taken around a deforming state: the initial
velocity is v0 + A (x - c) with the fixed non-symmetric rate matrix A below
(rotation, shear and dilation of about 0.1 m/s at the box faces) and c the
initial centre of mass, and the loss weighs every final position:

    L = sum_p w_p . x_p(T),   w = default_rng(33).normal(size=(n, 3)) / n

with the shared v0 (zero at the operating point) and the per-particle Young's
moduli as parameters.  Differences are taken with the
FP64 oracle (oracle/liboracle.so): v0 components with the absolute step 1e-4,
E of 16 sampled particles (default_rng(34)) with the relative step 1e-4
(E (1 +- 1e-4)), and, as a check of the step, one E parameter and one v0
component again with twice the step.  Each evaluation is 512 FP64 steps of
the full workload (about 4 minutes on one core); the evaluations run in
parallel processes.  Run in the build container:

    python tests/golden/make_rollout_fd.py [processes]

Output: tests/golden/rollout_fd.npz.
"""
from __future__ import annotations

import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

EPS = 1e-4
N_E = 16
A_RATE = np.array([[0.8, 0.3, -0.2], [-0.4, -0.6, 0.25], [0.2, -0.35, 0.5]])   # 1/s


def setup():
    from physgrad_amd.scenes import rollout
    sc = rollout()
    w = np.random.default_rng(33).normal(size=(sc.n, 3)) / sc.n
    idx = np.sort(np.random.default_rng(34).choice(sc.n, N_E, replace=False))
    return sc, w, idx


def loss(task):
    kind, i, h = task
    from oracle import oracle as O
    sc, w, _ = setup()
    prm = O.scene_params(sc)
    x0 = sc.x.astype(np.float64)
    v = (x0 - x0.mean(0)) @ A_RATE.T
    E = sc.E_particles.astype(np.float64).copy()
    if kind == "v":
        v[:, i] += h
    elif kind == "E":
        E[i] *= 1.0 + h
    out = O.mls_run(prm, x0, v, None, None, sc.p_mass, E, sc.steps)
    return task, float((w * out["x"]).sum())


def main():
    procs = int(sys.argv[1]) if len(sys.argv) > 1 else 6
    sc, w, idx = setup()
    tasks = [("base", 0, 0.0)]
    for c in range(3):
        tasks += [("v", c, EPS), ("v", c, -EPS)]
    for i in idx:
        tasks += [("E", int(i), EPS), ("E", int(i), -EPS)]
    check = [("v", 0), ("E", int(idx[0]))]
    for kind, i in check:
        tasks += [(kind, i, 2 * EPS), (kind, i, -2 * EPS)]
    t0 = time.time()
    with mp.Pool(procs) as pool:
        res = {}
        for task, val in pool.imap_unordered(loss, tasks):
            res[task] = val
            print(f"{len(res)}/{len(tasks)} {task} {val!r} ({time.time() - t0:.0f} s)", flush=True)
    dv0 = np.array([(res[("v", c, EPS)] - res[("v", c, -EPS)]) / (2 * EPS) for c in range(3)])
    # relative step: dL/dE_p = (L(E(1+h)) - L(E(1-h))) / (2 h E_p)
    E = sc.E_particles.astype(np.float64)
    dE = np.array([(res[("E", int(i), EPS)] - res[("E", int(i), -EPS)]) / (2 * EPS * E[i]) for i in idx])
    d2 = []
    for kind, i in check:
        scale = 1.0 if kind == "v" else E[i]
        d2.append((res[(kind, i, 2 * EPS)] - res[(kind, i, -2 * EPS)]) / (4 * EPS * scale))
    np.savez(os.path.join(HERE, "rollout_fd.npz"), loss=res[("base", 0, 0.0)], dv0=dv0, E_index=idx, dE=dE,
             eps=EPS, check_kind=np.array([k for k, _ in check]), check_index=np.array([i for _, i in check]),
             check_2eps=np.array(d2), weights_seed=33, index_seed=34, A_rate=A_RATE)
    print("dv0", dv0)
    print("dE", dE)
    print("step check (eps, 2 eps):", [(dv0[0], d2[0]), (dE[0], d2[1])])


if __name__ == "__main__":
    main()
