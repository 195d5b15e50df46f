"""Write the golden vectors of the reference CPU path.

Runs the reference's own MPMSolver<float,64> and SpatialHashGrid, compiled from
their headers where they lie (oracle/Makefile target `ref` -> oracle/_ref), on
stored inputs, and saves inputs and outputs as .npz fixtures here.  Needs
/root/reference; run in the build container:

    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def small_scenes():
    """(name, dims, cell, origin, dt, gravity, x, v, m) for the step fixtures."""
    rng = np.random.default_rng(20240601)
    scenes = []
    # two particles, fractional parts on both sides of 1/2 in every axis
    x2 = np.array([[0.40 + 0.2 / 16, 0.45 + 0.3 / 16, 0.50 + 0.1 / 16],
                   [0.40 + 0.7 / 16, 0.45 + 0.9 / 16, 0.50 + 0.6 / 16]], np.float32)
    v2 = np.array([[0.0, -1.0, 0.0], [0.3, 0.2, -0.5]], np.float32)
    scenes.append(("two", (16, 16, 16), (1 / 16,) * 3, (0.0, 0.0, 0.0), 1e-3, (0.0, -9.8, 0.0),
                   x2, v2, np.array([1e-3, 2e-3], np.float32)))
    # 512 particles in a box with an anisotropic grid and a shifted origin
    n = 512
    x = (np.array([0.25, 0.3, 0.35]) + 0.3 * rng.random((n, 3))).astype(np.float32)
    v = rng.normal(0.0, 0.5, (n, 3)).astype(np.float32)
    m = (1e-3 * (0.5 + rng.random(n))).astype(np.float32)
    scenes.append(("box512", (20, 18, 16), (0.05, 0.055, 0.0625), (0.01, -0.02, 0.0), 5e-4, (0.0, -9.81, 0.0),
                   x, v, m))
    return scenes


def write_step_fixtures():
    for name, dims, cell, origin, dt, grav, x, v, m in small_scenes():
        for order in (1, 2, 3):
            out = {"dims": np.array(dims, np.int32), "cell": np.array(cell, np.float32),
                   "origin": np.array(origin, np.float32), "dt": np.float32(dt),
                   "gravity": np.array(grav, np.float32), "order": np.int32(order),
                   "x0": x, "v0": v, "m": m}
            for steps in (1, 2, 10):
                prm = O.ref_params(dims, cell, origin, dt, grav, order)
                rx, rv, gm, gp, gv, _ = O.reference_run(prm, x, v, m, steps)
                out[f"x_{steps}"] = rx
                out[f"v_{steps}"] = rv
                out[f"grid_mass_{steps}"] = gm
                out[f"grid_mom_{steps}"] = gp
                out[f"grid_vel_{steps}"] = gv
            np.savez_compressed(os.path.join(HERE, f"refcompat_{name}_o{order}.npz"), **out)


def single_box_inputs():
    """64^3 grid, dx = 1/64, 32,768 particles uniform in [0.3,0.5]^3 (seed 1), v = (0,-1,0)."""
    rng = np.random.default_rng(1)
    n = 32768
    x = (0.3 + 0.2 * rng.random((n, 3))).astype(np.float32)
    v = np.tile(np.array([0.0, -1.0, 0.0], np.float32), (n, 1))
    m = np.full(n, np.float32(0.008 / n), np.float32)  # rho = 1, V0 = 0.2^3 / N
    return x, v, m


def write_single_box_fixture(nsteps=200):
    x, v, m = single_box_inputs()
    prm = O.ref_params((64, 64, 64), (1 / 64,) * 3, (0, 0, 0), 1e-4, (0, -9.8, 0), 2)
    sample = np.linspace(0, x.shape[0] - 1, 64).astype(np.int64)
    sums = np.zeros((nsteps, 10), np.float64)
    samples = {}
    cx, cv = x.copy(), v.copy()
    for s in range(1, nsteps + 1):
        cx, cv, gm, gp, gv, _ = O.reference_run(prm, cx, cv, m, 1)
        sums[s - 1] = [gm.astype(np.float64).sum(), *gp.astype(np.float64).sum(0),
                       *cx.astype(np.float64).sum(0), *cv.astype(np.float64).sum(0)]
        if s in (1, 10, 50, 100, 200):
            samples[f"x_{s}"] = cx[sample]
            samples[f"v_{s}"] = cv[sample]
    np.savez_compressed(os.path.join(HERE, "refcompat_single_box.npz"), x0=x, v0=v, m=m, sample=sample,
                        sums=sums, **samples)


def write_hash_fixtures():
    rng = np.random.default_rng(7)
    for name, cell, lo, hi in (("64", 1 / 64, (0, 0, 0), (1, 1, 1)), ("128", 1 / 128, (0, 0, 0), (1, 1, 1)),
                               ("odd", 0.07, (-0.5, 0.1, 0.0), (0.6, 0.9, 0.77))):
        n = 4096
        lo_a, hi_a = np.array(lo, np.float32), np.array(hi, np.float32)
        pos = (lo_a + (hi_a - lo_a) * rng.random((n, 3))).astype(np.float32)
        pos[:16] = hi_a - 1e-6  # clamp path: coordinates at the upper domain edge
        ids, idx, st, ct = O.reference_spatial_hash(pos, cell, lo_a, hi_a)
        np.savez_compressed(os.path.join(HERE, f"hash_{name}.npz"), pos=pos, cell=np.float32(cell),
                            lo=lo_a, hi=hi_a, ids=ids, idx=idx, start=st, count=ct)


if __name__ == "__main__":
    if not O.ref_available():
        O.build(ref=True)
    write_step_fixtures()
    write_hash_fixtures()
    write_single_box_fixture()
    print("fixtures written to", HERE)
