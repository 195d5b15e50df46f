"""ref-compat oracle pinned against the reference's own CPU solver.

The golden vectors come from the reference MPMSolver<float,64> compiled from
its header (tests/golden/make_golden.py); the C restatement must reproduce
them bit for bit.  Closed forms pin the semantics independently of the
fixtures: the reference stencil's per-axis weight sum s(f) and the
single-particle update v' = S (v + g dt).
"""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN

STEP_FILES = sorted(glob.glob(os.path.join(GOLDEN, "refcompat_*_o*.npz")))


@pytest.mark.parametrize("path", STEP_FILES, ids=[os.path.basename(p) for p in STEP_FILES])
def test_restatement_bitwise_vs_reference_fixtures(oracle_lib, path):
    O = oracle_lib
    f = np.load(path)
    assert len(STEP_FILES) >= 6
    for steps in (1, 2, 10):
        prm = O.ref_params(f["dims"], f["cell"], f["origin"], f["dt"], f["gravity"], int(f["order"]))
        x, v, gm, gp, gv = O.refcompat_run(prm, f["x0"], f["v0"], f["m"], steps)
        for got, key in ((x, "x"), (v, "v"), (gm, "grid_mass"), (gp, "grid_mom"), (gv, "grid_vel")):
            want = f[f"{key}_{steps}"]
            assert np.array_equal(got.reshape(want.shape).view(np.uint32), want.view(np.uint32)), (key, steps)


def test_single_box_checksums_and_samples(oracle_lib):
    """Single-box geometry (64^3, 32,768 particles): per-step sums and sampled
    particles of the reference for 200 steps, bit for bit on the samples."""
    O = oracle_lib
    f = np.load(os.path.join(GOLDEN, "refcompat_single_box.npz"))
    prm = O.ref_params((64, 64, 64), (1 / 64,) * 3, (0, 0, 0), 1e-4, (0, -9.8, 0), 2)
    x, v, m = f["x0"], f["v0"], f["m"]
    checkpoints = (1, 10, 50)
    last = 0
    for s in checkpoints:
        x, v, gm, gp, gv = O.refcompat_run(prm, x, v, m, s - last)
        last = s
        sums = [gm.astype(np.float64).sum(), *gp.astype(np.float64).sum(0),
                *x.astype(np.float64).sum(0), *v.astype(np.float64).sum(0)]
        np.testing.assert_array_equal(np.array(sums), f["sums"][s - 1])
        assert np.array_equal(x[f["sample"]], f[f"x_{s}"])
        assert np.array_equal(v[f["sample"]], f[f"v_{s}"])


def _axis_sum(fx):
    """Reference stencil weight sum per axis: nodes base-1..base+1, f = frac part."""
    def N(t):
        a = abs(t)
        return 0.75 - t * t if a < 0.5 else (0.5 * (1.5 - a) ** 2 if a < 1.5 else 0.0)
    return sum(N(fx - d) for d in (-1, 0, 1))


@pytest.mark.parametrize("f", [0.1, 0.3, 0.49, 0.5, 0.7, 0.95])
def test_stencil_defect_closed_form(f):
    # s(f) = 1 for f < 1/2, 1 - (f - 1/2)^2 / 2 for f >= 1/2
    want = 1.0 if f < 0.5 else 1.0 - 0.5 * (f - 0.5) ** 2
    assert abs(_axis_sum(f) - want) < 1e-12


def test_single_particle_closed_form(oracle_lib):
    """One isolated particle: v' = S (v + g dt), x' = x + v' dt, grid mass = m S,
    with S the product of the per-axis weight sums (all weights > 1e-10 here)."""
    O = oracle_lib
    dx = 1 / 16
    fr = np.array([0.2, 0.7, 0.55])
    x = ((np.array([6, 7, 8]) + fr) * dx).astype(np.float32)[None]
    v = np.array([[0.3, -1.0, 0.25]], np.float32)
    m = np.array([2e-3], np.float32)
    dt, g = 1e-3, np.array([0.0, -9.8, 0.0])
    prm = O.ref_params((16, 16, 16), (dx,) * 3, (0, 0, 0), dt, g, 2)
    x1, v1, gm, _, _ = O.refcompat_run(prm, x, v, m, 1)
    fr32 = (x[0].astype(np.float64) / dx) - np.floor(x[0].astype(np.float64) / dx)
    S = np.prod([_axis_sum(f) for f in fr32])
    np.testing.assert_allclose(v1[0], S * (v[0] + g * dt), rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(x1[0], x[0] + v1[0] * dt, rtol=1e-7)
    np.testing.assert_allclose(gm.sum(), m[0] * S, rtol=2e-6)


def test_mutation_caught_by_fixture(oracle_lib):
    """A planted mistake (gravity sign flipped) in the oracle's inputs must fail
    the fixture comparison: the fixtures pin the gravity term."""
    O = oracle_lib
    f = np.load(os.path.join(GOLDEN, "refcompat_box512_o2.npz"))
    g = -f["gravity"]
    prm = O.ref_params(f["dims"], f["cell"], f["origin"], f["dt"], g, 2)
    _, v, _, _, _ = O.refcompat_run(prm, f["x0"], f["v0"], f["m"], 1)
    assert not np.array_equal(v, f["v_1"])
