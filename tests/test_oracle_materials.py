"""Pins of the oracle's per-particle materials and per-node boundary flags
(oracle/mlsmpm_ext.c), CPU only.

  * With every per-particle array equal to the scene's values, or absent, the
    extended step is bit-identical to the pinned single-material step.
  * P2G is linear in the particles, so the grid momentum of a mixed scene
    (fixed-corotated, neo-Hookean and fluid particles with their own Poisson
    ratios, viscosities and volumes) is the sum of the grids of its particles
    run alone: this pins that each particle uses its own material.
  * The Newtonian fluid stress (src/mpm_kernels.cu:131-149, tau = J sigma):
    zero for a rigid rotation rate, symmetric, dissipative (tau : C >= 0),
    linear in the viscosity and proportional to J; and a fluid in a rigid
    rotation field keeps that field exactly, like an APIC transfer without
    stress.
  * Boundary flags: DIRICHLET_Y zeroes exactly the y velocity of every
    particle and leaves x, z as in the unflagged step; STICKY zeroes all;
    SLIP removes exactly the outgoing components; NEUMANN bits change nothing.
"""
import numpy as np
import pytest

DIMS = (16, 16, 16)
DX = 1.0 / 16
BC_DX, BC_DY, BC_DZ, BC_NX, BC_NY, BC_NZ, BC_STICKY, BC_SLIP = 1, 2, 4, 8, 16, 32, 64, 128


def params(O, g=(0, 0, 0), bc=0, material=0, nu=0.3, p_vol=2e-5):
    return O.mls_params(DIMS, DX, (0, 0, 0), 1e-4, g, bc, material, nu, p_vol)


def random_state(rng, n, lo=0.35, span=0.3):
    x = lo + span * rng.random((n, 3))
    v = rng.normal(0, 0.5, (n, 3))
    C = rng.normal(0, 2.0, (n, 3, 3))
    F = np.eye(3) + 0.1 * rng.normal(size=(n, 3, 3))
    return x, v, C, F


@pytest.mark.parametrize("material", [0, 1])
def test_ext_reduces_to_single_material_step(oracle_lib, material):
    O = oracle_lib
    rng = np.random.default_rng(80 + material)
    n = 200
    x, v, C, F = random_state(rng, n)
    prm = params(O, g=(0, -9.8, 0), bc=3, material=material)
    E = 1e4 * (0.5 + rng.random(n))
    m = 1e-3 * (0.5 + rng.random(n))
    want = O.mls_run(prm, x, v, C, F, m, E, 3, want_grid=True)
    for kw in ({}, {"model": np.full(n, material), "nu": np.full(n, 0.3), "vol": np.full(n, 2e-5),
                    "visc": np.zeros(n), "flags": np.zeros(np.prod(DIMS), np.uint32)},
               {"flags": np.full(np.prod(DIMS), BC_NX | BC_NY | BC_NZ, np.uint32)}):
        got = O.mls_run_ext(prm, x, v, C, F, m, E, 3, want_grid=True, **kw)
        for k in ("x", "v", "C", "F", "grid_m", "grid_v"):
            assert np.array_equal(got[k], want[k]), (k, list(kw))


def test_mixed_materials_superpose(oracle_lib):
    O = oracle_lib
    rng = np.random.default_rng(82)
    n = 12
    x, v, C, F = random_state(rng, n)
    prm = params(O)
    model = np.array([0, 1, 2] * 4)
    nu = rng.uniform(0.1, 0.45, n)
    visc = rng.uniform(0.5, 2.0, n)
    vol = 1e-5 * (0.5 + rng.random(n))
    E = 1e4 * (0.5 + rng.random(n))
    m = 1e-3 * (0.5 + rng.random(n))
    mixed = O.mls_run_ext(prm, x, v, C, F, m, E, 1, model=model, nu=nu, visc=visc, vol=vol, want_grid=True)
    gm = np.zeros(np.prod(DIMS))
    gp = np.zeros((np.prod(DIMS), 3))
    for q in range(n):
        one = O.mls_run_ext(prm, x[q:q + 1], v[q:q + 1], C[q:q + 1], F[q:q + 1], m[q:q + 1], E[q:q + 1], 1,
                            model=model[q:q + 1], nu=nu[q:q + 1], visc=visc[q:q + 1], vol=vol[q:q + 1],
                            want_grid=True)
        gm += one["grid_m"]
        gp += one["grid_v"] * one["grid_m"][:, None]
    np.testing.assert_allclose(mixed["grid_m"], gm, rtol=1e-13, atol=1e-20)
    pm = mixed["grid_v"] * mixed["grid_m"][:, None]
    np.testing.assert_allclose(pm, gp, rtol=1e-10, atol=1e-12 * np.abs(gp).max())
    # and each material matters: the same scene with one model for all differs
    other = O.mls_run_ext(prm, x, v, C, F, m, E, 1, model=np.zeros(n, int), nu=nu, visc=visc, vol=vol,
                          want_grid=True)
    assert np.abs(other["grid_v"] - mixed["grid_v"]).max() > 1e-6


def test_fluid_stress_properties(oracle_lib):
    O = oracle_lib
    rng = np.random.default_rng(83)
    for _ in range(10):
        F = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
        J = np.linalg.det(F)
        if J < 0.3:
            continue
        W = rng.normal(size=(3, 3))
        W = W - W.T                      # rigid rotation rate
        S = rng.normal(size=(3, 3))
        C = S + W
        tw = O.kirchhoff_ext(2, 0, 0, 1.7, F, W)
        assert np.abs(tw).max() < 1e-12
        t = O.kirchhoff_ext(2, 0, 0, 1.7, F, C)
        np.testing.assert_allclose(t, t.T, atol=1e-12)
        assert (t * C).sum() >= 0.0
        np.testing.assert_allclose(O.kirchhoff_ext(2, 0, 0, 3.4, F, C), 2 * t, rtol=1e-12)
        np.testing.assert_allclose(O.kirchhoff_ext(2, 0, 0, 1.7, np.eye(3), C) * J, t, rtol=1e-12, atol=1e-12)
        # trace-free shear only: the model carries no pressure (as the reference's)
        np.testing.assert_allclose(np.trace(O.kirchhoff_ext(2, 0, 0, 1.0, np.eye(3), W + np.diag([1, -1, 0]))),
                                   0.0, atol=1e-12)


def test_fluid_keeps_a_rigid_rotation(oracle_lib):
    O = oracle_lib
    rng = np.random.default_rng(84)
    n = 300
    x = 0.35 + 0.3 * rng.random((n, 3))
    W = np.array([[0, -0.7, 0.3], [0.7, 0, -0.2], [-0.3, 0.2, 0]])
    v = (x - 0.5) @ W.T
    C = np.tile(W, (n, 1, 1))
    prm = params(O)
    out = O.mls_run_ext(prm, x, v, C, None, 1e-3, 0.0, 1, model=np.full(n, 2), visc=np.full(n, 5.0))
    np.testing.assert_allclose(out["v"], v, atol=1e-12)
    np.testing.assert_allclose(out["C"], C, atol=1e-10)


def test_boundary_flags(oracle_lib):
    O = oracle_lib
    rng = np.random.default_rng(85)
    n = 200
    x, v, C, F = random_state(rng, n)
    prm = params(O, g=(0, -9.8, 0))
    G = int(np.prod(DIMS))
    E, m = 1e4, 1e-3
    free = O.mls_run_ext(prm, x, v, C, F, m, E, 1)
    dy = O.mls_run_ext(prm, x, v, C, F, m, E, 1, flags=np.full(G, BC_DY, np.uint32))
    assert np.all(dy["v"][:, 1] == 0.0)
    np.testing.assert_allclose(dy["v"][:, [0, 2]], free["v"][:, [0, 2]], rtol=0, atol=0)
    st = O.mls_run_ext(prm, x, v, C, F, m, E, 1, flags=np.full(G, BC_STICKY, np.uint32))
    assert np.all(st["v"] == 0.0) and np.all(st["x"] == x)
    # SLIP: a blob in the lower-x half moving to -x and +y: x velocity removed
    # (it points out through the x = 0 wall), y kept (it points away from y = 0)
    xs = 0.1 + 0.25 * rng.random((n, 3))
    vs = np.tile([-1.0, 0.5, 0.0], (n, 1))
    sl = O.mls_run_ext(prm, xs, vs, None, None, m, 0.0, 1, flags=np.full(G, BC_SLIP, np.uint32))
    g0 = O.mls_run_ext(O.mls_params(DIMS, DX, (0, 0, 0), 1e-4, (0, 0, 0), 0, 0, 0.3, 2e-5), xs, vs, None, None, m,
                       0.0, 1, flags=np.full(G, BC_SLIP, np.uint32))
    np.testing.assert_allclose(g0["v"], np.tile([0.0, 0.5, 0.0], (n, 1)), atol=1e-12)
    assert np.all(sl["v"][:, 0] == 0.0)
    # the same blob in the upper-x half moving to -x (away from the x = 1 wall) keeps its x velocity
    up = O.mls_run_ext(O.mls_params(DIMS, DX, (0, 0, 0), 1e-4, (0, 0, 0), 0, 0, 0.3, 2e-5), xs + [0.5, 0, 0], vs,
                       None, None, m, 0.0, 1, flags=np.full(G, BC_SLIP, np.uint32))
    np.testing.assert_allclose(up["v"], vs, atol=1e-12)
