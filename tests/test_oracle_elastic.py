"""Pins of the oracle's elastic force from the strain energy (CPU, FP64).

The MLS-MPM stress term is the gradient of the elastic energy with respect to
the grid node positions (Hu et al. 2018, Sec. 4; Jiang et al. 2016 course
notes, Sec. 9).  These checks are written from the energy densities alone:

  fixed-corotated  psi(F) = mu |F - R|_F^2 + lam/2 (J - 1)^2       (Stomakhin et al. 2012)
  neo-Hookean      psi(F) = mu/2 (tr F^T F - 3) - mu ln J + lam/2 (ln J)^2

with R the rotation of numpy's SVD, and the MLS deformation update of a
particle when grid node i moves from x_i to x_i + d:

  F_p(d) = (I + (4/dx^2) w_ip d (x_i - x_p)^T) F_p,     E(d) = sum_p V0 psi(F_p(d))

so that one stress-only P2G (v = 0, C = 0, g = 0) must leave the momentum
p_i = -dt dE/dd at d = 0 on every node.  The derivatives are central finite
differences; the weights are the textbook quadratic B-spline N evaluated at
the node offsets.  A flipped sign or scale of the stress term, (F - R) F^T
written as F^T (F - R), F^T F in place of F F^T, or a wrong node weight each
fail one of these tests (the random F are far from symmetric, so the
transposed forms differ by O(1)).

The launcher oracle (the reference's float4 transfer kernels) is pinned per
node the same way: every node's mass after one single-particle P2G equals
m N(fx - i) N(fy - j) N(fz - k) for the quadratic and cubic splines.
"""
import ctypes

import numpy as np
import pytest

DIMS = (16, 16, 16)
DX = 1.0 / 16


def bspline2(x):
    """Quadratic B-spline N(x) (support |x| < 1.5)."""
    a = np.abs(x)
    return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def bspline3(x):
    """Cubic B-spline N(x) (support |x| < 2)."""
    a = np.abs(x)
    return np.where(a < 1.0, 0.5 * a ** 3 - a * a + 2.0 / 3.0, np.where(a < 2.0, (2.0 - a) ** 3 / 6.0, 0.0))


def lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def psi(material, mu, lam, F):
    J = np.linalg.det(F)
    if material == 0:
        U, _, Vt = np.linalg.svd(F)
        R = U @ Vt
        return mu * np.sum((F - R) ** 2) + 0.5 * lam * (J - 1) ** 2
    lnJ = np.log(J)
    return 0.5 * mu * (np.trace(F.T @ F) - 3) - mu * lnJ + 0.5 * lam * lnJ ** 2


def random_F(rng, scale):
    while True:
        F = np.eye(3) + scale * rng.normal(size=(3, 3))
        if np.linalg.det(F) > 0.4 and np.abs(F - F.T).max() > 0.2:
            return F


def fd_piola(material, mu, lam, F, h=1e-6):
    P = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            Fp, Fm = F.copy(), F.copy()
            Fp[i, j] += h
            Fm[i, j] -= h
            P[i, j] = (psi(material, mu, lam, Fp) - psi(material, mu, lam, Fm)) / (2 * h)
    return P


@pytest.mark.parametrize("material", [0, 1], ids=["fixed_corotated", "neo_hookean"])
def test_kirchhoff_is_energy_gradient_times_FT(oracle_lib, material):
    """tau(F) = (d psi / d F) F^T for random non-symmetric F."""
    O = oracle_lib
    rng = np.random.default_rng(40 + material)
    mu, lam = lame(2e4, 0.3)
    for _ in range(12):
        F = random_F(rng, 0.35)
        P = fd_piola(material, mu, lam, F)
        tau = O.kirchhoff(material, mu, lam, F)
        scale = np.abs(P @ F.T).max()
        np.testing.assert_allclose(tau, P @ F.T, atol=1e-6 * scale)
        # the transposed product is far away: the check discriminates
        assert np.abs(F.T @ P - P @ F.T).max() > 1e-2 * scale


def stress_only_scene(rng, material, n):
    x = 0.35 + 0.3 * rng.random((n, 3))
    F = np.stack([random_F(rng, 0.25) for _ in range(n)])
    E = 1e4 * (0.5 + rng.random(n))
    mass = 1e-3 * (0.5 + rng.random(n))
    return x, F, E, mass


@pytest.mark.parametrize("material", [0, 1], ids=["fixed_corotated", "neo_hookean"])
def test_stress_p2g_is_minus_dt_energy_gradient(oracle_lib, material):
    """One stress-only P2G leaves p_i = -dt dE/dx_i on every node touched."""
    O = oracle_lib
    rng = np.random.default_rng(50 + material)
    n, nu, dt, p_vol = 5, 0.3, 1e-4, 2e-5
    x, F, E, mass = stress_only_scene(rng, material, n)
    prm = O.mls_params(DIMS, DX, (0, 0, 0), dt, (0, 0, 0), 0, material, nu, p_vol)
    out = O.mls_run(prm, x, np.zeros((n, 3)), None, F, mass, E, 1, want_grid=True)
    gm = out["grid_m"]
    p_grid = out["grid_v"] * gm[:, None]          # g = 0, no walls: v_i = p_i / m_i
    nodes = np.nonzero(gm > 0)[0]
    assert len(nodes) >= 27
    # per particle: base cell, and the weight of a node (textbook B-spline)
    xs = x / DX
    base = np.floor(xs - 0.5).astype(int)
    mu_lam = [lame(E[q], nu) for q in range(n)]

    def energy(node_ijk, d):
        e = 0.0
        xi = np.array(node_ijk, float) * DX
        for q in range(n):
            off = np.array(node_ijk) - base[q]
            if np.any(off < 0) or np.any(off > 2):
                Fq = F[q]
            else:
                w = np.prod(bspline2(xs[q] - np.array(node_ijk)))
                Fq = (np.eye(3) + (4 / DX ** 2) * w * np.outer(d, xi - x[q])) @ F[q]
            e += p_vol * psi(material, *mu_lam[q], Fq)
        return e

    nx, ny = DIMS[0], DIMS[1]
    worst = 0.0
    pmax = np.abs(p_grid[nodes]).max()
    assert pmax > 0
    for nid in nodes:
        ijk = (nid % nx, (nid // nx) % ny, nid // (nx * ny))
        grad = np.zeros(3)
        h = 1e-7
        for c in range(3):
            d = np.zeros(3)
            d[c] = h
            grad[c] = (energy(ijk, d) - energy(ijk, -d)) / (2 * h)
        worst = max(worst, np.abs(p_grid[nid] + dt * grad).max())
    assert worst <= 1e-6 * pmax, (worst, pmax)


def test_stress_p2g_sign_and_scale_single_particle(oracle_lib):
    """A single stretched particle (F = s I, s > 1) is in tension and pulls
    its nodes inwards: the stress momentum on a node is
    -dt V0 (4/dx^2) w tau (x_i - x_p) with tau = t I, t > 0."""
    O = oracle_lib
    s, E, nu, dt, p_vol = 1.1, 1e4, 0.2, 1e-4, 3e-5
    mu, lam = lame(E, nu)
    xp = np.array([[0.503, 0.497, 0.51]])
    F = (s * np.eye(3))[None]
    for material in (0, 1):
        prm = O.mls_params(DIMS, DX, (0, 0, 0), dt, (0, 0, 0), 0, material, nu, p_vol)
        out = O.mls_run(prm, xp, np.zeros((1, 3)), None, F, 1e-3, E, 1, want_grid=True)
        gm = out["grid_m"]
        p_grid = out["grid_v"] * gm[:, None]
        t = (2 * mu * (s - 1) * s + lam * (s ** 3 - 1) * s ** 3) if material == 0 else \
            (mu * (s * s - 1) + lam * 3 * np.log(s))
        for nid in np.nonzero(gm > 0)[0]:
            ijk = np.array([nid % 16, (nid // 16) % 16, nid // 256])
            r = ijk * DX - xp[0]
            w = np.prod(bspline2(xp[0] / DX - ijk))
            assert t > 0
            np.testing.assert_allclose(p_grid[nid], -dt * p_vol * 4 / DX ** 2 * w * t * r, rtol=1e-9, atol=1e-18)


_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


@pytest.mark.parametrize("order", [2, 3])
def test_launcher_p2g_node_weights(oracle_lib, order):
    """Single-particle launcher P2G: each node (base + offset) holds
    m N(f - offset) per axis product; no other node is touched."""
    L = oracle_lib.lib()
    L.orc_launch_p2g.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p, ctypes.c_int, ctypes.c_int]
    rng = np.random.default_rng(60 + order)
    o = np.array([0.05, -0.1, 0.0], np.float32)
    h = np.array([0.06, 0.07, 0.05], np.float32)
    d = np.array([12, 10, 14], np.int32)
    N = bspline2 if order == 2 else bspline3
    offs = np.arange(-1, 2) if order == 2 else np.arange(-2, 2)
    for _ in range(6):
        pos = np.zeros((1, 4), np.float32)
        pos[0, :3] = o + (3 + (d - 7) * rng.random(3)) * h
        pos[0, 3] = 2e-3
        vel = np.zeros((1, 4), np.float32)
        G = int(np.prod(d))
        gm = np.zeros(G, np.float32)
        gp = np.zeros(3 * G, np.float32)
        L.orc_launch_p2g(pos.reshape(-1), vel.reshape(-1), gm, gp, o, h, d, 1, order)
        t = (pos[0, :3].astype(np.float64) - o) / h
        b = np.floor(t).astype(int)
        f = t - b
        want = np.zeros(G)
        for di in offs:
            for dj in offs:
                for dk in offs:
                    w = N(f[0] - di) * N(f[1] - dj) * N(f[2] - dk)
                    if w > 1e-10:
                        i, j, k = b[0] + di, b[1] + dj, b[2] + dk
                        want[i + j * d[0] + k * d[0] * d[1]] = 2e-3 * w
        np.testing.assert_allclose(gm, want, rtol=0, atol=1e-6 * want.max())
