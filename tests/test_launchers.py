"""Reference launchers (launchOptimizedP2G/G2P/FullyFusedG2P2G semantics).

CPU: the oracle restatement against closed forms (the CUDA source does not
compile anywhere in this pipeline, so these closed forms are the pin):
partition-of-unity defect of the reference stencil, normalised G2P returning
the grid velocity of a single particle, the FLIP blend.
GPU: the HIP launchers against the oracle at rel 1e-5 (float atomics reorder
the P2G sums only).
"""
import ctypes

import numpy as np
import pytest

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def olib(O):
    L = O.lib()
    L.orc_launch_p2g.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p, ctypes.c_int, ctypes.c_int]
    L.orc_launch_g2p.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p, ctypes.c_int, ctypes.c_float,
                                 ctypes.c_float, ctypes.c_int]
    L.orc_launch_fused.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p, ctypes.c_int,
                                   ctypes.c_float, _f32p, ctypes.c_float, ctypes.c_int]
    return L


def scene(rng, n, dims=(12, 10, 14)):
    o = np.array([0.05, -0.1, 0.0], np.float32)
    h = np.array([0.06, 0.07, 0.05], np.float32)
    d = np.array(dims, np.int32)
    lo = o + 2 * h
    hi = o + (d - 3) * h
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = lo + (hi - lo) * rng.random((n, 3))
    pos[:, 3] = 1e-3 * (0.5 + rng.random(n))
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.5, (n, 3))
    return pos, vel, o, h, d


@pytest.mark.parametrize("order", [2, 3])
def test_single_particle_closed_form(oracle_lib, order):
    """One particle: grid mass = m S (S = weight sum over nodes with w > 1e-10),
    grid velocity = v everywhere it has mass, so the normalised G2P returns v and
    the blend gives v; position moves by v dt."""
    L = olib(oracle_lib)
    rng = np.random.default_rng(order)
    pos, vel, o, h, d = scene(rng, 1)
    G = int(np.prod(d))
    gm = np.zeros(G, np.float32)
    gp = np.zeros(3 * G, np.float32)
    L.orc_launch_p2g(pos.reshape(-1), vel.reshape(-1), gm, gp, o, h, d, 1, order)
    f = (pos[0, :3].astype(np.float64) - o) / h
    f -= np.floor(f)

    def w(x):
        a = abs(x)
        if order == 2:
            return 0.75 - x * x if a < 0.5 else (0.5 * (1.5 - a) ** 2 if a < 1.5 else 0.0)
        return 2 / 3 - a * a + 0.5 * a ** 3 if a < 1 else ((2 - a) ** 3 / 6 if a < 2 else 0.0)
    offs = (-1, 0, 1) if order == 2 else (-2, -1, 0, 1)
    S = np.prod([sum(w(f[a] - k) for k in offs) for a in range(3)])
    np.testing.assert_allclose(gm.sum(), pos[0, 3] * S, rtol=1e-5)
    gv = np.where(gm[:, None] > 1e-10, gp.reshape(-1, 3) / np.maximum(gm[:, None], 1e-30), 0).astype(np.float32)
    p2, v2 = pos.copy(), vel.copy()
    L.orc_launch_g2p(p2.reshape(-1), v2.reshape(-1), gm, gv.reshape(-1), o, h, d, 1, 1e-3, 0.3, order)
    np.testing.assert_allclose(v2[0, :3], vel[0, :3], rtol=2e-5)
    np.testing.assert_allclose(p2[0, :3], pos[0, :3] + 1e-3 * v2[0, :3], rtol=1e-6)


def test_flip_blend_and_fused_gravity(oracle_lib):
    """Fused launch, one particle: gravity enters the scattered velocity, the
    gathered grid velocity is v + g dt, and the blend is flip*v + (1-flip)(v + g dt)."""
    L = olib(oracle_lib)
    rng = np.random.default_rng(11)
    pos, vel, o, h, d = scene(rng, 1)
    G = int(np.prod(d))
    gm, gp, gv = np.zeros(G, np.float32), np.zeros(3 * G, np.float32), np.zeros(3 * G, np.float32)
    g = np.array([0.0, -9.8, 0.0], np.float32)
    flip, dt = 0.95, 1e-3
    v0 = vel[0, :3].copy()
    L.orc_launch_fused(pos.reshape(-1), vel.reshape(-1), gm, gp, gv, o, h, d, 1, dt, g, flip, 2)
    np.testing.assert_allclose(vel[0, :3], flip * v0 + (1 - flip) * (v0 + g * dt), rtol=1e-5, atol=1e-7)


def _gpu_call(fn, arrays, *scalars_after):
    import torch
    dev = {k: torch.as_tensor(v).cuda() for k, v in arrays.items()}
    return dev


@pytest.mark.gpu
@pytest.mark.parametrize("order", [2, 3])
def test_gpu_launchers_vs_oracle(gpu, oracle_lib, order):
    import torch
    from physgrad_amd import _lib
    Lg = _lib.load()
    L = olib(oracle_lib)
    rng = np.random.default_rng(20 + order)
    n = 4000
    pos, vel, o, h, d = scene(rng, n)
    G = int(np.prod(d))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    o_c = (ctypes.c_float * 3)(*o)
    h_c = (ctypes.c_float * 3)(*h)
    d_c = (ctypes.c_int * 3)(*d)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    def rel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))

    # P2G into a zeroed grid
    gm, gp = np.zeros(G, np.float32), np.zeros(3 * G, np.float32)
    L.orc_launch_p2g(pos.reshape(-1), vel.reshape(-1), gm, gp, o, h, d, n, order)
    dpos, dvel = torch.as_tensor(pos).cuda(), torch.as_tensor(vel).cuda()
    dgm, dgp = torch.zeros(G, device="cuda"), torch.zeros(3 * G, device="cuda")
    _lib.check(Lg.pg_launch_optimized_p2g(P(dpos), P(dvel), P(dgm), P(dgp), o_c, h_c, d_c, n, order, stream))
    torch.cuda.synchronize()
    assert rel(dgm.cpu().numpy(), gm) <= 1e-5
    assert rel(dgp.cpu().numpy(), gp) <= 1e-5
    # G2P from the oracle's grid velocity
    gv = np.where(gm[:, None] > 1e-10, gp.reshape(-1, 3) / np.maximum(gm[:, None], 1e-30), 0)
    gv = gv.astype(np.float32).reshape(-1)
    p2, v2 = pos.copy(), vel.copy()
    L.orc_launch_g2p(p2.reshape(-1), v2.reshape(-1), gm, gv, o, h, d, n, 1e-3, 0.9, order)
    dgm2, dgv = torch.as_tensor(gm).cuda(), torch.as_tensor(gv).cuda()
    _lib.check(Lg.pg_launch_optimized_g2p(P(dpos), P(dvel), P(dgm2), P(dgv), o_c, h_c, d_c, n, 1e-3, 0.9, order,
                                          stream))
    torch.cuda.synchronize()
    assert rel(dpos.cpu().numpy(), p2) <= 1e-6
    assert rel(dvel.cpu().numpy(), v2) <= 1e-5
    # fused
    p3, v3 = pos.copy(), vel.copy()
    gm3, gp3, gv3 = np.zeros(G, np.float32), np.zeros(3 * G, np.float32), np.zeros(3 * G, np.float32)
    grav = np.array([0.0, -9.8, 0.0], np.float32)
    L.orc_launch_fused(p3.reshape(-1), v3.reshape(-1), gm3, gp3, gv3, o, h, d, n, 1e-3, grav, 0.95, order)
    dpos, dvel = torch.as_tensor(pos).cuda(), torch.as_tensor(vel).cuda()
    dgm, dgp, dgv = (torch.full((G,), 7.0, device="cuda"), torch.full((3 * G,), 7.0, device="cuda"),
                     torch.full((3 * G,), 7.0, device="cuda"))
    g_c = (ctypes.c_float * 3)(*grav)
    _lib.check(Lg.pg_launch_fully_fused_g2p2g(P(dpos), P(dvel), P(dgm), P(dgp), P(dgv), o_c, h_c, d_c, n, 1e-3, g_c,
                                              0.95, order, stream))
    torch.cuda.synchronize()
    assert rel(dgm.cpu().numpy(), gm3) <= 1e-5
    assert rel(dvel.cpu().numpy(), v3) <= 1e-5
    assert rel(dpos.cpu().numpy(), p3) <= 1e-6


def _olib_g2p2g(O):
    L = olib(O)
    L.orc_launch_g2p2g.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p, ctypes.c_int,
                                   ctypes.c_float, _f32p, ctypes.c_int, ctypes.c_float, ctypes.c_int]
    return L


def _fixture_particles(f):
    n = len(f["m"])
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = f["x0"]
    pos[:, 3] = f["m"]
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = f["v0"]
    return pos, vel


@pytest.mark.parametrize("name", ["box512_o2", "box512_o3", "two_o2", "two_o3"])
def test_launcher_class_pic_matches_reference_solver(oracle_lib, name):
    """G2P2GKernelLauncher::launchG2P2G in pure-PIC mode is the reference CPU
    MPMSolver::step up to where gravity enters (particle velocity instead of
    grid momentum): the oracle restatement must reproduce the reference
    solver's fixtures (written by src/material_point_method.h itself) after
    1, 2 and 10 steps."""
    import os
    from conftest import GOLDEN
    L = _olib_g2p2g(oracle_lib)
    f = np.load(os.path.join(GOLDEN, f"refcompat_{name}.npz"))
    pos, vel = _fixture_particles(f)
    o, h, d = (f["origin"].astype(np.float32), f["cell"].astype(np.float32), f["dims"].astype(np.int32))
    grav = f["gravity"].astype(np.float32)
    G = int(np.prod(d))
    gm, gp, gv = np.zeros(G, np.float32), np.zeros(3 * G, np.float32), np.zeros(3 * G, np.float32)
    n = len(f["m"])

    def rel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))

    for s in range(1, 11):
        L.orc_launch_g2p2g(pos.reshape(-1), vel.reshape(-1), gm, gp, gv, o, h, d, n, float(f["dt"]), grav, 0,
                           0.95, int(f["order"]))
        if s in (1, 2, 10):
            assert rel(pos[:, :3], f[f"x_{s}"]) <= 1e-6, s
            assert rel(vel[:, :3], f[f"v_{s}"]) <= 1e-5, s
            assert rel(gm, f[f"grid_mass_{s}"]) <= 1e-6, s
            assert rel(gv.reshape(-1, 3), f[f"grid_vel_{s}"]) <= 1e-5, s


def test_launcher_class_blend(oracle_lib):
    """use_affine_pic blends the old velocity with the un-normalised gather: for
    one particle the gather returns S (v + g dt), S = sum of weights."""
    L = _olib_g2p2g(oracle_lib)
    rng = np.random.default_rng(5)
    pos, vel, o, h, d = scene(rng, 1)
    G = int(np.prod(d))
    gm, gp, gv = np.zeros(G, np.float32), np.zeros(3 * G, np.float32), np.zeros(3 * G, np.float32)
    g = np.array([0.0, -9.8, 0.0], np.float32)
    flip, dt = 0.8, 1e-3
    f = (pos[0, :3].astype(np.float64) - o) / h
    f -= np.floor(f)

    def w(x):
        a = abs(x)
        return 0.75 - x * x if a < 0.5 else (0.5 * (1.5 - a) ** 2 if a < 1.5 else 0.0)
    S = np.prod([sum(w(f[a] - k) for k in (-1, 0, 1)) for a in range(3)])
    v0 = vel[0, :3].astype(np.float64)
    x0 = pos[0, :3].astype(np.float64)
    L.orc_launch_g2p2g(pos.reshape(-1), vel.reshape(-1), gm, gp, gv, o, h, d, 1, dt, g, 1, flip, 2)
    want_v = flip * v0 + (1 - flip) * S * (v0 + g * dt)
    np.testing.assert_allclose(vel[0, :3], want_v, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(pos[0, :3], x0 + dt * want_v, rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("memory", [0, 1], ids=["host", "device"])
@pytest.mark.parametrize("affine", [0, 1], ids=["pic", "blend"])
def test_gpu_launcher_class_vs_oracle(gpu, oracle_lib, order, memory, affine):
    import torch
    from physgrad_amd import _lib
    Lg = _lib.load()
    L = _olib_g2p2g(oracle_lib)
    rng = np.random.default_rng(40 + order + 2 * memory + 4 * affine)
    n = 3000
    pos, vel, o, h, d = scene(rng, n)
    G = int(np.prod(d))
    grav = np.array([0.0, -9.8, 0.0], np.float32)
    print("order", order, "memory", memory, "affine", affine, flush=True)
    p1, v1 = pos.copy(), vel.copy()
    gm, gp, gv = np.zeros(G, np.float32), np.zeros(3 * G, np.float32), np.zeros(3 * G, np.float32)
    L.orc_launch_g2p2g(p1.reshape(-1), v1.reshape(-1), gm, gp, gv, o, h, d, n, 1e-3, grav, affine, 0.9, order)
    o_c = (ctypes.c_float * 3)(*o)
    h_c = (ctypes.c_float * 3)(*h)
    d_c = (ctypes.c_int * 3)(*d)
    g_c = (ctypes.c_float * 3)(*grav)
    if memory == 1:
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        dp, dv = torch.as_tensor(pos).cuda(), torch.as_tensor(vel).cuda()
        dgm, dgp, dgv = (torch.full((G,), 3.0, device="cuda"), torch.full((3 * G,), 3.0, device="cuda"),
                         torch.full((3 * G,), 3.0, device="cuda"))
        _lib.check(Lg.pg_launch_g2p2g(P(dp), P(dv), P(dgm), P(dgp), P(dgv), o_c, h_c, d_c, n, 1e-3, g_c, affine, 0.9,
                                      order, 1, stream))
        torch.cuda.synchronize()
        got = dp.cpu().numpy(), dv.cpu().numpy(), dgm.cpu().numpy(), dgv.cpu().numpy()
    else:
        hp, hv = pos.copy(), vel.copy()
        hgm, hgp, hgv = np.full(G, 3.0, np.float32), np.full(3 * G, 3.0, np.float32), np.full(3 * G, 3.0, np.float32)
        A = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        _lib.check(Lg.pg_launch_g2p2g(A(hp), A(hv), A(hgm), A(hgp), A(hgv), o_c, h_c, d_c, n, 1e-3, g_c, affine,
                                      0.9, order, 0, None))
        got = hp, hv, hgm, hgv

    def rel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))

    assert rel(got[0], p1) <= 1e-6
    assert rel(got[1], v1) <= 1e-5
    assert rel(got[2], gm) <= 1e-5
    assert rel(got[3], gv) <= 1e-5
