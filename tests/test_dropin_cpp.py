"""The C++ drop-in of the reference CPU solver (include/physgrad/material_point_method.h)
compiles against the C-ABI and, on the GPU, reproduces the reference fixtures."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "dropin_refcompat.cpp")


def build(tmp_path):
    exe = os.path.join(str(tmp_path), "dropin_refcompat")
    lib_dir = os.path.join(ROOT, "physgrad_amd")
    cmd = ["g++", "-std=c++17", "-O2", f"-I{os.path.join(ROOT, 'include')}", SRC, f"-L{lib_dir}",
           "-lphysgrad_mpm", f"-Wl,-rpath,{lib_dir}", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_dropin_compiles_and_links(tmp_path):
    exe = build(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_dropin_matches_reference_fixture(gpu, tmp_path):
    f = np.load(os.path.join(GOLDEN, "refcompat_box512_o2.npz"))
    exe = build(tmp_path)
    n = len(f["m"])
    steps = 10
    inp = os.path.join(str(tmp_path), "in.bin")
    out = os.path.join(str(tmp_path), "out.bin")
    with open(inp, "wb") as fh:
        np.array([*f["dims"], int(f["order"]), steps, n], np.int32).tofile(fh)
        np.array([*f["cell"], *f["origin"], f["dt"], *f["gravity"]], np.float32).tofile(fh)
        f["x0"].astype(np.float32).tofile(fh)
        f["v0"].astype(np.float32).tofile(fh)
        f["m"].astype(np.float32).tofile(fh)
    r = subprocess.run([exe, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    G = int(np.prod(f["dims"]))
    data = np.fromfile(out, np.float32)
    x, v = data[:3 * n], data[3 * n:6 * n]
    gm = data[6 * n:6 * n + G]
    gp = data[6 * n + G:6 * n + 4 * G]

    def rel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))

    assert rel(x, f["x_10"].reshape(-1)) <= 1e-5
    assert rel(v, f["v_10"].reshape(-1)) <= 1e-5
    assert rel(gm, f["grid_mass_10"]) <= 1e-5
    assert rel(gp, f["grid_mom_10"].reshape(-1)) <= 1e-5
    assert "steps 10" in r.stdout


# ---- GPU solver (mpm_solver.h), launcher (mpm_g2p2g_kernels.h) and
# differentiable system (differentiable_mpm.h) drop-ins

def build_driver(tmp_path, name):
    exe = os.path.join(str(tmp_path), name)
    lib_dir = os.path.join(ROOT, "physgrad_amd")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "tests", "cpp", name + ".cpp"), f"-L{lib_dir}", "-lphysgrad_mpm",
           f"-Wl,-rpath,{lib_dir}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("name", ["dropin_gpu_solver", "dropin_differentiable"])
def test_gpu_dropins_compile_and_link(tmp_path, name):
    assert os.path.exists(build_driver(tmp_path, name))


def test_solver_headers_are_exclusive(tmp_path):
    """The two solver drop-ins define the same types (as the reference's two
    headers do): including both must stop the build with a clear message."""
    src = os.path.join(str(tmp_path), "both.cpp")
    with open(src, "w") as fh:
        fh.write('#include "physgrad/material_point_method.h"\n#include "physgrad/mpm_solver.h"\nint main(){}\n')
    r = subprocess.run(["g++", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", "-c", src, "-o",
                        os.path.join(str(tmp_path), "both.o")], capture_output=True, text=True)
    assert r.returncode != 0 and "include one of them" in r.stderr


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.gpu
def test_gpu_solver_dropin_vs_oracle(gpu, oracle_lib, tmp_path):
    """MPMSolver<float> (jello: neo-Hookean E 1e4 nu 0.3 rho 1100) for 5 steps
    against the FP64 MLS-MPM oracle with the same model."""
    from oracle import oracle as O
    from physgrad_amd.scenes import Scene
    exe = build_driver(tmp_path, "dropin_gpu_solver")
    rng = np.random.default_rng(12)
    nx, n, steps, dt = 16, 512, 5, 1e-4
    x = (0.3 + 0.3 * rng.random((n, 3))).astype(np.float32)
    v = rng.normal(0, 0.3, (n, 3)).astype(np.float32)
    vol = 0.3 ** 3 / n
    m = np.full(n, 1100.0 * vol, np.float32)
    inp, out = os.path.join(str(tmp_path), "in.bin"), os.path.join(str(tmp_path), "out.bin")
    with open(inp, "wb") as fh:
        np.array([nx, n, steps], np.int32).tofile(fh)
        np.array([dt], np.float32).tofile(fh)
        x.tofile(fh)
        v.tofile(fh)
        m.tofile(fh)
    print("gpu solver drop-in", n, "particles", steps, "steps", flush=True)
    r = subprocess.run([exe, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "threw 1" in r.stdout and f"steps {steps}" in r.stdout, r.stdout
    d = np.fromfile(out, np.float32)
    gx, gv, gF, gs = d[:3 * n], d[3 * n:6 * n], d[6 * n:15 * n], d[15 * n:21 * n]
    gm = d[21 * n:]
    sc = Scene("dropin", (nx,) * 3, 1.0 / nx, dt, steps, 1, 1e4, 0.3, 1100.0, vol, x, v, gravity=(0.0, -9.81, 0.0))
    want = O.mls_run(O.scene_params(sc), x, v, None, None, m.astype(np.float64), 1e4, steps)
    # positions to FP32 resolution; the 5-step displacement (~1e-4) carries
    # the rounding of x itself (ulp 6e-8 at 0.5), hence the looser bar
    assert _rel(gx, want["x"].reshape(-1)) <= 1e-6
    assert _rel(gx - x.reshape(-1), want["x"].reshape(-1) - x.reshape(-1)) <= 5e-3
    assert _rel(gv, want["v"].reshape(-1)) <= 1e-4
    I = np.tile(np.eye(3).reshape(-1), n)
    assert _rel(gF - I, want["F"].reshape(-1) - I) <= 5e-3
    assert _rel(gF, want["F"].reshape(-1)) <= 1e-6
    # Cauchy stress of the returned F, neo-Hookean (mpm_kernels.cu:97-128)
    F = gF.reshape(n, 3, 3).astype(np.float64)
    mu, lam = 1e4 / 2.6, 1e4 * 0.3 / (1.3 * 0.4)
    J = np.linalg.det(F)
    B = F @ F.transpose(0, 2, 1)
    S = mu / J[:, None, None] * (B - np.eye(3)) + (lam * np.log(J) / J)[:, None, None] * np.eye(3)
    want_s = np.stack([S[:, 0, 0], S[:, 1, 1], S[:, 2, 2], S[:, 0, 1], S[:, 0, 2], S[:, 1, 2]], 1)
    assert _rel(gs, want_s.reshape(-1)) <= 1e-4
    assert abs(gm.sum() / m.sum() - 1) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("order,affine", [(2, 1), (3, 0)])
def test_launcher_dropin_vs_oracle(gpu, oracle_lib, tmp_path, order, affine):
    from test_launchers import _olib_g2p2g, scene
    L = _olib_g2p2g(oracle_lib)
    exe = build_driver(tmp_path, "dropin_differentiable")
    rng = np.random.default_rng(30 + order)
    n, steps = 2000, 3
    pos, vel, o, h, d = scene(rng, n)
    grav = np.array([0.0, -9.8, 0.0], np.float32)
    dt, flip = 1e-3, 0.9
    inp, out = os.path.join(str(tmp_path), "in.bin"), os.path.join(str(tmp_path), "out.bin")
    with open(inp, "wb") as fh:
        np.array([*d, n, steps, order, affine], np.int32).tofile(fh)
        np.array([*o, *h, dt, *grav, flip], np.float32).tofile(fh)
        pos.tofile(fh)
        vel[:, :3].copy().tofile(fh)
    print("launcher drop-in order", order, "affine", affine, flush=True)
    r = subprocess.run([exe, "launcher", inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    G = int(np.prod(d))
    gm, gp, gv = np.zeros(G, np.float32), np.zeros(3 * G, np.float32), np.zeros(3 * G, np.float32)
    p1, v1 = pos.copy(), vel.copy()
    for _ in range(steps):
        L.orc_launch_g2p2g(p1.reshape(-1), v1.reshape(-1), gm, gp, gv, o, h, d, n, dt, grav, affine, flip, order)
    res = np.fromfile(out, np.float32)
    assert _rel(res[:3 * n], p1[:, :3].reshape(-1)) <= 1e-6
    assert _rel(res[3 * n:6 * n], v1[:, :3].reshape(-1)) <= 1e-5
    assert _rel(res[6 * n:6 * n + G], gm) <= 1e-5
    assert _rel(res[6 * n + G:], gv) <= 1e-5


@pytest.mark.gpu
def test_differentiable_dropin_vs_oracle_reverse(gpu, oracle_lib, tmp_path):
    """DifferentiableMPMSystem: 7 recorded steps (checkpoint every 3), seed on
    the final x and v, gradients vs the FP64 oracle reverse at 1e-3; then the
    drop-in's verifyGradients (central differences of the same rollout)."""
    from physgrad_amd.scenes import Scene
    from test_gpu_adjoint import oracle_reverse
    exe = build_driver(tmp_path, "dropin_differentiable")
    rng = np.random.default_rng(21)
    nx, n, steps, K, material = 16, 256, 7, 3, 0
    dt, E0, nu, span = 2e-4, 2e4, 0.3, 0.35
    vol = span ** 3 / n
    x = (0.3 + span * rng.random((n, 3))).astype(np.float32)
    v = rng.normal(0, 0.3, (n, 3)).astype(np.float32)
    m = np.full(n, 1000.0 * vol, np.float32)
    Ep = np.exp(rng.uniform(np.log(5e3), np.log(5e4), n)).astype(np.float32)
    sx = rng.normal(size=(n, 3)).astype(np.float32)
    sv = rng.normal(size=(n, 3)).astype(np.float32)
    g = (0.0, -9.8, 0.0)
    inp, out = os.path.join(str(tmp_path), "in.bin"), os.path.join(str(tmp_path), "out.bin")
    with open(inp, "wb") as fh:
        np.array([nx, n, steps, K, material], np.int32).tofile(fh)
        np.array([dt, E0, nu, vol, *g], np.float32).tofile(fh)
        for a in (x, v, m, Ep, sx, sv):
            a.tofile(fh)
    print("differentiable drop-in", n, "particles", steps, "steps", flush=True)
    r = subprocess.run([exe, "diff", inp, out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    res = np.fromfile(out, np.float32)
    got = {"x": res[:3 * n], "v": res[3 * n:6 * n], "F": res[6 * n:15 * n], "C": res[15 * n:24 * n],
           "E": res[24 * n:]}
    sc = Scene("dropin", (nx,) * 3, 1.0 / nx, dt, steps, material, E0, nu, 1000.0, vol, x, v, gravity=g,
               E_particles=Ep)
    seeds = [sx.astype(np.float64), sv.astype(np.float64), np.zeros((n, 9)), np.zeros((n, 9))]
    (wx, wv, wC, wF), wE = oracle_reverse(oracle_lib, sc, x.astype(np.float64), v.astype(np.float64),
                                          np.zeros((n, 9)), np.tile(np.eye(3).reshape(1, 9), (n, 1)),
                                          Ep.astype(np.float64), steps, seeds)
    for key, w in (("x", wx), ("v", wv), ("C", wC), ("F", wF), ("E", wE)):
        err = _rel(got[key], np.asarray(w).reshape(-1))
        print(key, err)
        assert err <= 1e-3, (key, err)
    assert "verifyGradients passed" in r.stdout, r.stdout
