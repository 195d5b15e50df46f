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
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{os.path.join(ROOT, 'include')}", os.path.join(ROOT, "tests", "cpp", name + ".cpp"), f"-L{lib_dir}",
           "-lphysgrad_mpm", "-L/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{lib_dir}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("name", ["dropin_gpu_solver", "dropin_differentiable", "dropin_sparse"])
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


# MaterialDatabase defaults of the drop-in (the reference's, mpm_data_structures.h:
# 548-611): type -> (model, E, nu, density, viscosity) of the first entry of that type
DB = {0: (1, 1e4, 0.3, 1100.0, 0.0),          # ELASTIC -> jello, neo-Hookean
      2: (2, 0.0, 0.0, 1000.0, 0.001),        # FLUID -> water, Newtonian
      4: (1, 1.4e5, 0.2, 400.0, 0.0)}         # SNOW -> snow, neo-Hookean (no plasticity)
BC_DY, BC_STICKY = 2, 64


def _drive_gpu_solver(exe, tmp_path, nx, x, v, m, types, steps, dt, use_run=0, jplane=-1, outdir=None):
    n = len(m)
    inp, out = os.path.join(str(tmp_path), "in.bin"), os.path.join(str(tmp_path), "out.bin")
    with open(inp, "wb") as fh:
        np.array([nx, n, steps, use_run, jplane], np.int32).tofile(fh)
        np.array([dt], np.float32).tofile(fh)
        x.astype(np.float32).tofile(fh)
        v.astype(np.float32).tofile(fh)
        m.astype(np.float32).tofile(fh)
        types.astype(np.int32).tofile(fh)
    r = subprocess.run([exe, inp, out] + ([outdir] if outdir else []), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = np.fromfile(out, np.float32)
    G = nx ** 3
    res = dict(x=d[:3 * n].reshape(n, 3), v=d[3 * n:6 * n].reshape(n, 3), F=d[6 * n:15 * n].reshape(n, 3, 3),
               s=d[15 * n:21 * n].reshape(n, 6), gm=d[21 * n:21 * n + G],
               flags=np.fromfile(out, np.uint32)[21 * n + G:])
    return r.stdout, res


def _solver_scene(rng, n, types_pool):
    x = (0.3 + 0.3 * rng.random((n, 3))).astype(np.float32)
    v = rng.normal(0, 0.3, (n, 3)).astype(np.float32)
    types = rng.choice(types_pool, n).astype(np.int32)
    vol = 0.3 ** 3 / n
    rho = np.array([DB[t][3] for t in types])
    m = (rho * vol).astype(np.float32)
    return x, v, m, types


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["elastic", "mixed"])
def test_gpu_solver_dropin_vs_oracle(gpu, oracle_lib, tmp_path, kind):
    """MPMSolver<float> for 5 steps against the FP64 MLS-MPM oracle with the
    same per-particle materials and node flags.  elastic: all jello
    (neo-Hookean E 1e4 nu 0.3 rho 1100); mixed: jello, water (Newtonian
    fluid) and snow particles, plus a DIRICHLET_Y node plane at y = nx/2 set
    through setBoundaryCondition."""
    O = oracle_lib
    exe = build_driver(tmp_path, "dropin_gpu_solver")
    rng = np.random.default_rng(12 if kind == "elastic" else 13)
    nx, n, steps, dt = 16, 512, 5, 1e-4
    x, v, m, types = _solver_scene(rng, n, [0] if kind == "elastic" else [0, 2, 4])
    jplane = -1 if kind == "elastic" else nx // 2
    print("gpu solver drop-in", kind, n, "particles", steps, "steps", flush=True)
    stdout, res = _drive_gpu_solver(exe, tmp_path, nx, x, v, m, types, steps, dt, jplane=jplane)
    assert "threw 1" in stdout and f"steps {steps}" in stdout, stdout
    # enable_profiling: per-kernel event times reach getMetrics()
    assert float(stdout.split("avg_step_ms")[1].split()[0]) > 0, stdout
    # the driver adds particles grouped by type
    order = np.argsort(types, kind="stable")
    x, v, m, types = x[order], v[order], m[order], types[order]
    # node flags: sticky band of 3, the reference's y = 0 floor, the caller's plane
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(nx), np.arange(nx), indexing="ij")
    ii, jj, kk = (a.transpose(2, 1, 0).reshape(-1) for a in (ii, jj, kk))   # node (k ny + j) nx + i
    band = (np.minimum(np.minimum(ii, jj), kk) < 3) | (np.maximum(np.maximum(ii, jj), kk) >= nx - 3)
    want_flags = np.where(band, BC_STICKY, 0) | np.where(jj == 0, BC_DY, 0) | np.where(jj == jplane, BC_DY, 0)
    np.testing.assert_array_equal(res["flags"], want_flags.astype(np.uint32))
    prm = O.mls_params((nx,) * 3, 1.0 / nx, (0, 0, 0), dt, (0.0, -9.81, 0.0), 3, 1, 0.3, 0.0)
    model = np.array([DB[t][0] for t in types])
    E = np.array([DB[t][1] for t in types])
    nu = np.array([DB[t][2] for t in types])
    visc = np.array([DB[t][4] for t in types])
    vol = m.astype(np.float64) / np.array([DB[t][3] for t in types])
    want = O.mls_run_ext(prm, x, v, None, None, m.astype(np.float64), E, steps, model=model, nu=nu, visc=visc,
                         vol=vol, flags=res["flags"])
    gx, gv, gF = res["x"], res["v"], res["F"]
    # positions to FP32 resolution; the 5-step displacement (~1e-4) carries
    # the rounding of x itself (ulp 6e-8 at 0.5), hence the looser bar
    assert _rel(gx, want["x"]) <= 1e-6
    assert _rel(gx - x, want["x"] - x) <= 5e-3
    assert _rel(gv, want["v"]) <= 1e-4
    I = np.eye(3)
    assert _rel(gF - I, want["F"] - I) <= 5e-3
    assert _rel(gF, want["F"]) <= 1e-6
    # Cauchy stress: neo-Hookean of the returned F (mpm_kernels.cu:97-128);
    # water 2 visc sym(C) of the oracle's C (:131-149)
    F = gF.astype(np.float64)
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    J = np.linalg.det(F)
    B = F @ F.transpose(0, 2, 1)
    S = (mu / J)[:, None, None] * (B - I) + (lam * np.log(J) / J)[:, None, None] * I
    Cw = want["C"]
    Sf = visc[:, None, None] * (Cw + Cw.transpose(0, 2, 1))
    fluid = model == 2
    S[fluid] = Sf[fluid]
    want_s = np.stack([S[:, 0, 0], S[:, 1, 1], S[:, 2, 2], S[:, 0, 1], S[:, 0, 2], S[:, 1, 2]], 1)
    assert _rel(res["s"][~fluid], want_s[~fluid]) <= 1e-4
    if fluid.any():
        assert _rel(res["s"][fluid], want_s[fluid]) <= 1e-4
    assert abs(res["gm"].sum() / m.sum() - 1) <= 1e-5


@pytest.mark.gpu
def test_gpu_solver_dropin_run_writes_vtk(gpu, tmp_path):
    """run() to simulation_time = 6 dt with output_frequency 2 dt writes one
    legacy VTK file per output time into output_directory, read back with
    the package's VTK reader: n points, velocity, mass and J fields; the last
    file holds the final particles."""
    from physgrad_amd.state_io import read_vtk_points
    exe = build_driver(tmp_path, "dropin_gpu_solver")
    rng = np.random.default_rng(14)
    nx, n, steps, dt = 16, 300, 6, 1e-4
    x, v, m, types = _solver_scene(rng, n, [0, 2])
    outdir = os.path.join(str(tmp_path), "vtk")
    stdout, res = _drive_gpu_solver(exe, tmp_path, nx, x, v, m, types, steps, dt, use_run=1, outdir=outdir)
    # the solver's clock and output test, in float32 as the header computes them
    t, k, frames = np.float32(0), 0, []
    dt32, freq, end = np.float32(dt), np.float32(2) * np.float32(dt), np.float32(steps) * np.float32(dt)
    while t < end:
        t = np.float32(t + dt32)
        k += 1
        if int(np.float32(t / freq)) > int(np.float32(np.float32(t - dt32) / freq)):
            frames.append(k)
    assert f"steps {k} " in stdout, stdout
    assert sorted(os.listdir(outdir)) == [f"particles_{f:06d}.vtk" for f in frames]
    for f in frames:
        px, fields = read_vtk_points(os.path.join(outdir, f"particles_{f:06d}.vtk"))
        assert px.shape == (n, 3) and np.isfinite(px).all()
        assert set(fields) == {"velocity", "mass", "J"}
        np.testing.assert_array_equal(np.sort(fields["mass"]), np.sort(m))
    if frames[-1] == k:
        px, fields = read_vtk_points(os.path.join(outdir, f"particles_{k:06d}.vtk"))
        np.testing.assert_array_equal(px, res["x"])
        np.testing.assert_array_equal(fields["velocity"], res["v"])
        np.testing.assert_allclose(fields["J"], np.linalg.det(res["F"].astype(np.float64)), rtol=1e-5)


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


@pytest.mark.gpu
@pytest.mark.parametrize("cell,dmax,radius,n,nq", [(0.1, (1.0, 1.0, 1.0), 0.1, 5000, 300),
                                                   (0.13, (1.0, 0.7, 0.9), 0.3, 3000, 100)])
def test_sparse_dropin_hash_vs_oracle(gpu, oracle_lib, tmp_path, cell, dmax, radius, n, nq):
    """SpatialHashGrid<float> of the drop-in: buildHashTable (cell ids, order,
    starts, counts) and getNeighbors / getNeighborsBatch against the oracle,
    which is bit-identical to the reference's class."""
    O = oracle_lib
    exe = build_driver(tmp_path, "dropin_sparse")
    rng = np.random.default_rng(n)
    dmax = np.array(dmax, np.float32)
    pos = (rng.random((n, 3)) * dmax).astype(np.float32)
    q = (rng.random((nq, 3)) * dmax).astype(np.float32)
    inp, out = os.path.join(str(tmp_path), "in.bin"), os.path.join(str(tmp_path), "out.bin")
    with open(inp, "wb") as fh:
        np.array([cell, 0, 0, 0, *dmax, radius], np.float32).tofile(fh)
        np.array([n, nq], np.uint32).tofile(fh)
        pos.tofile(fh)
        q.tofile(fh)
    print("sparse drop-in hash", n, nq, flush=True)
    r = subprocess.run([exe, "hash", inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "threw 1" in r.stdout, r.stdout
    res = np.ceil(dmax / np.float32(cell)).astype(np.uint32)
    C = int(np.prod(res.astype(np.uint64)))
    ids = O.cell_ids(pos, cell, (0, 0, 0), res)
    idx, st, ct = O.counting_sort(ids, C)
    orp, onb = O.hash_neighbors(q, radius, cell, (0, 0, 0), res, idx, st, ct)
    d = np.fromfile(out, np.uint32)
    parts = np.split(d, np.cumsum([n, n, C, C, nq + 1, len(onb)]))
    for got, want in zip(parts[:6], (ids, idx, st, ct, orp, onb)):
        np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(parts[6], onb[orp[0]:orp[1]])


@pytest.mark.gpu
def test_sparse_dropin_matrix_vs_oracle(gpu, oracle_lib, tmp_path):
    """SparseMatrix<float> of the drop-in: addElement in row-major order,
    finalize, multiply on the GPU against the oracle's row-ordered sums, bit for
    bit (the reference's own class loses the rows and returns zeros)."""
    O = oracle_lib
    exe = build_driver(tmp_path, "dropin_sparse")
    rng = np.random.default_rng(70)
    nr, nc, nnz = 3000, 2500, 40000
    rows = np.sort(rng.integers(0, nr, nnz)).astype(np.uint32)
    rows[rows == 7] = 8            # an empty row
    rows = np.sort(rows)
    cols = rng.integers(0, nc, nnz).astype(np.uint32)
    vals = rng.normal(size=nnz).astype(np.float32)
    x = rng.normal(size=nc).astype(np.float32)
    inp, out = os.path.join(str(tmp_path), "in.bin"), os.path.join(str(tmp_path), "out.bin")
    with open(inp, "wb") as fh:
        np.array([nr, nc, nnz], np.uint32).tofile(fh)
        rows.tofile(fh)
        cols.tofile(fh)
        vals.tofile(fh)
        x.tofile(fh)
    print("sparse drop-in matrix", nr, nc, nnz, flush=True)
    r = subprocess.run([exe, "matrix", inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = np.fromfile(out, np.uint32)
    rp = d[:nr + 1]
    y = d[nr + 1:].view(np.float32)
    want_rp = np.zeros(nr + 1, np.uint32)
    want_rp[1:] = np.cumsum(np.bincount(rows, minlength=nr))
    np.testing.assert_array_equal(rp, want_rp)
    assert rp[8] == rp[7]
    want = O.csr_spmv(want_rp, cols, vals, x)
    np.testing.assert_array_equal(y.view(np.uint32), want.view(np.uint32))


# ---- checkpoint spacing of the differentiable drop-in (host only)

def test_checkpoint_interval_strategies(tmp_path):
    """checkpointInterval: NONE / no checkpointing keep one segment, UNIFORM takes
    checkpoint_every, LOGARITHMIC ceil(sqrt(steps)), ADAPTIVE the longest segment
    whose checkpoints + one segment's tape fit the budget (checked against a brute
    force over every spacing), the least memory when nothing fits."""
    import math
    exe = os.path.join(str(tmp_path), "checkpoint_interval")
    src = os.path.join(ROOT, "tests", "cpp", "checkpoint_interval.cpp")
    lib_dir = os.path.join(ROOT, "physgrad_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", f"-I{os.path.join(ROOT, 'include')}", src, f"-L{lib_dir}",
                    "-lphysgrad_mpm", f"-Wl,-rpath,{lib_dir}", "-o", exe], check=True, capture_output=True, text=True)
    NONE, UNIFORM, ADAPTIVE, LOG = 0, 1, 2, 3
    ck, st = 108.0 * 2**20, 116.0 * 2**20 + 32.0 * 128**3   # 1M particles on 128^3

    def mem(steps, K):
        return math.ceil(steps / K) * ck + (K + 1) * st

    def adaptive(steps, budget, maxc):
        fits = [K for K in range(steps, 0, -1) if mem(steps, K) <= budget * 1e9]
        capped = [K for K in fits if math.ceil(steps / K) <= maxc]
        if capped:
            return capped[0]
        if fits:
            return fits[0]
        return min(range(1, steps + 1), key=lambda K: (mem(steps, K), -K))

    cases = [(NONE, 256, 4.0, 100, 1, 32), (UNIFORM, 256, 4.0, 100, 0, 32), (UNIFORM, 256, 4.0, 100, 1, 32),
             (UNIFORM, 10, 4.0, 100, 1, 32), (LOG, 256, 4.0, 100, 1, 32), (LOG, 1000, 4.0, 100, 1, 32),
             (ADAPTIVE, 256, 4.0, 100, 1, 32), (ADAPTIVE, 256, 60.0, 100, 1, 32), (ADAPTIVE, 512, 2.0, 100, 1, 32),
             (ADAPTIVE, 512, 2.0, 4, 1, 32), (ADAPTIVE, 512, 0.01, 100, 1, 32), (ADAPTIVE, 1, 4.0, 100, 1, 32), (ADAPTIVE, 256, 10.0, 100, 1, 32)]
    args = []
    for c in cases:
        args += [str(a) for a in c] + [repr(ck), repr(st)]
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = [int(t) for t in r.stdout.split()]
    want = [256, 256, 32, 10, 16, 32, adaptive(256, 4.0, 100), adaptive(256, 60.0, 100), adaptive(512, 2.0, 100),
            adaptive(512, 2.0, 4), adaptive(512, 0.01, 100), 1, adaptive(256, 10.0, 100)]
    assert got == want
    assert want[7] == 256                         # a generous budget keeps one segment: no recompute
    assert mem(256, want[6]) == min(mem(256, K) for K in range(1, 257)) > 4e9   # nothing fits 4 GB
    assert 1 < want[12] < 256 and mem(256, want[12]) <= 10e9 < mem(256, want[12] + 1)
