"""GPU: per-particle materials (fixed-corotated, neo-Hookean, Newtonian fluid;
own Poisson ratio, viscosity and density) and per-node boundary flags
(DIRICHLET_X/Y/Z, STICKY, SLIP, NEUMANN), tiled and plain paths, forward
against the FP64 oracle with the same materials and flags
(oracle/mlsmpm_ext.c), and the adjoint against FP64 central finite
differences of that oracle."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TABLE = [
    {"model": 0, "youngs_modulus": 2e4, "poisson_ratio": 0.25, "density": 900.0, "viscosity": 0.0},
    {"model": 1, "youngs_modulus": 1e4, "poisson_ratio": 0.35, "density": 1100.0, "viscosity": 0.0},
    {"model": 2, "youngs_modulus": 0.0, "poisson_ratio": 0.0, "density": 1000.0, "viscosity": 0.8},
    {"model": 1, "youngs_modulus": 3e4, "poisson_ratio": 0.2, "density": 0.0, "viscosity": 0.0},
]
BC = {"dx": 1, "dy": 2, "dz": 4, "nx": 8, "ny": 16, "nz": 32, "sticky": 64, "slip": 128}


def scene(rng, n, side=16, per_particle_E=True):
    x = (0.25 + 0.5 * rng.random((n, 3))).astype(np.float32)
    v = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    C = rng.normal(0, 1.0, (n, 3, 3)).astype(np.float32)
    F = (np.eye(3) + 0.05 * rng.normal(size=(n, 3, 3))).astype(np.float32)
    m = (2e-4 * (0.5 + rng.random(n))).astype(np.float32)
    ids = rng.integers(0, len(TABLE), n).astype(np.uint8)
    E = (1e4 * (0.5 + rng.random(n))).astype(np.float32) if per_particle_E else None
    G = side ** 3
    flags = rng.choice([0, 0, 0, BC["dx"], BC["dy"] | BC["nz"], BC["dz"], BC["sticky"], BC["slip"],
                        BC["slip"] | BC["dy"]], G).astype(np.uint32)
    return x, v, C, F, m, ids, E, flags


def oracle_inputs(ids, m, E, p_vol):
    model = np.array([TABLE[i]["model"] for i in ids])
    nu = np.array([TABLE[i]["poisson_ratio"] for i in ids], np.float64)
    visc = np.array([TABLE[i]["viscosity"] for i in ids], np.float64)
    rho = np.array([TABLE[i]["density"] for i in ids], np.float64)
    vol = np.where(rho > 0, m.astype(np.float64) / np.where(rho > 0, rho, 1.0), p_vol)
    Ep = E.astype(np.float64) if E is not None else np.array([TABLE[i]["youngs_modulus"] for i in ids], np.float64)
    return dict(model=model, nu=nu, visc=visc, vol=vol), Ep


def params(impl, side=16, p_vol=1e-6, g=(0.0, -9.8, 0.0)):
    from physgrad_amd import _lib
    from physgrad_amd.solver import make_params
    return make_params(mode=_lib.MODE_MLS_MPM, impl=impl, dims=(side,) * 3, dx=1.0 / side, dt=1e-4, gravity=g,
                       material=0, youngs_modulus=1e4, poisson_ratio=0.3, density=1000.0, particle_volume=p_vol,
                       boundary_thickness=2, max_particles=4096)


def rel(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("impl", [0, 1], ids=["plain", "tiled"])
@pytest.mark.parametrize("perE", [True, False], ids=["E_per_particle", "E_from_table"])
def test_forward_mixed_materials_and_flags_vs_fp64(gpu, oracle_lib, impl, perE):
    from physgrad_amd.solver import MPMContext
    O = oracle_lib
    rng = np.random.default_rng(90 + impl + 2 * perE)
    n = 1500
    x, v, C, F, m, ids, E, flags = scene(rng, n, per_particle_E=perE)
    prm = params(impl)
    ctx = MPMContext(prm)
    ctx.set_particles(x, v, C, F, mass=m, E=E)
    ctx.set_materials(TABLE, ids)
    ctx.set_boundary(flags)
    extra, Ep = oracle_inputs(ids, m, E, prm.particle_volume)
    oprm = O.mls_params((16,) * 3, 1 / 16, (0, 0, 0), 1e-4, (0, -9.8, 0), 2, 0, 0.3, prm.particle_volume)
    xs, vs, Cs, Fs = (a.astype(np.float64) for a in (x, v, C, F))
    for k in range(3):
        print("impl", impl, "step", k, flush=True)
        ctx.step(1)
        want = O.mls_run_ext(oprm, xs, vs, Cs, Fs, m.astype(np.float64), Ep, 1, flags=flags, want_grid=True,
                             **extra)
        got = {kk: t.cpu().numpy() for kk, t in ctx.get_particles().items()}
        for key in ("x", "v", "C", "F"):
            err = rel(got[key], want[key])
            assert err <= (1e-5 if k == 0 else 1e-4), (key, k, err)
        xs, vs, Cs, Fs = want["x"], want["v"], want["C"], want["F"]
    ctx.close()


def test_batch_flags_index_the_concatenated_grids(gpu, oracle_lib):
    """Two scenes of different size in a batch: each scene's flags and
    materials apply to its own grid and particles."""
    from physgrad_amd.solver import MPMBatch, scene_desc
    O = oracle_lib
    rng = np.random.default_rng(95)
    sides = (16, 24)
    parts, flags_all, oprms, descs = [], [], [], []
    for s, side in enumerate(sides):
        n = 900 + 300 * s
        x, v, C, F, m, ids, E, flags = scene(rng, n, side=side)
        parts.append((x, v, C, F, m, ids, E, flags))
        flags_all.append(flags)
        descs.append(scene_desc((side,) * 3, 1.0 / side, dt=1e-4, gravity=(0, -9.8, 0), material=0,
                                youngs_modulus=1e4, poisson_ratio=0.3, density=1000.0, particle_volume=1e-6,
                                boundary_thickness=2, max_particles=n))
        oprms.append(O.mls_params((side,) * 3, 1.0 / side, (0, 0, 0), 1e-4, (0, -9.8, 0), 2, 0, 0.3, 1e-6))
    b = MPMBatch(descs)
    cat = lambda k: np.concatenate([p[k] for p in parts])  # noqa: E731
    b.set_particles([len(p[0]) for p in parts], cat(0), cat(1), cat(2), cat(3), mass=cat(4), E=cat(6))
    b.set_materials(TABLE, cat(5))
    b.set_boundary(np.concatenate(flags_all))
    b.step(1)
    got = {k: t.cpu().numpy() for k, t in b.get_particles().items()}
    off = 0
    for p, oprm in zip(parts, oprms):
        x, v, C, F, m, ids, E, flags = p
        n = len(x)
        extra, Ep = oracle_inputs(ids, m, E, 1e-6)
        want = O.mls_run_ext(oprm, x.astype(np.float64), v.astype(np.float64), C.astype(np.float64),
                             F.astype(np.float64), m.astype(np.float64), Ep, 1, flags=flags, **extra)
        for key in ("x", "v", "C", "F"):
            assert rel(got[key][off:off + n], want[key]) <= 1e-5, key
        off += n
    b.close()


@pytest.mark.parametrize("impl", [0, 1], ids=["plain", "tiled"])
def test_adjoint_mixed_materials_and_flags_vs_fp64_fd(gpu, oracle_lib, impl):
    """dL/d(x0, v0, C0, F0, E) of L = sum w . x_T + sum u . v_T after 3 steps
    (checkpoint every 2) against FP64 central differences of the oracle, on
    sampled entries of fluid and elastic particles."""
    from physgrad_amd.solver import MPMContext
    O = oracle_lib
    rng = np.random.default_rng(97 + impl)
    n, T = 96, 3
    x, v, C, F, m, ids, E, _ = scene(rng, n)
    # flags away from the particles' reach plus a Dirichlet/slip band under them, with
    # velocities far from zero so that no SLIP sign flips under the differences
    G = 16 ** 3
    flags = np.zeros(G, np.uint32)
    kk, jj, ii = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    flags[(jj.reshape(-1) == 5)] = BC["dy"]
    flags[(ii.reshape(-1) == 11)] = BC["slip"] | BC["dz"]
    v[:, 0] += 1.0
    prm = params(impl)
    w = rng.normal(size=(n, 3))
    u = rng.normal(size=(n, 3))
    extra, Ep = oracle_inputs(ids, m, E, prm.particle_volume)
    oprm = O.mls_params((16,) * 3, 1 / 16, (0, 0, 0), 1e-4, (0, -9.8, 0), 2, 0, 0.3, prm.particle_volume)

    def loss(x_, v_, C_, F_, E_):
        o = O.mls_run_ext(oprm, x_, v_, C_, F_, m.astype(np.float64), E_, T, flags=flags, **extra)
        return float((w * o["x"]).sum() + (u * o["v"]).sum())

    ctx = MPMContext(prm)
    ctx.set_particles(x, v, C, F, mass=m, E=E)
    ctx.set_materials(TABLE, ids)
    ctx.set_boundary(flags)
    ctx.forward_checkpointed(T, 2)
    g = ctx.backward(ax=torch.as_tensor(w, dtype=torch.float32), av=torch.as_tensor(u, dtype=torch.float32))
    g = {k: t.cpu().numpy().astype(np.float64) for k, t in g.items()}
    base = {"x": x.astype(np.float64), "v": v.astype(np.float64), "C": C.astype(np.float64),
            "F": F.astype(np.float64), "E": Ep.copy()}
    fluid = np.nonzero(np.array([TABLE[i]["model"] for i in ids]) == 2)[0]
    elastic = np.nonzero(np.array([TABLE[i]["model"] for i in ids]) != 2)[0]
    for name in ("x", "v", "C", "F", "E"):
        fd, gg = [], []
        pool = elastic if name == "E" else np.concatenate([fluid[:3], elastic[:3]])
        for q in pool[:6]:
            arr = base[name]
            idx = (q,) + tuple(int(rng.integers(0, s)) for s in arr.shape[1:])
            h = 1e-6 * max(1.0, abs(arr[idx]))
            ap, am = arr.copy(), arr.copy()
            ap[idx] += h
            am[idx] -= h
            kp = {k: (ap if k == name else a) for k, a in base.items()}
            km = {k: (am if k == name else a) for k, a in base.items()}
            fd.append((loss(kp["x"], kp["v"], kp["C"], kp["F"], kp["E"]) -
                       loss(km["x"], km["v"], km["C"], km["F"], km["E"])) / (2 * h))
            gg.append(g[name].reshape(arr.shape)[idx])
        fd, gg = np.array(fd), np.array(gg)
        print(name, "fd", fd, "adjoint", gg, flush=True)
        assert np.abs(gg - fd).max() <= 1e-3 * np.abs(fd).max(), name
    ctx.close()
