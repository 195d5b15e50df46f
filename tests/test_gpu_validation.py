"""GPU: state validation (non-finite particles, particles outside the grid,
the binning error word) and launch sizes that are not a multiple of the XCD
count (the persistent tile kernels must still take every work item)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def scene(rng, n=2000, material=1):
    from physgrad_amd.scenes import Scene
    x = (0.3 + 0.4 * rng.random((n, 3))).astype(np.float32)
    v = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    return Scene("val", (32, 32, 32), 1 / 32, 1e-4, 4, material, 2e4, 0.3, 1000.0, 0.4 ** 3 / n, x, v)


@pytest.mark.parametrize("impl", [0, 1], ids=["plain", "tiled"])
def test_nonfinite_particle_is_named(gpu, impl):
    from physgrad_amd._lib import PG_ERR_STATE, PgError
    from physgrad_amd.solver import MPMContext
    rng = np.random.default_rng(5)
    sc = scene(rng)
    ctx = MPMContext(sc.params(impl=impl))
    ctx.set_particles(sc.x, sc.v)
    ctx.step(3)
    assert ctx.check_state() == (-1, -1)
    x = sc.x.copy()
    x[37, 1] = np.nan
    v = sc.v.copy()
    v[1200, 2] = np.inf
    ctx.set_particles(x, v)
    with pytest.raises(PgError) as ei:
        ctx.check_state()
    assert ei.value.code == PG_ERR_STATE
    assert "2 particle(s)" in str(ei.value) and "particle 37" in str(ei.value)
    # a clean state again passes, and validation on every step catches the NaN
    ctx.set_particles(sc.x, sc.v)
    ctx.set_validation(True)
    ctx.step(2)
    ctx.set_particles(x, sc.v)
    with pytest.raises(PgError) as ei:
        ctx.step(1)
    assert ei.value.code == PG_ERR_STATE
    ctx.close()


def test_nonfinite_named_by_user_index_after_tiled_steps(gpu):
    """An isolated particle (more than 4 cells from any other) with a NaN in C
    poisons only its own nodes; after two tiled steps the storage order is the
    tile order, and the reported index is still the caller's."""
    from physgrad_amd._lib import PgError
    from physgrad_amd.solver import MPMContext
    rng = np.random.default_rng(6)
    sc = scene(rng)
    x, v = sc.x.copy(), sc.v.copy()
    x[1500] = (0.86, 0.86, 0.86)
    v[1500] = 0.0
    C = np.zeros((sc.n, 3, 3), np.float32)
    C[1500, 0, 1] = np.nan
    ctx = MPMContext(sc.params(impl=1))
    ctx.set_particles(x, v, C=C)
    ctx.step(2)
    with pytest.raises(PgError) as ei:
        ctx.check_state()
    assert "1 particle(s)" in str(ei.value) and "particle 1500" in str(ei.value)
    ctx.close()


def test_particle_outside_grid_is_reported_not_an_error(gpu):
    from physgrad_amd.solver import MPMContext
    rng = np.random.default_rng(7)
    sc = scene(rng)
    x = sc.x.copy()
    x[11] = (0.5, 1.2, 0.5)
    x[900] = (-0.1, 0.5, 0.5)
    for impl in (0, 1):
        ctx = MPMContext(sc.params(impl=impl))
        ctx.set_particles(x, sc.v)
        assert ctx.check_state() == (-1, 11)
        ctx.close()


def test_refcompat_validation(gpu):
    from physgrad_amd import _lib
    from physgrad_amd._lib import PgError
    from physgrad_amd.solver import MPMContext, make_params
    rng = np.random.default_rng(8)
    n = 500
    x = (0.3 + 0.4 * rng.random((n, 3))).astype(np.float32)
    v = np.zeros((n, 3), np.float32)
    ctx = MPMContext(make_params(mode=_lib.MODE_REF_COMPAT, dims=(32, 32, 32), dx=1 / 32, max_particles=n))
    ctx.set_particles(x, v)
    ctx.step(2)
    assert ctx.check_state() == (-1, -1)
    v[250, 0] = np.nan
    ctx.set_particles(x, v)
    with pytest.raises(PgError) as ei:
        ctx.check_state()
    assert "particle 250" in str(ei.value)
    ctx.close()


def run(sc, launch, steps, backward=False):
    from physgrad_amd.solver import MPMContext
    ctx = MPMContext(sc.params(impl=1))
    ctx.set_particles(sc.x, sc.v, E=sc.E_particles)
    if launch:
        ctx.launch = launch
    if backward:
        ctx.forward_checkpointed(steps, 2)
        # random weights on the final positions and velocities: every adjoint is O(1)
        rng = np.random.default_rng(19)
        ax = torch.as_tensor(rng.normal(size=(sc.n, 3)).astype(np.float32), device="cuda")
        av = torch.as_tensor(rng.normal(size=(sc.n, 3)).astype(np.float32), device="cuda")
        out = ctx.backward(ax=ax, av=av)
        res = {k: t.cpu().numpy() for k, t in out.items()}
    else:
        ctx.step(steps)
        res = {k: t.cpu().numpy() for k, t in ctx.get_particles().items()}
    ctx.check_state()
    ctx.close()
    return res


@pytest.mark.parametrize("launch", [8, 100, 500, 1001])
def test_launch_sizes_not_multiple_of_8(gpu, launch):
    """Workgroup counts that do not divide over the 8 XCDs (and one per XCD)
    take every work item: forward and backward equal the default launch to
    float rounding.  Two float32 runs whose summation orders differ (the
    binning's atomics depend on the launch and on timing) agree to about 1e-6
    (1e-5 for dL/dE, a sum with cancellation over the steps), measured here
    between two runs of the default launch too; a work item left out would
    change its particles' results by O(1)."""
    from physgrad_amd.scenes import Scene
    rng = np.random.default_rng(9)
    n = 20000
    x = (0.2 + 0.6 * rng.random((n, 3))).astype(np.float32)
    v = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    E = np.exp(rng.uniform(np.log(5e3), np.log(5e4), n)).astype(np.float32)
    sc = Scene("launch", (48, 48, 48), 1 / 48, 1e-4, 5, 0, 2e4, 0.3, 1000.0, 0.6 ** 3 / n, x, v, E_particles=E)
    print("launch", launch, flush=True)

    def errs(got, want, keys):
        return {k: float(np.abs(got[k].astype(np.float64) - want[k]).max() / np.abs(want[k]).max()) for k in keys}

    want = run(sc, 0, 5)
    again = errs(run(sc, 0, 5), want, ("x", "v", "C", "F"))
    got = errs(run(sc, launch, 5), want, ("x", "v", "C", "F"))
    print("launch", launch, "forward", got, "default run to run", again, flush=True)
    for k, err in got.items():
        assert err <= 1e-5, (k, err)
    want = run(sc, 0, 4, backward=True)
    again = errs(run(sc, 0, 4, backward=True), want, ("x", "v", "C", "F", "E"))
    got = errs(run(sc, launch, 4, backward=True), want, ("x", "v", "C", "F", "E"))
    print("launch", launch, "backward", got, "default run to run", again, flush=True)
    for k, err in got.items():
        assert err <= (1e-4 if k == "E" else 1e-5), (k, err)
