"""CPU: the oracle's neighbour lists (getNeighbors) and SPH-MPM coupling pairs
and forces (oracle/coupling.c) against the reference's own classes compiled
from its headers (oracle/_ref: SpatialHashGrid, SPHMPMCouplingSystem),
bit for bit, and against independent numpy restatements and closed forms."""
import numpy as np
import pytest

from oracle import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference sources)")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@needs_ref
@pytest.mark.parametrize("cell,dmax,radius,n,nq", [
    (0.1, (1.0, 1.0, 1.0), 0.1, 2000, 500),      # 10^3 cells: morton codes collide modulo 1000
    (0.125, (1.0, 1.0, 1.0), 0.1, 3000, 400),    # 8^3 cells: no collisions
    (0.13, (1.0, 0.7, 0.9), 0.3, 2000, 300),     # R = 3 (343 cells per query)
    (0.05, (1.0, 1.0, 1.0), 0.04, 5000, 700),
])
def test_neighbors_match_reference(cell, dmax, radius, n, nq):
    rng = np.random.default_rng(int(cell * 1000) + n)
    dmax = np.array(dmax, np.float32)
    pos = (rng.random((n, 3)) * dmax).astype(np.float32)
    # queries inside and up to 1/4 cell outside (the reference's conversion
    # is defined there), and some on cell boundaries
    q = (rng.random((nq, 3)) * (dmax + 0.5 * cell) - 0.25 * cell).astype(np.float32)
    q[:20] = np.round(q[:20] / cell) * cell
    ids, idx, st, ct = O.reference_spatial_hash(pos, cell, (0, 0, 0), dmax)
    res = np.ceil(dmax / np.float32(cell)).astype(np.uint32)
    rp, nb = O.hash_neighbors(q, radius, cell, (0, 0, 0), res, idx, st, ct)
    rrp, rnb = O.reference_hash_neighbors(pos, cell, (0, 0, 0), dmax, q, radius)
    np.testing.assert_array_equal(rp, rrp)
    np.testing.assert_array_equal(nb, rnb)


def test_neighbors_cover_every_point_within_radius():
    """Independent of the reference: every particle within radius of a query
    (numpy distances) is in its list (the lists are supersets)."""
    rng = np.random.default_rng(3)
    cell, radius = 0.1, 0.17
    pos = rng.random((3000, 3)).astype(np.float32)
    q = rng.random((200, 3)).astype(np.float32)
    res = np.array([10, 10, 10], np.uint32)
    ids = O.cell_ids(pos, cell, (0, 0, 0), res)
    idx, st, ct = O.counting_sort(ids, 1000)
    rp, nb = O.hash_neighbors(q, radius, cell, (0, 0, 0), res, idx, st, ct)
    d = np.linalg.norm(pos[None, :, :].astype(np.float64) - q[:, None, :], axis=2)
    for i in range(len(q)):
        row = set(nb[rp[i]:rp[i + 1]].tolist())
        assert set(np.nonzero(d[i] < radius)[0].tolist()) <= row


def _numpy_pairs(xs, xm, radius):
    """findCouplingPairs in numpy float32, the reference's operation order."""
    r = xm[None, :, :] - xs[:, None, :]
    d = np.sqrt((r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1]) + r[..., 2] * r[..., 2])
    s, m = np.nonzero(d < np.float32(radius))
    return s, m, (np.float32(radius) - d[s, m]).astype(np.float32)


def _scene(rng, ns, nm, dup=0):
    xs = rng.random((ns, 3)).astype(np.float32)
    xm = rng.random((nm, 3)).astype(np.float32)
    if dup:
        xm[:dup] = xs[:dup]        # coincident particles: distance 0, forces skipped
    vs = rng.normal(size=(ns, 3)).astype(np.float32)
    vm = rng.normal(size=(nm, 3)).astype(np.float32)
    # pairs without relative motion: close particles with equal velocities
    xm[dup:dup + 5] = xs[dup:dup + 5] + np.float32(0.01)
    vm[dup:dup + 5] = vs[dup:dup + 5]
    return xs, vs, xm, vm


def test_pairs_match_numpy_float32():
    rng = np.random.default_rng(4)
    xs, vs, xm, vm = _scene(rng, 400, 500, dup=3)
    rp, col, pen = O.coupling_pairs(xs, xm, 0.1)
    s, m, p = _numpy_pairs(xs, xm, 0.1)
    np.testing.assert_array_equal(np.repeat(np.arange(len(xs)), np.diff(rp)), s)
    np.testing.assert_array_equal(col, m)
    np.testing.assert_array_equal(_bits(pen), _bits(p))


@needs_ref
@pytest.mark.parametrize("two_way", [True, False])
def test_coupling_matches_reference(two_way):
    rng = np.random.default_rng(5 + two_way)
    xs, vs, xm, vm = _scene(rng, 300, 400, dup=4)
    rp, col, pen = O.coupling_pairs(xs, xm, 0.12)
    fs, ff, sf = O.coupling_forces(xs, vs, xm, vm, rp, col, pen, 1e5, 1e3, 0.3, two_way)
    r = O.reference_coupling(xs, vs, xm, vm, 0.12, 1e5, 1e3, 0.3, two_way)
    assert len(col) > 100
    np.testing.assert_array_equal(r["sph"], np.repeat(np.arange(len(xs)), np.diff(rp)))
    np.testing.assert_array_equal(r["mpm"], col)
    np.testing.assert_array_equal(_bits(r["penetration"]), _bits(pen))
    np.testing.assert_array_equal(_bits(r["solid_force"]), _bits(fs))
    np.testing.assert_array_equal(_bits(r["fluid_force"]), _bits(ff))
    np.testing.assert_array_equal(_bits(r["sph_force"]), _bits(sf))


def test_force_closed_forms():
    """Per pair, in float64 from the float32 inputs: solid = -fluid; the
    normal component is stiffness pen + damping (dv . n); the tangential
    component is -mu stiffness pen along the tangential relative velocity
    (zero when there is none); coincident particles get no force; the SPH
    reaction is the sum of its pairs' fluid forces."""
    rng = np.random.default_rng(6)
    xs, vs, xm, vm = _scene(rng, 300, 400, dup=4)
    k, c, mu, R = 1e5, 1e3, 0.3, 0.12
    rp, col, pen = O.coupling_pairs(xs, xm, R)
    fs, ff, sf = O.coupling_forces(xs, vs, xm, vm, rp, col, pen, k, c, mu, True)
    rows = np.repeat(np.arange(len(xs)), np.diff(rp))
    np.testing.assert_array_equal(fs, -ff)
    r = xm[col].astype(np.float64) - xs[rows]
    d = np.linalg.norm(r, axis=1)
    zero = d < 1e-6
    assert zero.sum() >= 4 and not fs[zero].any()
    n = r[~zero] / d[~zero, None]
    dv = vm[col][~zero].astype(np.float64) - vs[rows][~zero]
    vn = (dv * n).sum(1)
    f = fs[~zero].astype(np.float64)
    fn_want = k * pen[~zero] + c * vn
    np.testing.assert_allclose((f * n).sum(1), fn_want, rtol=1e-4, atol=1e-4 * np.abs(fn_want).max())
    t = dv - n * vn[:, None]
    ts = np.linalg.norm(t, axis=1)
    moving = ts > 1e-3
    ft = f - n * (f * n).sum(1)[:, None]
    want_t = -(mu * k * pen[~zero])[:, None] * t / np.maximum(ts, 1e-30)[:, None]
    np.testing.assert_allclose(ft[moving], want_t[moving], rtol=1e-3, atol=1e-3 * np.abs(want_t).max())
    assert (~moving).sum() >= 1 and np.abs(ft[~moving]).max() <= 1e-3 * np.abs(f).max()
    acc = np.zeros((len(xs), 3))
    np.add.at(acc, rows, ff.astype(np.float64))
    np.testing.assert_allclose(sf, acc, rtol=1e-5, atol=1e-5 * np.abs(acc).max())


def test_empty_inputs():
    rp, col, pen = O.coupling_pairs(np.zeros((0, 3), np.float32), np.ones((5, 3), np.float32), 0.1)
    assert rp.tolist() == [0] and len(col) == 0
    rp, col, pen = O.coupling_pairs(np.ones((4, 3), np.float32), np.zeros((0, 3), np.float32), 0.1)
    assert rp.tolist() == [0] * 5 and len(col) == 0


def test_csr_spmv_matches_scipy():
    import scipy.sparse as sp
    rng = np.random.default_rng(8)
    nr, nc = 200, 150
    A = sp.random(nr, nc, density=0.05, random_state=1, format="csr", dtype=np.float64)
    A.data = A.data.astype(np.float32).astype(np.float64)
    x = rng.normal(size=nc).astype(np.float32)
    y = O.csr_spmv(A.indptr.astype(np.uint32), A.indices.astype(np.uint32), A.data.astype(np.float32), x)
    np.testing.assert_allclose(y, A @ x.astype(np.float64), rtol=1e-5, atol=1e-5 * np.abs(A @ x).max())
    # float sums in stored order: bit-equal to a numpy float32 restatement
    want = np.zeros(nr, np.float32)
    for r in range(nr):
        s = np.float32(0)
        for k in range(A.indptr[r], A.indptr[r + 1]):
            s = np.float32(s + np.float32(np.float32(A.data[k]) * x[A.indices[k]]))
        want[r] = s
    np.testing.assert_array_equal(y.view(np.uint32), want.view(np.uint32))


@needs_ref
def test_reference_sparse_matrix_loses_its_rows():
    """The reference SparseMatrix (addElement drops the row, finalize leaves the
    row pointers at zero): every product is zero.  The drop-in records rows."""
    rng = np.random.default_rng(9)
    rows = np.sort(rng.integers(0, 20, 100)).astype(np.uint32)
    cols = rng.integers(0, 10, 100).astype(np.uint32)
    vals = rng.normal(size=100).astype(np.float32) + 3
    rp, y = O.reference_sparse_matrix(20, 10, rows, cols, vals, np.ones(10, np.float32))
    assert not rp.any() and not y.any()
