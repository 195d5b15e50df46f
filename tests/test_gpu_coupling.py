"""GPU: neighbour lists of the spatial hash (pg_hash_neighbors_*) and the
SPH-MPM coupling pairs and forces (pg_coupling_*) against the oracle
(oracle/coupling.c, itself pinned bit-exact to the reference's
SpatialHashGrid::getNeighbors and SPHMPMCouplingSystem), bit for bit: same
CSR rows, same order, same float bits."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32)


def _np(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("cell,dmax,radius,n,nq", [
    (0.1, (1.0, 1.0, 1.0), 0.1, 20000, 5000),     # hash slots collide (10^3 cells)
    (0.125, (1.0, 1.0, 1.0), 0.1, 30000, 4000),
    (0.13, (1.0, 0.7, 0.9), 0.3, 5000, 1000),     # R = 3
    (0.05, (1.0, 1.0, 1.0), 0.04, 1, 3),
])
def test_neighbors_match_oracle(gpu, oracle_lib, cell, dmax, radius, n, nq):
    from physgrad_amd.spatial import SpatialHashGrid
    rng = np.random.default_rng(n + nq)
    dmax = np.array(dmax, np.float32)
    pos = (rng.random((n, 3)) * dmax).astype(np.float32)
    q = (rng.random((nq, 3)) * (dmax + 4 * cell) - 2 * cell).astype(np.float32)   # also far outside
    print("neighbours", cell, radius, n, nq, flush=True)
    h = SpatialHashGrid(cell, (0, 0, 0), dmax, max_particles=n)
    h.buildHashTable(torch.as_tensor(pos))
    rp, nb = h.neighbors(torch.as_tensor(q), radius)
    idx, st, ct = (_np(t) for t in (h.getParticleIndices(), h.getCellStartIndices(), h.getCellParticleCounts()))
    # the hash itself equals the oracle's
    ids = O.cell_ids(pos, cell, (0, 0, 0), h.res)
    oidx, ost, oct_ = O.counting_sort(ids, h.total_cells)
    np.testing.assert_array_equal(idx, oidx)
    np.testing.assert_array_equal(ct, oct_)
    orp, onb = O.hash_neighbors(q, radius, cell, (0, 0, 0), h.res, oidx, ost, oct_)
    np.testing.assert_array_equal(_np(rp), orp)
    np.testing.assert_array_equal(_np(nb), onb)
    assert h.getNeighbors(*q[0].tolist(), radius) == onb[orp[0]:orp[1]].tolist()


def _run_pairs(xs, vs, xm, vm, radius, lo=(0, 0, 0), hi=(1, 1, 1), two_way=True, f0=None):
    from physgrad_amd.spatial import SPHMPMCoupling
    c = SPHMPMCoupling(coupling_radius=radius, domain_min=lo, domain_max=hi, enable_two_way_coupling=two_way)
    n = c.findCouplingPairs(torch.as_tensor(xs), torch.as_tensor(xm))
    rp, col, pen = c.getCouplingPairs()
    sf = None if f0 is None else torch.as_tensor(f0).cuda().contiguous()
    out = c.computeCouplingForces(torch.as_tensor(xs), torch.as_tensor(vs), torch.as_tensor(xm), torch.as_tensor(vm),
                                  sph_force=sf)
    res = dict(n=n, rp=_np(rp), col=_np(col), pen=_np(pen), fs=_np(out["solid_force"]),
               ff=_np(out["fluid_force"]), sf=_np(out["sph_force"]), energy=c.getTotalCouplingEnergy())
    c.close()
    return res


def _check_vs_oracle(res, xs, vs, xm, vm, radius, two_way=True, f0=None):
    orp, ocol, open_ = O.coupling_pairs(xs, xm, radius)
    assert res["n"] == len(ocol)
    np.testing.assert_array_equal(res["rp"], orp)
    np.testing.assert_array_equal(res["col"], ocol)
    np.testing.assert_array_equal(_bits(res["pen"]), _bits(open_))
    fs, ff, sf = O.coupling_forces(xs, vs, xm, vm, orp, ocol, open_, 1e5, 1e3, 0.3, two_way, sph_force=f0)
    np.testing.assert_array_equal(_bits(res["fs"]), _bits(fs))
    np.testing.assert_array_equal(_bits(res["ff"]), _bits(ff))
    np.testing.assert_array_equal(_bits(res["sf"]), _bits(sf))
    want_e = 0.5 * 1e5 * float((open_.astype(np.float64) ** 2).sum())
    assert abs(res["energy"] - want_e) <= 1e-9 * max(want_e, 1.0)


@pytest.mark.parametrize("two_way", [True, False])
def test_pairs_and_forces_match_oracle(gpu, oracle_lib, two_way):
    rng = np.random.default_rng(40 + two_way)
    ns, nm, radius = 3000, 4000, 0.06
    xs = rng.random((ns, 3)).astype(np.float32)
    xm = rng.random((nm, 3)).astype(np.float32)
    xm[:5] = xs[:5]                                    # coincident: forces skipped
    xm[5:10] = xs[5:10] + np.float32(0.01)             # no relative motion below
    vs = rng.normal(size=(ns, 3)).astype(np.float32)
    vm = rng.normal(size=(nm, 3)).astype(np.float32)
    vm[5:10] = vs[5:10]
    f0 = rng.normal(size=(ns, 3)).astype(np.float32)
    print("pairs two_way", two_way, flush=True)
    res = _run_pairs(xs, vs, xm, vm, radius, two_way=two_way, f0=f0)
    assert res["n"] > 1000
    _check_vs_oracle(res, xs, vs, xm, vm, radius, two_way=two_way, f0=f0)


def test_pairs_outside_the_box_and_colliding_slots(gpu, oracle_lib):
    """MPM and SPH particles beyond the binning box (clamped into its border
    cells) and a box whose cell count makes hash slots collide."""
    rng = np.random.default_rng(43)
    ns, nm, radius = 2000, 3000, 0.11
    xs = (rng.random((ns, 3)) * 1.6 - 0.3).astype(np.float32)
    xm = (rng.random((nm, 3)) * 1.6 - 0.3).astype(np.float32)
    vs = rng.normal(size=(ns, 3)).astype(np.float32)
    vm = rng.normal(size=(nm, 3)).astype(np.float32)
    print("pairs outside / collisions", flush=True)
    res = _run_pairs(xs, vs, xm, vm, radius, lo=(0.0, 0.1, 0.2), hi=(0.97, 1.0, 0.93))
    _check_vs_oracle(res, xs, vs, xm, vm, radius)


@pytest.mark.parametrize("per_row", [3000, 20000], ids=["lds_cap", "merge"])
def test_long_rows(gpu, oracle_lib, per_row):
    """Rows beyond the in-LDS sort (2048 pairs) and beyond one LDS block
    (8192: merge passes), mixed with short rows."""
    rng = np.random.default_rng(44 + per_row)
    radius = 0.05
    centres = np.array([[0.3, 0.3, 0.3], [0.7, 0.6, 0.5]], np.float32)
    blob = [c + (rng.normal(size=(per_row, 3)) * 0.01).astype(np.float32) for c in centres]
    xm = np.concatenate([rng.random((2000, 3)).astype(np.float32)] + blob)
    perm = rng.permutation(len(xm))
    xm = xm[perm]                                      # blob members spread over the index range
    xs = np.concatenate([centres, rng.random((500, 3)).astype(np.float32)])
    vs = rng.normal(size=xs.shape).astype(np.float32)
    vm = rng.normal(size=xm.shape).astype(np.float32)
    print("long rows", per_row, flush=True)
    res = _run_pairs(xs, vs, xm, vm, radius)
    assert np.diff(res["rp"])[:2].min() > per_row * 0.9
    _check_vs_oracle(res, xs, vs, xm, vm, radius)


def test_empty_and_pairless(gpu, oracle_lib):
    from physgrad_amd.spatial import SPHMPMCoupling
    c = SPHMPMCoupling(coupling_radius=0.01)
    z = torch.zeros((0, 3))
    assert c.findCouplingPairs(z, torch.rand(100, 3)) == 0
    rp, col, pen = c.getCouplingPairs()
    assert _np(rp).tolist() == [0] and col.numel() == 0
    assert c.findCouplingPairs(torch.rand(10, 3), z) == 0
    assert _np(c.getCouplingPairs()[0]).tolist() == [0] * 11
    xs = torch.tensor([[0.1, 0.1, 0.1]])
    xm = torch.tensor([[0.9, 0.9, 0.9]])
    assert c.findCouplingPairs(xs, xm) == 0
    out = c.computeCouplingForces(xs, xs, xm, xm)
    assert out["solid_force"].numel() == 0 and not out["sph_force"].any()
    c.close()


def test_bad_config_is_rejected(gpu):
    from physgrad_amd._lib import PgError
    from physgrad_amd.spatial import SPHMPMCoupling
    with pytest.raises(PgError, match="coupling_radius"):
        SPHMPMCoupling(coupling_radius=0.0)
    with pytest.raises(PgError, match="2\\^28"):
        SPHMPMCoupling(coupling_radius=1e-4, domain_max=(1.0, 1.0, 1.0))


def test_million_particles_properties(gpu, oracle_lib):
    """1M SPH and 1M MPM particles: the CSR is well formed, rows are ascending,
    every pair is within the radius, and sampled rows equal the oracle's brute
    force over all MPM particles."""
    rng = np.random.default_rng(45)
    n = 1 << 20
    radius = 0.012
    xs = rng.random((n, 3)).astype(np.float32)
    xm = rng.random((n, 3)).astype(np.float32)
    print("1M x 1M pairs", flush=True)
    res = _run_pairs(xs, np.zeros_like(xs), xm, np.zeros_like(xm), radius)
    rp, col, pen = res["rp"], res["col"], res["pen"]
    assert rp[0] == 0 and rp[-1] == res["n"] and (np.diff(rp.astype(np.int64)) >= 0).all()
    expect = n * n * (4 / 3 * np.pi * radius ** 3)
    assert 0.9 * expect < res["n"] < 1.1 * expect
    rows = np.repeat(np.arange(n), np.diff(rp))
    same = rows[1:] == rows[:-1]
    assert (col[1:][same] > col[:-1][same]).all()
    assert (pen > 0).all() and (pen <= radius).all()
    sample = rng.choice(n, 64, replace=False)
    orp, ocol, opn = O.coupling_pairs(xs[sample], xm, radius)
    for i, s in enumerate(sample):
        np.testing.assert_array_equal(col[rp[s]:rp[s + 1]], ocol[orp[i]:orp[i + 1]])
        np.testing.assert_array_equal(_bits(pen[rp[s]:rp[s + 1]]), _bits(opn[orp[i]:orp[i + 1]]))


def test_csr_spmv_vs_oracle_and_column_guard(gpu, oracle_lib):
    """pg_csr_spmv over the coupling pairs as a matrix (rows: SPH particles,
    columns: MPM particles, values: penetrations) against the oracle's
    row-ordered sums, bit for bit; a column index past x makes its row NaN."""
    import ctypes
    from physgrad_amd import _lib
    from physgrad_amd.spatial import SPHMPMCoupling
    rng = np.random.default_rng(46)
    ns, nm = 4000, 5000
    xs = torch.rand(ns, 3)
    xm = torch.rand(nm, 3)
    c = SPHMPMCoupling(coupling_radius=0.05)
    c.findCouplingPairs(xs, xm)
    rp, col, pen = c.getCouplingPairs()
    c.close()
    x = torch.as_tensor(rng.normal(size=nm).astype(np.float32)).cuda()
    rp32, col32 = rp.to(torch.int32).cuda(), col.to(torch.int32).cuda()
    y = torch.empty(ns, dtype=torch.float32, device="cuda")
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = _lib.load()
    assert L.pg_csr_spmv(p(rp32), p(col32), p(pen), p(x), nm, p(y), ns, st) == 0
    want = oracle_lib.csr_spmv(_np(rp), _np(col), _np(pen), _np(x))
    np.testing.assert_array_equal(_np(y).view(np.uint32), want.view(np.uint32))
    # the same with x one column short: rows that reach the last column are NaN
    assert L.pg_csr_spmv(p(rp32), p(col32), p(pen), p(x), nm - 1, p(y), ns, st) == 0
    rows_last = np.repeat(np.arange(ns), np.diff(_np(rp)))[_np(col) == nm - 1]
    got = _np(y)
    assert np.isnan(got[rows_last]).all()
    others = np.setdiff1d(np.arange(ns), rows_last)
    np.testing.assert_array_equal(got[others].view(np.uint32), want[others].view(np.uint32))
