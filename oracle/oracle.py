"""ctypes front-end of the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the physgrad_amd package.

Functions follow the C sources in this directory; each C function cites the
reference file:line it restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_mpm.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")

FIXED_COROTATED = 0
NEO_HOOKEAN = 1
FLUID = 2


class RefParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float),
                ("ox", ctypes.c_float), ("oy", ctypes.c_float), ("oz", ctypes.c_float),
                ("dt", ctypes.c_float),
                ("gx", ctypes.c_float), ("gy", ctypes.c_float), ("gz", ctypes.c_float),
                ("order", ctypes.c_int)]


class MlsParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double),
                ("ox", ctypes.c_double), ("oy", ctypes.c_double), ("oz", ctypes.c_double),
                ("dt", ctypes.c_double),
                ("gx", ctypes.c_double), ("gy", ctypes.c_double), ("gz", ctypes.c_double),
                ("bc", ctypes.c_int), ("material", ctypes.c_int),
                ("nu", ctypes.c_double), ("p_vol", ctypes.c_double)]


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and the reference driver when ref=True)."""
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.check_call(["make", "-s", "-C", HERE] + targets)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_ref_run.argtypes = [ctypes.POINTER(RefParams), ctypes.c_int, _f32p, _f32p, _f32p,
                                  ctypes.c_int, _f32p, _f32p, _f32p]
        L.orc_mls_run_f64.argtypes = [ctypes.POINTER(MlsParams), ctypes.c_int, _f64p, _f64p, _f64p,
                                      _f64p, _f64p, _f64p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_mls_run_f32.argtypes = [ctypes.POINTER(MlsParams), ctypes.c_int, _f32p, _f32p, _f32p,
                                      _f32p, _f32p, _f32p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_mls_run_f32_mt.argtypes = [ctypes.POINTER(MlsParams), ctypes.c_int, _f32p, _f32p, _f32p,
                                         _f32p, _f32p, _f32p, ctypes.c_int, ctypes.c_int]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_kirchhoff_f64.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _f64p, _f64p]
        L.orc_polar_f64.argtypes = [_f64p, _f64p, _f64p]
        L.orc_polar_f64.restype = ctypes.c_int
        L.orc_mls_adjoint_step_f64.argtypes = [ctypes.POINTER(MlsParams), ctypes.c_int,
                                               _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                               _f64p, _f64p, _f64p, _f64p, _f64p]
        _ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.orc_mls_run_ext_f64.argtypes = [ctypes.POINTER(MlsParams), ctypes.c_int, _f64p, _f64p, _f64p, _f64p,
                                          _f64p, _f64p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p]
        L.orc_kirchhoff_ext_f64.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            _f64p, _f64p, _f64p]
        L.orc_morton3.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint]
        L.orc_morton3.restype = ctypes.c_uint
        L.orc_cell_ids.argtypes = [_f32p, ctypes.c_uint, ctypes.c_float, _f32p, _u32p, _u32p]
        L.orc_counting_sort.argtypes = [_u32p, ctypes.c_uint, ctypes.c_uint, _u32p, _u32p, _u32p]
        L.orc_hash_neighbors.argtypes = [_f32p, ctypes.c_uint, ctypes.c_float, ctypes.c_float, _f32p, _u32p, _u32p,
                                         _u32p, _u32p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_hash_neighbors.restype = ctypes.c_ulonglong
        L.orc_coupling_pairs.argtypes = [_f32p, ctypes.c_uint, _f32p, ctypes.c_uint, ctypes.c_float,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_coupling_pairs.restype = ctypes.c_ulonglong
        L.orc_coupling_forces.argtypes = [_f32p, _f32p, ctypes.c_uint, _f32p, _f32p, _u32p, _u32p, _f32p,
                                          ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                          _f32p, _f32p, _f32p]
        L.orc_csr_spmv.argtypes = [_u32p, _u32p, _f32p, _f32p, ctypes.c_uint, _f32p]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own CPU solver (oracle/_ref), or None when not built."""
    global _ref
    if _ref is None and ref_available():
        R = ctypes.CDLL(REF_PATH)
        R.ref_mpm_run.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_float] * 6 + [ctypes.c_float] * 4 +
                                  [ctypes.c_int, ctypes.c_int, _f32p, _f32p, _f32p, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_double)])
        R.ref_spatial_hash.argtypes = [_f32p, ctypes.c_uint, ctypes.c_float, _f32p, _f32p,
                                       _u32p, _u32p, _u32p, _u32p, ctypes.POINTER(ctypes.c_uint)]
        R.ref_hash_neighbors.argtypes = [_f32p, ctypes.c_uint, ctypes.c_float, _f32p, _f32p, _f32p, ctypes.c_uint,
                                         ctypes.c_float, _u32p, _u32p, ctypes.c_ulonglong]
        R.ref_hash_neighbors.restype = ctypes.c_longlong
        R.ref_coupling.argtypes = [_f32p, _f32p, ctypes.c_uint, _f32p, _f32p, ctypes.c_uint, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int, _u32p, _u32p,
                                   _f32p, _f32p, _f32p, _f32p, ctypes.c_ulonglong]
        R.ref_coupling.restype = ctypes.c_longlong
        R.ref_sparse_matrix.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, _u32p, _u32p, _f32p, _f32p,
                                        _f32p, _u32p]
        _ref = R
    return _ref


# ----------------------------------------------------------------------------- ref-compat

def ref_params(dims, cell, origin, dt, gravity, order=2) -> RefParams:
    return RefParams(int(dims[0]), int(dims[1]), int(dims[2]),
                     float(cell[0]), float(cell[1]), float(cell[2]),
                     float(origin[0]), float(origin[1]), float(origin[2]), float(dt),
                     float(gravity[0]), float(gravity[1]), float(gravity[2]), int(order))


def refcompat_run(prm: RefParams, x, v, m, nsteps):
    """Restated MPMSolver<float,64>::step x nsteps. Returns (x, v, grid_mass, grid_mom, grid_vel)."""
    x = np.ascontiguousarray(x, np.float32).reshape(-1).copy()
    v = np.ascontiguousarray(v, np.float32).reshape(-1).copy()
    m = np.ascontiguousarray(m, np.float32).reshape(-1)
    G = prm.nx * prm.ny * prm.nz
    gm = np.zeros(G, np.float32)
    gp = np.zeros(3 * G, np.float32)
    gv = np.zeros(3 * G, np.float32)
    lib().orc_ref_run(ctypes.byref(prm), m.size, x, v, m, int(nsteps), gm, gp, gv)
    return x.reshape(-1, 3), v.reshape(-1, 3), gm, gp.reshape(-1, 3), gv.reshape(-1, 3)


def reference_run(prm: RefParams, x, v, m, nsteps, want_grid=True):
    """The reference's compiled MPMSolver<float,64> (oracle/_ref). Returns like refcompat_run + seconds."""
    R = ref()
    if R is None:
        raise RuntimeError("oracle/_ref/libref_mpm.so is not built")
    x = np.ascontiguousarray(x, np.float32).reshape(-1).copy()
    v = np.ascontiguousarray(v, np.float32).reshape(-1).copy()
    m = np.ascontiguousarray(m, np.float32).reshape(-1)
    G = prm.nx * prm.ny * prm.nz
    gm = np.zeros(G, np.float32) if want_grid else None
    gp = np.zeros(3 * G, np.float32) if want_grid else None
    gv = np.zeros(3 * G, np.float32) if want_grid else None
    secs = ctypes.c_double(0.0)
    ptr = (lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None)
    R.ref_mpm_run(prm.nx, prm.ny, prm.nz, prm.dx, prm.dy, prm.dz, prm.ox, prm.oy, prm.oz,
                  prm.dt, prm.gx, prm.gy, prm.gz, prm.order, m.size, x, v, m, int(nsteps),
                  ptr(gm), ptr(gp), ptr(gv), ctypes.byref(secs))
    out = (x.reshape(-1, 3), v.reshape(-1, 3), gm, None if gp is None else gp.reshape(-1, 3),
           None if gv is None else gv.reshape(-1, 3))
    return out + (secs.value,)


# ----------------------------------------------------------------------------- MLS-MPM

def mls_params(dims, dx, origin, dt, gravity, bc, material, nu, p_vol) -> MlsParams:
    return MlsParams(int(dims[0]), int(dims[1]), int(dims[2]), float(dx),
                     float(origin[0]), float(origin[1]), float(origin[2]), float(dt),
                     float(gravity[0]), float(gravity[1]), float(gravity[2]),
                     int(bc), int(material), float(nu), float(p_vol))


def scene_params(sc) -> MlsParams:
    """MlsParams of a scene object with the attributes dims, dx, origin, dt,
    gravity, bc, material, nu, p_vol (physgrad_amd.scenes.Scene)."""
    return mls_params(sc.dims, sc.dx, sc.origin, sc.dt, sc.gravity, sc.bc, sc.material, sc.nu, sc.p_vol)


def _state(x, v, C, F, dtype):
    n = np.asarray(x).reshape(-1, 3).shape[0]
    x = np.array(x, dtype).reshape(-1).copy()
    v = np.array(v, dtype).reshape(-1).copy()
    C = np.zeros(9 * n, dtype) if C is None else np.array(C, dtype).reshape(-1).copy()
    F = np.tile(np.eye(3, dtype=dtype).reshape(-1), n) if F is None else np.array(F, dtype).reshape(-1).copy()
    return n, x, v, C, F


def mls_run(prm: MlsParams, x, v, C, F, mass, E, nsteps, dtype=np.float64, want_grid=False):
    """MLS-MPM forward x nsteps in float64 (oracle) or float32 (CPU port).

    Returns dict(x, v, C, F[, grid_m, grid_v])."""
    n, x, v, C, F = _state(x, v, C, F, dtype)
    mass = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, dtype), (n,)), dtype)
    E = np.ascontiguousarray(np.broadcast_to(np.asarray(E, dtype), (n,)), dtype)
    G = prm.nx * prm.ny * prm.nz
    gm = np.zeros(G, dtype) if want_grid else None
    gv = np.zeros(3 * G, dtype) if want_grid else None
    ptr = (lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None)
    fn = lib().orc_mls_run_f64 if dtype == np.float64 else lib().orc_mls_run_f32
    fn(ctypes.byref(prm), n, x, v, C, F, mass, E, int(nsteps), ptr(gm), ptr(gv))
    out = dict(x=x.reshape(n, 3), v=v.reshape(n, 3), C=C.reshape(n, 3, 3), F=F.reshape(n, 3, 3))
    if want_grid:
        out["grid_m"] = gm
        out["grid_v"] = gv.reshape(-1, 3)
    return out


def mls_run_ext(prm: MlsParams, x, v, C, F, mass, E, nsteps, model=None, nu=None, visc=None, vol=None,
                flags=None, want_grid=False):
    """FP64 MLS-MPM with per-particle model / Poisson ratio / viscosity /
    volume and per-node boundary flags (None: the scene's values, no flags).
    Returns dict(x, v, C, F[, grid_m, grid_v])."""
    n, x, v, C, F = _state(x, v, C, F, np.float64)
    arr = lambda a, dt: None if a is None else np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt), (n,)))  # noqa: E731
    mass = arr(mass, np.float64)
    E = arr(E, np.float64)
    model, nu, visc, vol = arr(model, np.int32), arr(nu, np.float64), arr(visc, np.float64), arr(vol, np.float64)
    G = prm.nx * prm.ny * prm.nz
    fl = None if flags is None else np.ascontiguousarray(np.asarray(flags, np.uint32).reshape(G))
    gm = np.zeros(G) if want_grid else None
    gv = np.zeros(3 * G) if want_grid else None
    ptr = (lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None)
    lib().orc_mls_run_ext_f64(ctypes.byref(prm), n, x, v, C, F, mass, E, ptr(model), ptr(nu), ptr(visc), ptr(vol),
                              ptr(fl), int(nsteps), ptr(gm), ptr(gv))
    out = dict(x=x.reshape(n, 3), v=v.reshape(n, 3), C=C.reshape(n, 3, 3), F=F.reshape(n, 3, 3))
    if want_grid:
        out["grid_m"] = gm
        out["grid_v"] = gv.reshape(-1, 3)
    return out


def kirchhoff_ext(model, mu, lam, visc, F, C):
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    C = np.ascontiguousarray(C, np.float64).reshape(9)
    tau = np.zeros(9, np.float64)
    lib().orc_kirchhoff_ext_f64(int(model), float(mu), float(lam), float(visc), F, C, tau)
    return tau.reshape(3, 3)


def mls_run_mt(prm: MlsParams, x, v, C, F, mass, E, nsteps, nthreads: int = 0):
    """FP32 MLS-MPM forward x nsteps on all host cores (OpenMP; nthreads 0 =
    the OpenMP default).  Returns dict(x, v, C, F)."""
    n, x, v, C, F = _state(x, v, C, F, np.float32)
    mass = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float32), (n,)), np.float32)
    E = np.ascontiguousarray(np.broadcast_to(np.asarray(E, np.float32), (n,)), np.float32)
    lib().orc_mls_run_f32_mt(ctypes.byref(prm), n, x, v, C, F, mass, E, int(nsteps), int(nthreads))
    return dict(x=x.reshape(n, 3), v=v.reshape(n, 3), C=C.reshape(n, 3, 3), F=F.reshape(n, 3, 3))


def num_threads() -> int:
    """Threads the multi-threaded CPU port uses by default."""
    return int(lib().orc_num_threads())


def mls_adjoint_step(prm: MlsParams, x, v, C, F, mass, E, ax, av, aC, aF):
    """One FP64 reverse step; returns (ax, av, aC, aF, aE) adjoints of the step's input state."""
    n, x, v, C, F = _state(x, v, C, F, np.float64)
    mass = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float64), (n,)))
    E = np.ascontiguousarray(np.broadcast_to(np.asarray(E, np.float64), (n,)))
    ax = np.array(ax, np.float64).reshape(-1).copy()
    av = np.array(av, np.float64).reshape(-1).copy()
    aC = np.array(aC, np.float64).reshape(-1).copy()
    aF = np.array(aF, np.float64).reshape(-1).copy()
    aE = np.zeros(n, np.float64)
    lib().orc_mls_adjoint_step_f64(ctypes.byref(prm), n, x, v, C, F, mass, E, ax, av, aC, aF, aE)
    return ax.reshape(n, 3), av.reshape(n, 3), aC.reshape(n, 3, 3), aF.reshape(n, 3, 3), aE


def kirchhoff(material, mu, lam, F):
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    tau = np.zeros(9, np.float64)
    lib().orc_kirchhoff_f64(int(material), float(mu), float(lam), F, tau)
    return tau.reshape(3, 3)


def polar(F):
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    R = np.zeros(9, np.float64)
    S = np.zeros(9, np.float64)
    it = lib().orc_polar_f64(F, R, S)
    return R.reshape(3, 3), S.reshape(3, 3), it


# ----------------------------------------------------------------------------- cell keys

def cell_ids(pos, cell_size, dmin, res):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
    n = pos.size // 3
    ids = np.zeros(n, np.uint32)
    lib().orc_cell_ids(pos, n, float(cell_size), np.asarray(dmin, np.float32),
                       np.asarray(res, np.uint32), ids)
    return ids


def counting_sort(ids, total_cells):
    ids = np.ascontiguousarray(ids, np.uint32)
    n = ids.size
    idx = np.zeros(n, np.uint32)
    st = np.zeros(total_cells, np.uint32)
    ct = np.zeros(total_cells, np.uint32)
    lib().orc_counting_sort(ids, n, int(total_cells), idx, st, ct)
    return idx, st, ct


def reference_spatial_hash(pos, cell_size, dmin, dmax):
    R = ref()
    if R is None:
        raise RuntimeError("oracle/_ref/libref_mpm.so is not built")
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
    n = pos.size // 3
    dmin = np.asarray(dmin, np.float32)
    dmax = np.asarray(dmax, np.float32)
    res = np.ceil((dmax - dmin) / np.float32(cell_size)).astype(np.uint64)
    tot = int(np.prod(res))
    ids = np.zeros(n, np.uint32)
    idx = np.zeros(n, np.uint32)
    st = np.zeros(tot, np.uint32)
    ct = np.zeros(tot, np.uint32)
    tc = ctypes.c_uint(0)
    R.ref_spatial_hash(pos, n, float(cell_size), dmin, dmax, ids, idx, st, ct, ctypes.byref(tc))
    return ids, idx, st, ct


# ----------------------------------------------------------------------------- neighbours and coupling

def _f3(a):
    return np.ascontiguousarray(a, np.float32).reshape(-1)


def hash_neighbors(queries, radius, cell_size, dmin, res, particle_indices, cell_start, cell_count):
    """SpatialHashGrid::getNeighbors for every query over a built hash.
    Returns (row_ptr [nq+1] uint32, neighbours uint32)."""
    q = _f3(queries)
    nq = q.size // 3
    dmin = np.asarray(dmin, np.float32)
    res = np.asarray(res, np.uint32)
    pi, st, ct = (np.ascontiguousarray(a, np.uint32) for a in (particle_indices, cell_start, cell_count))
    counts = np.zeros(nq, np.uint32)
    total = lib().orc_hash_neighbors(q, nq, float(radius), float(cell_size), dmin, res, pi, st, ct,
                                     counts.ctypes.data_as(ctypes.c_void_p), None)
    out = np.zeros(max(int(total), 1), np.uint32)
    lib().orc_hash_neighbors(q, nq, float(radius), float(cell_size), dmin, res, pi, st, ct, None,
                             out.ctypes.data_as(ctypes.c_void_p))
    row_ptr = np.zeros(nq + 1, np.uint32)
    row_ptr[1:] = np.cumsum(counts, dtype=np.uint64).astype(np.uint32)
    return row_ptr, out[:int(total)]


def coupling_pairs(sph_x, mpm_x, radius):
    """findCouplingPairs: (row_ptr [ns+1], col, penetration), SPH-major, MPM ascending."""
    xs, xm = _f3(sph_x), _f3(mpm_x)
    ns, nm = xs.size // 3, xm.size // 3
    counts = np.zeros(ns, np.uint32)
    k = lib().orc_coupling_pairs(xs, ns, xm, nm, float(radius), counts.ctypes.data_as(ctypes.c_void_p), None, None)
    col = np.zeros(max(int(k), 1), np.uint32)
    pen = np.zeros(max(int(k), 1), np.float32)
    lib().orc_coupling_pairs(xs, ns, xm, nm, float(radius), None, col.ctypes.data_as(ctypes.c_void_p),
                             pen.ctypes.data_as(ctypes.c_void_p))
    row_ptr = np.zeros(ns + 1, np.uint32)
    row_ptr[1:] = np.cumsum(counts, dtype=np.uint64).astype(np.uint32)
    return row_ptr, col[:int(k)], pen[:int(k)]


def coupling_forces(sph_x, sph_v, mpm_x, mpm_v, row_ptr, col, penetration, stiffness, damping, friction,
                    two_way=True, sph_force=None):
    """computeCouplingForces: (solid_force [nnz,3], fluid_force [nnz,3], sph_force [ns,3])."""
    xs, vs, xm, vm = _f3(sph_x), _f3(sph_v), _f3(mpm_x), _f3(mpm_v)
    ns = xs.size // 3
    rp = np.ascontiguousarray(row_ptr, np.uint32)
    col = np.ascontiguousarray(col, np.uint32)
    pen = np.ascontiguousarray(penetration, np.float32)
    nnz = int(rp[-1])
    fs = np.zeros(3 * max(nnz, 1), np.float32)
    ff = np.zeros(3 * max(nnz, 1), np.float32)
    sf = np.zeros(3 * ns, np.float32) if sph_force is None else _f3(sph_force).copy()
    lib().orc_coupling_forces(xs, vs, ns, xm, vm, rp, np.resize(col, max(nnz, 1)), np.resize(pen, max(nnz, 1)),
                              float(stiffness), float(damping), float(friction), int(bool(two_way)), fs, ff, sf)
    return fs[:3 * nnz].reshape(-1, 3), ff[:3 * nnz].reshape(-1, 3), sf.reshape(-1, 3)


def reference_hash_neighbors(pos, cell_size, dmin, dmax, queries, radius, cap=1 << 24):
    """The reference's compiled getNeighbors (oracle/_ref): (row_ptr, neighbours)."""
    R = ref()
    if R is None:
        raise RuntimeError("oracle/_ref/libref_mpm.so is not built")
    pos, q = _f3(pos), _f3(queries)
    nq = q.size // 3
    counts = np.zeros(nq, np.uint32)
    out = np.zeros(cap, np.uint32)
    k = R.ref_hash_neighbors(pos, pos.size // 3, float(cell_size), np.asarray(dmin, np.float32),
                             np.asarray(dmax, np.float32), q, nq, float(radius), counts, out, cap)
    if k < 0:
        raise RuntimeError("neighbour lists exceed the capacity")
    row_ptr = np.zeros(nq + 1, np.uint32)
    row_ptr[1:] = np.cumsum(counts, dtype=np.uint64).astype(np.uint32)
    return row_ptr, out[:k].copy()


def reference_coupling(sph_x, sph_v, mpm_x, mpm_v, radius, stiffness, damping, friction, two_way=True,
                       cap=1 << 22):
    """The reference's compiled SPHMPMCouplingSystem: dict(sph, mpm, penetration,
    solid_force, fluid_force, sph_force)."""
    R = ref()
    if R is None:
        raise RuntimeError("oracle/_ref/libref_mpm.so is not built")
    xs, vs, xm, vm = _f3(sph_x), _f3(sph_v), _f3(mpm_x), _f3(mpm_v)
    ns, nm = xs.size // 3, xm.size // 3
    ps, pm = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
    pen = np.zeros(cap, np.float32)
    fs, ff = np.zeros(3 * cap, np.float32), np.zeros(3 * cap, np.float32)
    sf = np.zeros(3 * max(ns, 1), np.float32)
    k = R.ref_coupling(xs, vs, ns, xm, vm, nm, float(radius), float(stiffness), float(damping), float(friction),
                       int(bool(two_way)), ps, pm, pen, fs, ff, sf, cap)
    if k < 0:
        raise RuntimeError("pairs exceed the capacity")
    return dict(sph=ps[:k].copy(), mpm=pm[:k].copy(), penetration=pen[:k].copy(),
                solid_force=fs[:3 * k].reshape(-1, 3).copy(), fluid_force=ff[:3 * k].reshape(-1, 3).copy(),
                sph_force=sf[:3 * ns].reshape(-1, 3).copy())


def csr_spmv(row_ptr, col, val, x):
    """SparseMatrix::multiply: y = A x, each row summed in stored order (float32)."""
    rp = np.ascontiguousarray(row_ptr, np.uint32)
    nrows = rp.size - 1
    col = np.ascontiguousarray(col, np.uint32)
    val = np.ascontiguousarray(val, np.float32)
    y = np.zeros(max(nrows, 1), np.float32)
    lib().orc_csr_spmv(rp, np.resize(col, max(col.size, 1)), np.resize(val, max(val.size, 1)),
                       np.ascontiguousarray(x, np.float32), nrows, y)
    return y[:nrows]


def reference_sparse_matrix(nrows, ncols, rows, cols, vals, x):
    """The reference's compiled SparseMatrix<float,uint32_t> fed with
    addElement(rows[k], cols[k], vals[k]) in the given order, finalize(),
    multiply(x): (row_pointers [nrows+1], y [nrows])."""
    R = ref()
    if R is None:
        raise RuntimeError("oracle/_ref/libref_mpm.so is not built")
    rows = np.ascontiguousarray(rows, np.uint32)
    cols = np.ascontiguousarray(cols, np.uint32)
    vals = np.ascontiguousarray(vals, np.float32)
    y = np.zeros(max(nrows, 1), np.float32)
    rp = np.zeros(nrows + 1, np.uint32)
    R.ref_sparse_matrix(nrows, ncols, rows.size, rows, cols, vals, np.ascontiguousarray(x, np.float32), y, rp)
    return rp, y[:nrows]
