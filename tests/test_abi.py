"""C-ABI shape tests (no GPU compute): the library loads, exports every entry
point declared in include/*.h, and rejects bad arguments before touching HIP."""
import ctypes
import glob
import os
import re

import pytest

from conftest import ROOT


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(pg_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "pg_mpm_create" in names and "pg_mpm_step" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    from physgrad_amd import _lib
    L = _lib.load()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.exported_names()) <= set(declared_symbols())


def test_params_struct_layout():
    from physgrad_amd import _lib
    p = _lib.PgMpmParams()
    _lib.load().pg_mpm_default_params(ctypes.byref(p))
    assert (p.nx, p.ny, p.nz) == (64, 64, 64)
    assert p.interpolation_order == 2 and abs(p.gy + 9.81) < 1e-6 and abs(p.dt - 1e-4) < 1e-10
    assert ctypes.sizeof(p) == 23 * 4


@pytest.mark.parametrize("field,value", [("nx", 0), ("dt", 0.0), ("dx", -1.0), ("interpolation_order", 5),
                                         ("max_particles", -3), ("mode", 7)])
def test_create_rejects_bad_params(field, value):
    from physgrad_amd import _lib
    L = _lib.load()
    p = _lib.PgMpmParams()
    L.pg_mpm_default_params(ctypes.byref(p))
    setattr(p, field, value)
    h = ctypes.c_void_p()
    rc = L.pg_mpm_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == _lib.PG_ERR_INVALID
    assert not h.value
    assert L.pg_mpm_last_error(None)


def test_mls_rejects_anisotropic_cell():
    from physgrad_amd import _lib
    L = _lib.load()
    p = _lib.PgMpmParams()
    L.pg_mpm_default_params(ctypes.byref(p))
    p.mode = _lib.MODE_MLS_MPM
    p.dy = p.dx * 2
    h = ctypes.c_void_p()
    assert L.pg_mpm_create(ctypes.byref(p), ctypes.byref(h)) == _lib.PG_ERR_INVALID
    assert b"cubic" in L.pg_mpm_last_error(None)


def test_null_context_is_an_error():
    from physgrad_amd import _lib
    L = _lib.load()
    assert L.pg_mpm_step(None, 1, None) == _lib.PG_ERR_INVALID
    assert L.pg_mpm_num_particles(None) == -1


def test_spatial_hash_argument_checks():
    from physgrad_amd import _lib
    L = _lib.load()
    res = (ctypes.c_uint * 3)(4, 4, 4)
    lo = (ctypes.c_float * 3)(0, 0, 0)
    assert L.pg_spatial_hash_build(None, 0, 0.0, lo, res, None, None, None, None, None, 0, None) == \
        _lib.PG_ERR_INVALID
    assert L.pg_spatial_hash_workspace_size(1000) >= 3 * 4000


@pytest.mark.parametrize("side", [257, 264])
def test_tiled_path_rejects_grids_over_256_nodes_per_axis(side):
    """The tiled path refuses more than 64 tiles (256 nodes) per axis before
    touching HIP (single-scene and batch contexts)."""
    from physgrad_amd import _lib
    L = _lib.load()
    p = _lib.PgMpmParams()
    L.pg_mpm_default_params(ctypes.byref(p))
    p.mode, p.impl = _lib.MODE_MLS_MPM, _lib.IMPL_TILED
    p.nx, p.ny, p.nz = 64, side, 64
    p.dx = p.dy = p.dz = 1.0 / side
    h = ctypes.c_void_p()
    assert L.pg_mpm_create(ctypes.byref(p), ctypes.byref(h)) == _lib.PG_ERR_INVALID
    assert b"256 nodes per axis" in L.pg_mpm_last_error(None)
    d = _lib.PgSceneDesc()
    d.nx, d.ny, d.nz, d.dx, d.dt = side, 32, 32, 1.0 / side, 1e-4
    d.poisson_ratio, d.particle_volume, d.max_particles = 0.3, 1e-6, 16
    arr = (_lib.PgSceneDesc * 1)(d)
    assert L.pg_mpm_batch_create(arr, 1, ctypes.byref(h)) == _lib.PG_ERR_INVALID
    assert b"256 nodes per axis" in L.pg_mpm_last_error(None)
