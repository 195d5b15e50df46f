"""ctypes binding of libphysgrad_mpm.so (the C-ABI in include/physgrad_mpm.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback in the product path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphysgrad_mpm.so")

PG_OK = 0
PG_ERR_INVALID = 1
PG_ERR_HIP = 2
PG_ERR_NOMEM = 3
PG_ERR_STATE = 4

MODE_REF_COMPAT = 0
MODE_MLS_MPM = 1
MAT_FIXED_COROTATED = 0
MAT_NEO_HOOKEAN = 1
IMPL_PLAIN = 0
IMPL_TILED = 1
HOST = 0
DEVICE = 1


class PgMpmParams(ctypes.Structure):
    """Mirror of struct pg_mpm_params."""
    _fields_ = [("mode", ctypes.c_int), ("impl", ctypes.c_int),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float),
                ("ox", ctypes.c_float), ("oy", ctypes.c_float), ("oz", ctypes.c_float),
                ("dt", ctypes.c_float),
                ("gx", ctypes.c_float), ("gy", ctypes.c_float), ("gz", ctypes.c_float),
                ("interpolation_order", ctypes.c_int),
                ("material", ctypes.c_int),
                ("youngs_modulus", ctypes.c_float), ("poisson_ratio", ctypes.c_float),
                ("density", ctypes.c_float), ("particle_volume", ctypes.c_float),
                ("boundary_thickness", ctypes.c_int), ("max_particles", ctypes.c_int)]


class PgSceneDesc(ctypes.Structure):
    """Mirror of struct pg_scene_desc (one scene of a batch context)."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_float),
                ("ox", ctypes.c_float), ("oy", ctypes.c_float), ("oz", ctypes.c_float),
                ("dt", ctypes.c_float),
                ("gx", ctypes.c_float), ("gy", ctypes.c_float), ("gz", ctypes.c_float),
                ("material", ctypes.c_int),
                ("youngs_modulus", ctypes.c_float), ("poisson_ratio", ctypes.c_float),
                ("density", ctypes.c_float), ("particle_volume", ctypes.c_float),
                ("boundary_thickness", ctypes.c_int), ("max_particles", ctypes.c_int)]


class PgMaterialDesc(ctypes.Structure):
    """Mirror of struct pg_material_desc (one entry of a material table)."""
    _fields_ = [("model", ctypes.c_int), ("youngs_modulus", ctypes.c_float), ("poisson_ratio", ctypes.c_float),
                ("density", ctypes.c_float), ("viscosity", ctypes.c_float)]


MAT_FLUID = 2
BC_DIRICHLET_X = 1
BC_DIRICHLET_Y = 2
BC_DIRICHLET_Z = 4
BC_NEUMANN_X = 8
BC_NEUMANN_Y = 16
BC_NEUMANN_Z = 32
BC_STICKY = 64
BC_SLIP = 128


class PgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[pg error {code}] {msg}")
        self.code = code


_vp = ctypes.c_void_p
_SIGNATURES = {
    "pg_mpm_default_params": (None, [ctypes.POINTER(PgMpmParams)]),
    "pg_mpm_create": (ctypes.c_int, [ctypes.POINTER(PgMpmParams), ctypes.POINTER(_vp)]),
    "pg_mpm_destroy": (ctypes.c_int, [_vp]),
    "pg_mpm_last_error": (ctypes.c_char_p, [_vp]),
    "pg_mpm_set_particles": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "pg_mpm_get_particles": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "pg_mpm_num_particles": (ctypes.c_int, [_vp]),
    "pg_mpm_get_grid": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "pg_mpm_step": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "pg_mpm_get_time": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)]),
    "pg_mpm_set_time": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_longlong]),
    "pg_mpm_forward_checkpointed": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp]),
    "pg_mpm_backward": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pg_mpm_set_profiling": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pg_mpm_kernel_time": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_longlong)]),
    "pg_mpm_check_state": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _vp]),
    "pg_mpm_set_validation": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pg_mpm_set_materials": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp]),
    "pg_mpm_set_boundary": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    "pg_mpm_batch_create": (ctypes.c_int, [ctypes.POINTER(PgSceneDesc), ctypes.c_int, ctypes.POINTER(_vp)]),
    "pg_mpm_batch_set_particles": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "pg_mpm_num_scenes": (ctypes.c_int, [_vp]),
    "pg_mpm_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pg_mpm_get_launch": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "pg_mpm_set_launch": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pg_mpm_autotune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp]),
    "pg_mpm_set_work_item": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pg_mpm_set_block": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pg_mpm_get_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "pg_launch_optimized_p2g": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "pg_launch_optimized_g2p": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_float,
                                               ctypes.c_float, ctypes.c_int, _vp]),
    "pg_launch_fully_fused_g2p2g": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int,
                                                   ctypes.c_float, _vp, ctypes.c_float, ctypes.c_int, _vp]),
    "pg_launch_g2p2g": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_float, _vp,
                                       ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, _vp]),
    "pg_spatial_hash_workspace_size": (ctypes.c_size_t, [ctypes.c_uint]),
    "pg_spatial_hash_build": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_float, _vp, _vp, _vp, _vp, _vp, _vp,
                                             _vp, ctypes.c_size_t, _vp]),
    "pg_spatial_hash_build_noalloc": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_float, _vp, _vp, _vp, _vp, _vp,
                                                     _vp, _vp]),
    "pg_hash_neighbors_workspace_size": (ctypes.c_size_t, [ctypes.c_uint]),
    "pg_hash_neighbors_count": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp, _vp,
                                               ctypes.POINTER(ctypes.c_ulonglong), _vp, ctypes.c_size_t, _vp]),
    "pg_hash_neighbors_fill": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp, _vp,
                                              _vp, _vp, _vp, _vp]),
    "pg_csr_spmv": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_uint, _vp, ctypes.c_uint, _vp]),
    "pg_coupling_default_config": (None, [_vp]),
    "pg_coupling_create": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "pg_coupling_destroy": (None, [_vp]),
    "pg_coupling_last_error": (ctypes.c_char_p, [_vp]),
    "pg_coupling_find_pairs": (ctypes.c_int, [_vp, _vp, ctypes.c_uint, _vp, ctypes.c_uint,
                                              ctypes.POINTER(ctypes.c_ulonglong), _vp]),
    "pg_coupling_get_pairs": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "pg_coupling_compute_forces": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
}


class PgCouplingConfig(ctypes.Structure):
    """Mirror of struct pg_coupling_config."""
    _fields_ = [("coupling_stiffness", ctypes.c_float), ("coupling_damping", ctypes.c_float),
                ("friction_coefficient", ctypes.c_float), ("coupling_radius", ctypes.c_float),
                ("enable_two_way_coupling", ctypes.c_int), ("domain_min", ctypes.c_float * 3),
                ("domain_max", ctypes.c_float * 3)]

SPATIAL_HASH_NOALLOC_MAX = 1 << 20

_lib = None


def load():
    """Load libphysgrad_mpm.so; raises if it is missing (build with physgrad_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m physgrad_amd.build` "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_names():
    return list(_SIGNATURES)


def check(rc: int, ctx=None):
    if rc != PG_OK:
        msg = load().pg_mpm_last_error(ctx)
        raise PgError(rc, msg.decode() if msg else "unknown error")
