"""Spatial hash neighbour lists and SPH-MPM coupling on the GPU.

Host-side mirrors of the reference's classes over the C-ABI:
  SpatialHashGrid      SpatialHashGrid<T, uint32_t> (src/sparse_data_structures.h:44-294):
                       buildHashTable (:155-190) -> pg_spatial_hash_build,
                       getNeighbors (:193-228) -> pg_hash_neighbors_count/fill
                       (one point, or many at once as a CSR list), getStats (:231-269).
  SPHMPMCoupling       SPHMPMCouplingSystem<T> (src/sph_mpm_coupling.h:391-579):
                       findCouplingPairs (:454-486), computeCouplingForces
                       (:488-545), getNumCouplingPairs, getTotalCouplingEnergy,
                       getCouplingForces -> pg_coupling_*.
Arrays are torch tensors on the GPU (numpy arrays are copied there); results
are device tensors.  Everything runs in the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import PgError, load


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(a, dtype=torch.float32, cols: Optional[int] = None, name: str = "array") -> torch.Tensor:
    t = torch.as_tensor(a)
    t = t.to(device="cuda", dtype=dtype).contiguous()
    if cols is not None:
        if t.numel() % cols:
            raise ValueError(f"{name}: {t.numel()} values are not a multiple of {cols}")
        t = t.reshape(-1, cols)
    return t


def _p(t: Optional[torch.Tensor]):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else None


def _check(rc: int):
    if rc != _lib.PG_OK:
        msg = load().pg_mpm_last_error(None)
        raise PgError(rc, msg.decode() if msg else "unknown error")


def _u32(t: torch.Tensor) -> torch.Tensor:
    """int32 storage of uint32 values read as int64 (exact for all uint32)."""
    return t.to(torch.int64) & 0xFFFFFFFF


class SpatialHashGrid:
    """SpatialHashGrid<float, uint32_t>: cell size, domain box, Morton hash of
    the truncated cell coordinates modulo the cell count, stable sort of the
    particle indices by cell."""

    def __init__(self, cell_size: float, domain_min: Sequence[float], domain_max: Sequence[float],
                 max_particles: int = 10_000_000):
        self.cell_size = float(cell_size)
        self.domain_min = np.asarray(domain_min, np.float32)
        self.domain_max = np.asarray(domain_max, np.float32)
        self.max_particles = int(max_particles)
        self.res = np.ceil((self.domain_max - self.domain_min) / np.float32(self.cell_size)).astype(np.uint32)
        self.total_cells = int(np.prod(self.res.astype(np.uint64)))
        self.cell_start = torch.zeros(self.total_cells, dtype=torch.int32, device="cuda")
        self.cell_count = torch.zeros(self.total_cells, dtype=torch.int32, device="cuda")
        self.particle_indices = torch.zeros(0, dtype=torch.int32, device="cuda")
        self.cell_ids = torch.zeros(0, dtype=torch.int32, device="cuda")
        self.num_particles = 0

    def buildHashTable(self, positions, num_particles: Optional[int] = None):
        pos = _dev(positions, cols=3, name="positions")
        n = pos.shape[0] if num_particles is None else int(num_particles)
        if n > pos.shape[0] or n > self.max_particles:
            raise ValueError(f"{n} particles: more than given or than max_particles")
        L = load()
        self.particle_indices = torch.empty(n, dtype=torch.int32, device="cuda")
        self.cell_ids = torch.empty(n, dtype=torch.int32, device="cuda")
        ws = torch.empty(int(L.pg_spatial_hash_workspace_size(n)), dtype=torch.uint8, device="cuda")
        lo = (ctypes.c_float * 3)(*self.domain_min.tolist())
        res = (ctypes.c_uint * 3)(*self.res.tolist())
        _check(L.pg_spatial_hash_build(_p(pos), n, self.cell_size, lo, res, _p(self.cell_ids),
                                       _p(self.particle_indices), _p(self.cell_start), _p(self.cell_count),
                                       _p(ws), ws.numel(), _stream()))
        self.num_particles = n

    def neighbors(self, queries, radius: float):
        """getNeighbors for every row of queries [nq, 3] at once: (row_ptr
        [nq+1], neighbours) as int64 tensors; row q lists what getNeighbors
        returns for query q, in its order."""
        q = _dev(queries, cols=3, name="queries")
        nq = q.shape[0]
        L = load()
        lo = (ctypes.c_float * 3)(*self.domain_min.tolist())
        res = (ctypes.c_uint * 3)(*self.res.tolist())
        row_ptr = torch.empty(nq + 1, dtype=torch.int32, device="cuda")
        ws = torch.empty(max(int(L.pg_hash_neighbors_workspace_size(nq)), 1), dtype=torch.uint8, device="cuda")
        total = ctypes.c_ulonglong(0)
        _check(L.pg_hash_neighbors_count(_p(q), nq, float(radius), self.cell_size, lo, res, _p(self.cell_count),
                                         _p(row_ptr), ctypes.byref(total), _p(ws), ws.numel(), _stream()))
        out = torch.empty(max(int(total.value), 1), dtype=torch.int32, device="cuda")
        _check(L.pg_hash_neighbors_fill(_p(q), nq, float(radius), self.cell_size, lo, res,
                                        _p(self.particle_indices), _p(self.cell_start), _p(self.cell_count),
                                        _p(row_ptr), _p(out), _stream()))
        return _u32(row_ptr), _u32(out[:int(total.value)])

    def getNeighbors(self, x: float, y: float, z: float, radius: float):
        """Neighbours of one point, as the reference's list."""
        _, nb = self.neighbors([[x, y, z]], radius)
        return nb.cpu().tolist()

    def getStats(self):
        """HashTableStats (sparse_data_structures.h:231-269) from the device counts."""
        ct = _u32(self.cell_count)
        occ = int((ct > 0).sum())
        mx = int(ct.max()) if ct.numel() else 0
        tot = int(ct.sum())
        mem = (2 * self.total_cells + 2 * self.max_particles) * 4
        return {"average_particles_per_cell": tot / max(1, occ), "load_factor": occ / self.total_cells,
                "max_particles_per_cell": mx, "occupied_cells": occ, "memory_usage_mb": mem / (1024.0 * 1024.0)}

    def getCellSize(self):
        return self.cell_size

    def getTotalCells(self):
        return self.total_cells

    def getParticleIndices(self):
        return _u32(self.particle_indices)

    def getCellStartIndices(self):
        return _u32(self.cell_start)

    def getCellParticleCounts(self):
        return _u32(self.cell_count)


class SPHMPMCoupling:
    """SPHMPMCouplingSystem's pair search and contact forces (CouplingConfig
    fields as keyword arguments; domain_min/domain_max bound the binning of
    the MPM particles, e.g. the MPM grid's box)."""

    def __init__(self, coupling_radius: float = 0.15, coupling_stiffness: float = 1e5,
                 coupling_damping: float = 1e3, friction_coefficient: float = 0.3,
                 enable_two_way_coupling: bool = True, domain_min=(0.0, 0.0, 0.0), domain_max=(1.0, 1.0, 1.0)):
        L = load()
        cfg = _lib.PgCouplingConfig()
        L.pg_coupling_default_config(ctypes.byref(cfg))
        cfg.coupling_radius = float(coupling_radius)
        cfg.coupling_stiffness = float(coupling_stiffness)
        cfg.coupling_damping = float(coupling_damping)
        cfg.friction_coefficient = float(friction_coefficient)
        cfg.enable_two_way_coupling = int(bool(enable_two_way_coupling))
        for a in range(3):
            cfg.domain_min[a] = float(domain_min[a])
            cfg.domain_max[a] = float(domain_max[a])
        self.config = cfg
        h = ctypes.c_void_p()
        _check(L.pg_coupling_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._L = L
        self.num_pairs = 0
        self.ns = 0
        self._pen = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.pg_coupling_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int):
        if rc != _lib.PG_OK:
            msg = self._L.pg_coupling_last_error(self._h)
            raise PgError(rc, msg.decode() if msg else "unknown error")

    def findCouplingPairs(self, sph_positions, mpm_positions) -> int:
        xs = _dev(sph_positions, cols=3, name="sph_positions")
        xm = _dev(mpm_positions, cols=3, name="mpm_positions")
        n = ctypes.c_ulonglong(0)
        self._chk(self._L.pg_coupling_find_pairs(self._h, _p(xs), xs.shape[0], _p(xm), xm.shape[0],
                                                 ctypes.byref(n), _stream()))
        self.num_pairs = int(n.value)
        self.ns = xs.shape[0]
        self._forces = None
        return self.num_pairs

    def getCouplingPairs(self):
        """(row_ptr [ns+1], mpm index [npairs]) as int64, penetration [npairs] float32:
        pairs in the reference's order, SPH index major."""
        row_ptr = torch.empty(self.ns + 1, dtype=torch.int32, device="cuda")
        col = torch.empty(max(self.num_pairs, 1), dtype=torch.int32, device="cuda")
        pen = torch.empty(max(self.num_pairs, 1), dtype=torch.float32, device="cuda")
        self._chk(self._L.pg_coupling_get_pairs(self._h, _p(row_ptr), _p(col), _p(pen), _stream()))
        k = self.num_pairs
        return _u32(row_ptr), _u32(col[:k]), pen[:k]

    def computeCouplingForces(self, sph_positions, sph_velocities, mpm_positions, mpm_velocities,
                              sph_force: Optional[torch.Tensor] = None):
        """Forces of the pairs of the last findCouplingPairs.  sph_force [ns, 3]
        (float32 CUDA tensor, updated in place with the fluid forces when
        two-way coupling is on; zeros when None).  Returns dict(solid_force,
        fluid_force [npairs, 3], sph_force)."""
        xs = _dev(sph_positions, cols=3)
        vs = _dev(sph_velocities, cols=3)
        xm = _dev(mpm_positions, cols=3)
        vm = _dev(mpm_velocities, cols=3)
        if sph_force is None:
            sph_force = torch.zeros((xs.shape[0], 3), dtype=torch.float32, device="cuda")
        if sph_force.dtype != torch.float32 or not sph_force.is_cuda or not sph_force.is_contiguous():
            raise ValueError("sph_force must be a contiguous float32 CUDA tensor")
        k = max(self.num_pairs, 1)
        fs = torch.empty((k, 3), dtype=torch.float32, device="cuda")
        ff = torch.empty((k, 3), dtype=torch.float32, device="cuda")
        self._chk(self._L.pg_coupling_compute_forces(self._h, _p(xs), _p(vs), _p(xm), _p(vm), _p(sph_force),
                                                     _p(fs), _p(ff), _stream()))
        self._forces = {"solid_force": fs[:self.num_pairs], "fluid_force": ff[:self.num_pairs],
                        "sph_force": sph_force}
        return self._forces

    def getCouplingForces(self):
        """The pairs and the forces of the last computeCouplingForces, field by
        field as the reference's CouplingForce (sph_mpm_coupling.h:378-386):
        fluid_particle_id, solid_particle_id, penetration_depth, contact_area,
        solid_force, fluid_force."""
        row_ptr, col, pen = self.getCouplingPairs()
        rows = torch.repeat_interleave(torch.arange(self.ns, device="cuda"), row_ptr[1:] - row_ptr[:-1])
        f = getattr(self, "_forces", None)
        return {"fluid_particle_id": rows, "solid_particle_id": col, "penetration_depth": pen,
                "contact_area": self.contactArea(),
                "solid_force": None if f is None else f["solid_force"],
                "fluid_force": None if f is None else f["fluid_force"]}

    def getNumCouplingPairs(self) -> int:
        return self.num_pairs

    def getTotalCouplingEnergy(self) -> float:
        """sum over pairs of 0.5 stiffness penetration^2 (sph_mpm_coupling.h:562-574),
        summed in float64 on the device."""
        _, _, pen = self.getCouplingPairs()
        return float(0.5 * self.config.coupling_stiffness * (pen.double() ** 2).sum())

    def contactArea(self) -> float:
        """contact_area of every pair: pi radius^2 (sph_mpm_coupling.h:479)."""
        r = np.float32(self.config.coupling_radius)
        return float(np.float32(3.14159265359) * r * r)
