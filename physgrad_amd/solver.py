"""Host-side mirror of the reference MPM solver interfaces over the C-ABI.

`MPMContext` is the thin typed wrapper of one pg_mpm_ctx (device tensors in,
device tensors out, work on torch's current HIP stream).  `MPMSolver` keeps the
names and argument meaning of the reference CPU solver
(src/material_point_method.h:430-506: constructor(grid_dims, cell_size,
origin, params), addParticles, step, getParticles, getGrid, getCurrentTime,
getStepCount) so code written against it reads the same.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import PgMpmParams, PgSceneDesc, check, load


def _stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dptr(t: Optional[torch.Tensor]):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dev_f32(a, n_cols: int, name: str) -> Optional[torch.Tensor]:
    if a is None:
        return None
    t = torch.as_tensor(a)
    if t.dtype != torch.float32:
        t = t.to(torch.float32)
    t = t.to("cuda").contiguous()
    if n_cols > 1 and t.numel() % n_cols:
        raise ValueError(f"{name}: size {t.numel()} is not a multiple of {n_cols}")
    return t.reshape(-1, n_cols) if n_cols > 1 else t.reshape(-1)


def default_params() -> PgMpmParams:
    p = PgMpmParams()
    load().pg_mpm_default_params(ctypes.byref(p))
    return p


def make_params(**kw) -> PgMpmParams:
    """pg_mpm_params from keyword arguments; unspecified fields keep the defaults.

    Accepts dims=(nx,ny,nz), cell=(dx,dy,dz) or dx=float, origin=(ox,oy,oz),
    gravity=(gx,gy,gz) as conveniences besides the struct's own field names."""
    p = default_params()
    dims = kw.pop("dims", None)
    if dims is not None:
        p.nx, p.ny, p.nz = (int(d) for d in dims)
    cell = kw.pop("cell", None)
    if cell is not None:
        p.dx, p.dy, p.dz = (float(c) for c in cell)
    if "dx" in kw and not isinstance(kw["dx"], (tuple, list)):
        d = float(kw.pop("dx"))
        p.dx = p.dy = p.dz = d
    origin = kw.pop("origin", None)
    if origin is not None:
        p.ox, p.oy, p.oz = (float(o) for o in origin)
    grav = kw.pop("gravity", None)
    if grav is not None:
        p.gx, p.gy, p.gz = (float(g) for g in grav)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(f"unknown pg_mpm_params field {k!r}")
        setattr(p, k, v)
    return p


class MPMContext:
    """One pg_mpm_ctx. All data arguments are device tensors (or anything
    torch.as_tensor accepts, copied to the current HIP device)."""

    def __init__(self, params: PgMpmParams):
        self._L = load()
        self.params = params
        h = ctypes.c_void_p()
        check(self._L.pg_mpm_create(ctypes.byref(params), ctypes.byref(h)))
        self._h = h
        self._nodes = params.nx * params.ny * params.nz

    def close(self):
        if getattr(self, "_h", None):
            self._L.pg_mpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def num_particles(self) -> int:
        return int(self._L.pg_mpm_num_particles(self._h))

    @property
    def num_nodes(self) -> int:
        return self._nodes

    def set_graph(self, enable: bool = True):
        """Replay step() from captured hipGraphs (tiled MLS-MPM)."""
        check(self._L.pg_mpm_set_graph(self._h, 1 if enable else 0), self._h)

    @property
    def launch(self) -> int:
        """Workgroups of the persistent tile kernels."""
        v = ctypes.c_int()
        check(self._L.pg_mpm_get_launch(self._h, ctypes.byref(v)), self._h)
        return v.value

    @launch.setter
    def launch(self, workgroups: int):
        check(self._L.pg_mpm_set_launch(self._h, int(workgroups)), self._h)

    @property
    def work_item(self) -> int:
        """Particles per work item of the tile kernels (a tile's particles are split into
        work items of at most this many)."""
        return self.tuning()["work_item"]

    @work_item.setter
    def work_item(self, particles: int):
        check(self._L.pg_mpm_set_work_item(self._h, int(particles)), self._h)

    @property
    def block_threads(self) -> int:
        """Threads per tile workgroup (256 or 512)."""
        return self.tuning()["block_threads"]

    @block_threads.setter
    def block_threads(self, threads: int):
        check(self._L.pg_mpm_set_block(self._h, int(threads)), self._h)

    def tuning(self) -> dict:
        """Launch shape of the tiled path: tile workgroups, work-item size, tile edge
        (cells), threads per tile workgroup."""
        v = [ctypes.c_int() for _ in range(4)]
        check(self._L.pg_mpm_get_tuning(self._h, *[ctypes.byref(t) for t in v]), self._h)
        return {"tile_workgroups": v[0].value, "work_item": v[1].value, "tile_edge": v[2].value,
                "block_threads": v[3].value}

    def autotune(self, trial_steps: int = 10, use_cache: bool = True) -> int:
        """Pick the tile-kernel launch size, work-item size and workgroup threads by
        timing; returns the workgroup count (tuning() gives all)."""
        check(self._L.pg_mpm_autotune(self._h, int(trial_steps), 1 if use_cache else 0, _stream_ptr()), self._h)
        return self.launch

    def set_particles(self, x, v, C=None, F=None, mass=None, E=None):
        x = _dev_f32(x, 3, "x")
        v = _dev_f32(v, 3, "v")
        n = x.shape[0]
        C = _dev_f32(C, 9, "C")
        F = _dev_f32(F, 9, "F")
        mass = _dev_f32(mass, 1, "mass")
        E = _dev_f32(E, 1, "E")
        for name, t, cols in (("v", v, 3), ("C", C, 9), ("F", F, 9), ("mass", mass, 1), ("E", E, 1)):
            if t is not None and t.numel() != n * cols:
                raise ValueError(f"{name} has {t.numel()} values, expected {n * cols}")
        check(self._L.pg_mpm_set_particles(self._h, n, _dptr(x), _dptr(v), _dptr(C), _dptr(F), _dptr(mass),
                                           _dptr(E), _lib.DEVICE, _stream_ptr()), self._h)
        # per-particle constants travel with the particles unchanged: kept in the
        # caller's order for checkpoints
        self.particle_E = None if E is None else E.clone()
        self.particle_mass = None if mass is None else mass.clone()
        self.rollouts = getattr(self, "rollouts", 0)
        self.material_table, self.material_ids = None, None   # the library resets them too

    def get_particles(self):
        n = self.num_particles
        out = {k: torch.empty((n, c), dtype=torch.float32, device="cuda")
               for k, c in (("x", 3), ("v", 3), ("C", 9), ("F", 9))}
        check(self._L.pg_mpm_get_particles(self._h, _dptr(out["x"]), _dptr(out["v"]), _dptr(out["C"]),
                                           _dptr(out["F"]), _lib.DEVICE, _stream_ptr()), self._h)
        out["C"] = out["C"].reshape(n, 3, 3)
        out["F"] = out["F"].reshape(n, 3, 3)
        return out

    def get_grid(self, momentum: bool = False):
        G = self.num_nodes
        m = torch.empty(G, dtype=torch.float32, device="cuda")
        vel = torch.empty((G, 3), dtype=torch.float32, device="cuda")
        mom = torch.empty((G, 3), dtype=torch.float32, device="cuda") if momentum else None
        check(self._L.pg_mpm_get_grid(self._h, _dptr(m), _dptr(mom), _dptr(vel), _lib.DEVICE, _stream_ptr()),
              self._h)
        out = {"mass": m, "velocity": vel}
        if momentum:
            out["momentum"] = mom
        return out

    def step(self, nsteps: int = 1):
        check(self._L.pg_mpm_step(self._h, int(nsteps), _stream_ptr()), self._h)

    def set_materials(self, table, ids=None):
        """Per-particle materials: table = list of dicts (model, youngs_modulus,
        poisson_ratio, density, viscosity) or PgMaterialDesc; ids [n] material
        index per particle (caller order).  table None restores the scenes'
        materials.  Call after set_particles (which resets them)."""
        if table is None or len(table) == 0:
            check(self._L.pg_mpm_set_materials(self._h, None, 0, None, _lib.DEVICE, _stream_ptr()), self._h)
            self.material_table, self.material_ids = None, None
            return
        descs = []
        for t in table:
            if isinstance(t, _lib.PgMaterialDesc):
                descs.append(t)
                continue
            d = _lib.PgMaterialDesc()
            d.model = int(t.get("model", _lib.MAT_NEO_HOOKEAN))
            d.youngs_modulus = float(t.get("youngs_modulus", 0.0))
            d.poisson_ratio = float(t.get("poisson_ratio", 0.0))
            d.density = float(t.get("density", 0.0))
            d.viscosity = float(t.get("viscosity", 0.0))
            descs.append(d)
        arr = (_lib.PgMaterialDesc * len(descs))(*descs)
        idt = None
        if ids is not None:
            idt = torch.as_tensor(ids).to(torch.uint8).to("cuda").contiguous()
        check(self._L.pg_mpm_set_materials(self._h, arr, len(descs), _dptr(idt), _lib.DEVICE, _stream_ptr()),
              self._h)
        self.material_table = [{f: getattr(d, f) for f, _ in _lib.PgMaterialDesc._fields_} for d in descs]
        self.material_ids = None if idt is None else idt.clone()

    def set_boundary(self, flags=None):
        """Per-node boundary flags ([num_nodes] uint32 BoundaryType bits, node
        index (k ny + j) nx + i); None clears them."""
        if flags is None:
            check(self._L.pg_mpm_set_boundary(self._h, None, _lib.DEVICE, _stream_ptr()), self._h)
            self.boundary_flags = None
            return
        t = torch.as_tensor(flags).to(torch.int64).to(torch.int32).to("cuda").contiguous().reshape(-1)
        if t.numel() != self.num_nodes:
            raise ValueError(f"{t.numel()} flags for {self.num_nodes} grid nodes")
        check(self._L.pg_mpm_set_boundary(self._h, _dptr(t), _lib.DEVICE, _stream_ptr()), self._h)
        self.boundary_flags = t.clone()

    def check_state(self):
        """Validates the particle state; returns (first_nonfinite, first_outside),
        -1 for none.  Raises PgError (PG_ERR_STATE) for a non-finite state or a
        binning error, with the first bad particle named in the message."""
        a, b = ctypes.c_int(), ctypes.c_int()
        check(self._L.pg_mpm_check_state(self._h, ctypes.byref(a), ctypes.byref(b), _stream_ptr()), self._h)
        return a.value, b.value

    def set_validation(self, on: bool = True):
        """When on, every step() ends with check_state() (and synchronises)."""
        check(self._L.pg_mpm_set_validation(self._h, 1 if on else 0), self._h)

    def forward_checkpointed(self, nsteps: int, checkpoint_every: int = 32):
        """Advance nsteps steps keeping a checkpoint every `checkpoint_every` steps."""
        check(self._L.pg_mpm_forward_checkpointed(self._h, int(nsteps), int(checkpoint_every), _stream_ptr()),
              self._h)
        self.rollouts = getattr(self, "rollouts", 0) + 1

    def backward(self, ax=None, av=None, aC=None, aF=None):
        """Adjoints of the initial state (and per-particle E) from adjoints of the
        final state of the last forward_checkpointed.  Returns dict x, v, C, F, E."""
        n = self.num_particles
        ins = [None if t is None else _dev_f32(t, c, name).contiguous()
               for t, c, name in ((ax, 3, "ax"), (av, 3, "av"), (aC, 9, "aC"), (aF, 9, "aF"))]
        out = {k: torch.empty((n, c), dtype=torch.float32, device="cuda")
               for k, c in (("x", 3), ("v", 3), ("C", 9), ("F", 9), ("E", 1))}
        check(self._L.pg_mpm_backward(self._h, *[_dptr(t) for t in ins], _dptr(out["x"]), _dptr(out["v"]),
                                      _dptr(out["C"]), _dptr(out["F"]), _dptr(out["E"]), _stream_ptr()), self._h)
        out["C"] = out["C"].reshape(n, 3, 3)
        out["F"] = out["F"].reshape(n, 3, 3)
        out["E"] = out["E"].reshape(n)
        return out

    KERNELS = {"tile_p2g": 0, "tile_g2p2g": 1, "tile_g2p": 2, "grid": 3, "bin": 4, "plain_p2g": 5,
               "plain_g2p": 6, "clear": 7, "ref_step": 8, "adj_g2p": 9, "adj_grid": 10, "adj_p2g": 11,
               "adj_fused": 12}

    def set_profiling(self, on: bool):
        check(self._L.pg_mpm_set_profiling(self._h, 1 if on else 0), self._h)

    def kernel_times(self):
        """{kernel: (total_ms, launches)} of the launches timed since set_profiling(True)."""
        out = {}
        for name, k in self.KERNELS.items():
            t = ctypes.c_double()
            c = ctypes.c_longlong()
            check(self._L.pg_mpm_kernel_time(self._h, k, ctypes.byref(t), ctypes.byref(c)), self._h)
            if c.value:
                out[name] = (t.value, c.value)
        return out

    def time(self):
        t = ctypes.c_double()
        s = ctypes.c_longlong()
        check(self._L.pg_mpm_get_time(self._h, ctypes.byref(t), ctypes.byref(s)), self._h)
        return t.value, s.value

    def set_time(self, time: float, steps: int):
        """Sets the clock (resume from a checkpoint; set_particles resets it)."""
        check(self._L.pg_mpm_set_time(self._h, float(time), int(steps)), self._h)


def scene_desc(dims, dx, origin=(0.0, 0.0, 0.0), dt=1e-4, gravity=(0.0, -9.8, 0.0), material=0,
               youngs_modulus=1e4, poisson_ratio=0.2, density=1.0, particle_volume=1e-6, boundary_thickness=3,
               max_particles=1 << 16) -> PgSceneDesc:
    d = PgSceneDesc()
    d.nx, d.ny, d.nz = (int(v) for v in dims)
    d.dx = float(dx)
    d.ox, d.oy, d.oz = (float(v) for v in origin)
    d.dt = float(dt)
    d.gx, d.gy, d.gz = (float(v) for v in gravity)
    d.material = int(material)
    d.youngs_modulus = float(youngs_modulus)
    d.poisson_ratio = float(poisson_ratio)
    d.density = float(density)
    d.particle_volume = float(particle_volume)
    d.boundary_thickness = int(boundary_thickness)
    d.max_particles = int(max_particles)
    return d


class MPMBatch(MPMContext):
    """Many independent MLS-MPM scenes advanced by the same launches (one
    pg_mpm_ctx built by pg_mpm_batch_create).  Particles of all scenes are
    concatenated in scene order; get_particles returns them in that order and
    get_grid returns the concatenated grids."""

    def __init__(self, scenes: Sequence[PgSceneDesc]):
        self._L = load()
        self.scenes = list(scenes)
        arr = (PgSceneDesc * len(self.scenes))(*self.scenes)
        h = ctypes.c_void_p()
        check(self._L.pg_mpm_batch_create(arr, len(self.scenes), ctypes.byref(h)))
        self._h = h
        self.params = None
        self._nodes = sum(d.nx * d.ny * d.nz for d in self.scenes)
        self.counts = [0] * len(self.scenes)

    @property
    def num_scenes(self) -> int:
        return int(self._L.pg_mpm_num_scenes(self._h))

    def set_particles(self, counts, x, v, C=None, F=None, mass=None, E=None):
        counts = [int(c) for c in counts]
        if len(counts) != len(self.scenes):
            raise ValueError(f"{len(counts)} counts for {len(self.scenes)} scenes")
        x = _dev_f32(x, 3, "x")
        v = _dev_f32(v, 3, "v")
        n = x.shape[0]
        if n != sum(counts):
            raise ValueError(f"{n} particles but counts sum to {sum(counts)}")
        C = _dev_f32(C, 9, "C")
        F = _dev_f32(F, 9, "F")
        mass = _dev_f32(mass, 1, "mass")
        E = _dev_f32(E, 1, "E")
        for name, t, cols in (("v", v, 3), ("C", C, 9), ("F", F, 9), ("mass", mass, 1), ("E", E, 1)):
            if t is not None and t.numel() != n * cols:
                raise ValueError(f"{name} has {t.numel()} values, expected {n * cols}")
        carr = (ctypes.c_int * len(counts))(*counts)
        check(self._L.pg_mpm_batch_set_particles(self._h, carr, _dptr(x), _dptr(v), _dptr(C), _dptr(F),
                                                 _dptr(mass), _dptr(E), _lib.DEVICE,
                                                 _stream_ptr()), self._h)
        self.counts = counts
        self.particle_E = None if E is None else E.clone()
        self.particle_mass = None if mass is None else mass.clone()
        self.material_table, self.material_ids = None, None


@dataclass
class SimulationParams:
    """Reference MPMSolver<T,64>::SimulationParams (src/material_point_method.h:440-448)."""
    timestep: float = 1e-4
    gravity: Sequence[float] = (0.0, -9.81, 0.0)
    cfl_factor: float = 0.4
    use_affine_pic: bool = True
    use_flip: bool = True
    flip_ratio: float = 0.95
    interpolation_order: int = 2


class MPMSolver:
    """Drop-in for the reference CPU solver MPMSolver<float,64>
    (src/material_point_method.h:430-506) running the ref-compat kernels.

    As in the reference, use_affine_pic/use_flip/flip_ratio/cfl_factor and the
    addParticles material arguments do not change the step."""

    def __init__(self, grid_dims, cell_size, origin, params: SimulationParams = SimulationParams(),
                 max_particles: int = 1 << 20):
        self.params = params
        p = make_params(mode=_lib.MODE_REF_COMPAT, dims=grid_dims, cell=cell_size, origin=origin,
                        dt=float(params.timestep), gravity=params.gravity,
                        interpolation_order=int(params.interpolation_order), max_particles=int(max_particles))
        self._ctx = MPMContext(p)
        self._x = None
        self._v = None
        self._m = None

    def addParticles(self, positions, velocities, masses, volumes=None, young_modulus=1e6, poisson_ratio=0.3,
                     density=1000.0):
        """Appends particles (reference :454-477); F starts at I."""
        pos = torch.as_tensor(positions, dtype=torch.float32).reshape(-1, 3)
        vel = torch.as_tensor(velocities, dtype=torch.float32).reshape(-1, 3)
        ms = torch.as_tensor(masses, dtype=torch.float32).reshape(-1)
        if self._ctx.num_particles > 0:
            cur = self._ctx.get_particles()
            pos = torch.cat([cur["x"].cpu(), pos])
            vel = torch.cat([cur["v"].cpu(), vel])
            ms = torch.cat([self._m.cpu(), ms])
        self._m = ms.to("cuda")
        self._ctx.set_particles(pos, vel, mass=ms)

    def step(self):
        self._ctx.step(1)

    def getParticles(self):
        return self._ctx.get_particles()

    def getGrid(self):
        return self._ctx.get_grid(momentum=True)

    def getCurrentTime(self) -> float:
        return self._ctx.time()[0]

    def getStepCount(self) -> int:
        return self._ctx.time()[1]
