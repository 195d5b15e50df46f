"""State I/O: particle snapshots (VTK, npz) and disk checkpoints with resume.

The reference's solver only prints at its output cadence
(MPMSolver<T>::outputData, src/mpm_solver.h:501-504, called from run(),
:287-321) and records checkpoints as step indices in memory
(DifferentiableMPMSystem, src/differentiable_mpm.h:449-451).  Here:

  write_vtk(path, x, point_data)     legacy binary VTK polydata of the particles
                                     (one vertex per particle, scalar or vector
                                     point fields), readable by ParaView / VisIt
  save_checkpoint(ctx, path)         the whole solver state (parameters or scene
                                     table, particle x, v, C, F and per-particle
                                     E and mass in the caller's order, material
                                     table and ids, node boundary flags, clock)
                                     in one .npz file
  load_checkpoint(path)              a new context in the saved state; stepping
                                     it continues the saved run
  run(ctx, total_time, ...)          the reference's run() loop: steps until the
                                     clock reaches total_time and writes a
                                     snapshot each time it crosses a multiple of
                                     output_frequency

A restored tiled context bins its particles afresh, so it gathers each node's
P2G contributions in a different order than the saved run would have; the
continuation therefore agrees with an uninterrupted run to float rounding,
not bit for bit (the order of particles within a tile already depends on the
scheduling of the binning atomics).  The file round trip itself is exact.
"""
from __future__ import annotations

import os
from typing import Callable, Dict, Optional

import numpy as np

CHECKPOINT_VERSION = 1

_PARAM_FIELDS = ("mode", "impl", "nx", "ny", "nz", "dx", "dy", "dz", "ox", "oy", "oz", "dt", "gx", "gy", "gz",
                 "interpolation_order", "material", "youngs_modulus", "poisson_ratio", "density", "particle_volume",
                 "boundary_thickness", "max_particles")
_SCENE_FIELDS = ("nx", "ny", "nz", "dx", "ox", "oy", "oz", "dt", "gx", "gy", "gz", "material", "youngs_modulus",
                 "poisson_ratio", "density", "particle_volume", "boundary_thickness", "max_particles")


# ----------------------------------------------------------------------------- VTK

def write_vtk(path: str, x, point_data: Optional[Dict[str, np.ndarray]] = None, title: str = "physgrad_amd particles"):
    """Legacy binary VTK polydata: points x [n, 3], one vertex cell per point,
    and point fields ([n] scalars or [n, 3] vectors).  Big-endian float32, as
    the legacy format requires."""
    x = np.asarray(x, np.float32).reshape(-1, 3)
    n = x.shape[0]
    fields = dict(point_data or {})
    with open(path, "wb") as f:
        f.write(b"# vtk DataFile Version 3.0\n")
        f.write(title.replace("\n", " ")[:255].encode() + b"\n")
        f.write(b"BINARY\nDATASET POLYDATA\n")
        f.write(f"POINTS {n} float\n".encode())
        f.write(x.astype(">f4").tobytes())
        f.write(f"\nVERTICES {n} {2 * n}\n".encode())
        cells = np.empty((n, 2), ">i4")
        cells[:, 0] = 1
        cells[:, 1] = np.arange(n, dtype=np.int64)
        f.write(cells.tobytes())
        if fields:
            f.write(f"\nPOINT_DATA {n}\n".encode())
        for name, a in fields.items():
            a = np.asarray(a, np.float32)
            if a.ndim == 1 or (a.ndim == 2 and a.shape[1] == 1):
                f.write(f"SCALARS {name} float 1\nLOOKUP_TABLE default\n".encode())
                f.write(a.reshape(n).astype(">f4").tobytes())
            elif a.ndim == 2 and a.shape[1] == 3:
                f.write(f"VECTORS {name} float\n".encode())
                f.write(a.astype(">f4").tobytes())
            else:
                raise ValueError(f"point field {name!r}: shape {a.shape} is neither [n] nor [n, 3]")
            f.write(b"\n")


def read_vtk_points(path: str):
    """Points and point fields of a file written by write_vtk (for tests and tools)."""
    data = open(path, "rb").read()
    pos = data.index(b"POINTS ")
    line_end = data.index(b"\n", pos)
    n = int(data[pos:line_end].split()[1])
    x = np.frombuffer(data, ">f4", 3 * n, line_end + 1).reshape(n, 3).astype(np.float32)
    fields = {}
    for kind in (b"SCALARS ", b"VECTORS "):
        start = 0
        while True:
            k = data.find(b"\n" + kind, start)
            if k < 0:
                break
            hdr_end = data.index(b"\n", k + 1)
            name = data[k + 1:hdr_end].split()[1].decode()
            off = hdr_end + 1
            if kind == b"SCALARS ":
                off = data.index(b"\n", off) + 1   # LOOKUP_TABLE line
                fields[name] = np.frombuffer(data, ">f4", n, off).astype(np.float32)
            else:
                fields[name] = np.frombuffer(data, ">f4", 3 * n, off).reshape(n, 3).astype(np.float32)
            start = off
    return x, fields


def write_snapshot(ctx, path: str):
    """Snapshot of a context's particles: VTK (.vtk) with v, |v| and det F, or
    npz (any other suffix) with x, v, C, F and the clock."""
    p = {k: t.detach().cpu().numpy() for k, t in ctx.get_particles().items()}
    t, steps = ctx.time()
    if path.endswith(".vtk"):
        J = np.linalg.det(p["F"].astype(np.float64)).astype(np.float32)
        write_vtk(path, p["x"], {"velocity": p["v"], "speed": np.linalg.norm(p["v"], axis=1), "J": J},
                  title=f"physgrad_amd particles t={t:.6g} step={steps}")
    else:
        np.savez(path, time=t, steps=steps, **p)


# ----------------------------------------------------------------------------- checkpoints

def pack_checkpoint(kind: str, config: Dict[str, np.ndarray], particles: Dict[str, np.ndarray], time: float,
                    steps: int, counts=None, extra: Optional[Dict[str, np.ndarray]] = None) -> Dict[str, np.ndarray]:
    """The arrays of a checkpoint file: kind ("context" or "batch"), the
    parameter struct or scene table fields, particle arrays in the caller's
    order, the clock, for a batch the per-scene particle counts, and extra
    arrays stored as given (material table and ids, boundary flags)."""
    out = {"version": np.int32(CHECKPOINT_VERSION), "kind": np.array(kind), "time": np.float64(time),
           "steps": np.int64(steps)}
    for k, v in (extra or {}).items():
        if v is not None:
            out["x_" + k] = np.asarray(v)
    for k, v in config.items():
        out["cfg_" + k] = np.asarray(v)
    for k, v in particles.items():
        if v is not None:
            out["p_" + k] = np.ascontiguousarray(v, np.float32)
    if counts is not None:
        out["counts"] = np.asarray(counts, np.int64)
    return out


def unpack_checkpoint(arrays) -> dict:
    if int(arrays["version"]) != CHECKPOINT_VERSION:
        raise ValueError(f"checkpoint version {int(arrays['version'])}, expected {CHECKPOINT_VERSION}")
    out = {"kind": str(arrays["kind"]), "time": float(arrays["time"]), "steps": int(arrays["steps"]),
           "config": {k[4:]: arrays[k] for k in arrays.files if k.startswith("cfg_")},
           "particles": {k[2:]: arrays[k] for k in arrays.files if k.startswith("p_")},
           "counts": arrays["counts"].tolist() if "counts" in arrays.files else None,
           "extra": {k[2:]: arrays[k] for k in arrays.files if k.startswith("x_")}}
    return out


_MAT_FIELDS = ("model", "youngs_modulus", "poisson_ratio", "density", "viscosity")


def _extras(ctx) -> Dict[str, np.ndarray]:
    ex = {}
    table = getattr(ctx, "material_table", None)
    if table:
        for f in _MAT_FIELDS:
            ex["mat_" + f] = np.array([t[f] for t in table])
        ids = getattr(ctx, "material_ids", None)
        ex["mat_ids"] = None if ids is None else ids.detach().cpu().numpy().astype(np.uint8)
    flags = getattr(ctx, "boundary_flags", None)
    if flags is not None:
        ex["bc_flags"] = flags.detach().cpu().numpy().astype(np.uint32)
    return ex


def _restore_extras(ctx, ex: Dict[str, np.ndarray]):
    if "mat_model" in ex:
        n = len(ex["mat_model"])
        table = [{f: ex["mat_" + f][m].item() for f in _MAT_FIELDS} for m in range(n)]
        ctx.set_materials(table, ex.get("mat_ids"))
    if "bc_flags" in ex:
        ctx.set_boundary(ex["bc_flags"])


def save_checkpoint(ctx, path: str):
    """Writes the whole state of a context (single scene or batch) to `path` (.npz)."""
    from .solver import MPMBatch
    p = {k: t.detach().cpu().numpy() for k, t in ctx.get_particles().items()}
    E = getattr(ctx, "particle_E", None)
    m = getattr(ctx, "particle_mass", None)
    p["E"] = None if E is None else E.detach().cpu().numpy()
    p["mass"] = None if m is None else m.detach().cpu().numpy()
    t, steps = ctx.time()
    if isinstance(ctx, MPMBatch):
        cfg = {f: np.array([getattr(d, f) for d in ctx.scenes]) for f in _SCENE_FIELDS}
        arrays = pack_checkpoint("batch", cfg, p, t, steps, counts=ctx.counts, extra=_extras(ctx))
    else:
        cfg = {f: np.array(getattr(ctx.params, f)) for f in _PARAM_FIELDS}
        arrays = pack_checkpoint("context", cfg, p, t, steps, extra=_extras(ctx))
    tmp = path + ".tmp.npz"
    np.savez(tmp, **arrays)
    os.replace(tmp, path)


def load_checkpoint(path: str):
    """A new context (MPMContext or MPMBatch) holding the state saved at `path`."""
    from .solver import MPMBatch, MPMContext, PgSceneDesc, default_params
    with np.load(path, allow_pickle=False) as f:
        ck = unpack_checkpoint(f)
    cfg, p = ck["config"], ck["particles"]
    if ck["kind"] == "batch":
        descs = []
        for s in range(len(cfg["nx"])):
            d = PgSceneDesc()
            for k in _SCENE_FIELDS:
                setattr(d, k, type(getattr(d, k))(cfg[k][s].item()))
            descs.append(d)
        ctx = MPMBatch(descs)
        ctx.set_particles(ck["counts"], p["x"], p["v"], p.get("C"), p.get("F"), p.get("mass"), p.get("E"))
    else:
        prm = default_params()
        for k in _PARAM_FIELDS:
            setattr(prm, k, type(getattr(prm, k))(cfg[k].item()))
        ctx = MPMContext(prm)
        ctx.set_particles(p["x"], p["v"], p.get("C"), p.get("F"), p.get("mass"), p.get("E"))
    _restore_extras(ctx, ck["extra"])
    ctx.set_time(ck["time"], ck["steps"])
    return ctx


# ----------------------------------------------------------------------------- run loop

def run(ctx, total_time: float, output_frequency: float = 0.0, output_dir: Optional[str] = None,
        fmt: str = "vtk", steps_per_call: int = 1, callback: Optional[Callable] = None) -> int:
    """Steps `ctx` until its clock reaches total_time (MPMSolver<T>::run,
    src/mpm_solver.h:287-321).  Each time the clock crosses a multiple of
    output_frequency a snapshot is written to output_dir (frame_NNNNNN.vtk or
    .npz) and callback(ctx, frame) is called.  Returns the number of frames."""
    t, steps = ctx.time()
    dt = float(ctx.params.dt) if ctx.params is not None else float(ctx.scenes[0].dt)
    if output_dir:
        os.makedirs(output_dir, exist_ok=True)
    frames = 0
    while t < total_time - 0.5 * dt:
        k = max(1, min(steps_per_call, int(round((total_time - t) / dt))))
        ctx.step(k)
        t_new, steps = ctx.time()
        # the clock sums a float32 dt, so multiples of the frequency are matched
        # with a relative slack of 1e-6
        if output_frequency > 0 and int(t_new / output_frequency + 1e-6) > int(t / output_frequency + 1e-6):
            if output_dir:
                write_snapshot(ctx, os.path.join(output_dir, f"frame_{frames:06d}.{fmt}"))
            if callback is not None:
                callback(ctx, frames)
            frames += 1
        t = t_new
    return frames
