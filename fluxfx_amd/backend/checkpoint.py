"""Versioned checkpoints ``fluxfx-checkpoint-1`` (backend/checkpoint.py:14-69).

Same archive as the reference: a compressed .npz holding a JSON ``metadata``
string (format tag, settings, grid, time, steps, last dt) and float32 flat
x-fast arrays density, temperature, pressure, velocity0..2 (+ fuel and flame
with combustion).  Settings this build adds (tolerance solvers, learned
preconditioner) and a lifted grid cap travel under ``metadata['extensions']``,
which the reference's loader ignores; without it the defaults apply.  Writes
go to a temporary file first, so a failed write keeps the previous checkpoint.
"""
import json
from dataclasses import asdict
from math import isfinite, prod
from pathlib import Path

import numpy as np

from ..physics.config import GridSpec, validate_dt
from ..physics.pressure import EXTENSION_FIELDS, PressureSettings
from .projected import DenseProjectedSmoke

FORMAT = "fluxfx-checkpoint-1"


def save_checkpoint(solver, path):
    if solver.faulted or solver.solids is not None:
        raise ValueError("Checkpoint v1 requires a healthy solver without colliders")
    options = asdict(solver.settings)
    extensions = {k: options.pop(k) for k in EXTENSION_FIELDS if k in options}
    extensions["max_dim"] = solver.grid.max_dim
    meta = {"format": FORMAT, "settings": options,
            "grid": {"shape": list(solver.grid.shape), "extent": list(solver.grid.extent)},
            "time": solver.time, "steps": solver.steps, "last_dt": solver.projector.last_dt,
            "extensions": extensions}
    dev = solver.device
    host = lambda f: f.tensor.detach().cpu().numpy().astype(np.float32).ravel()
    arrays = {"density": host(solver.density), "temperature": host(solver.temperature),
              "pressure": host(solver.projector.pressure)}
    arrays.update({f"velocity{i}": host(f) for i, f in enumerate(solver._velocity)})
    if solver.combustion:
        arrays.update(fuel=host(solver.combustion.fuel), flame=host(solver.combustion.flame))
    if any(not np.isfinite(a).all() for a in arrays.values()):
        raise ValueError("Nonfinite checkpoint state")
    path = Path(path)
    tmp = path.with_name(path.name + ".tmp")
    try:
        with tmp.open("wb") as stream:
            np.savez_compressed(stream, metadata=json.dumps(meta), **arrays)
        tmp.replace(path)
    finally:
        if tmp.exists():
            tmp.unlink()


def load_checkpoint(path, device=None):
    with np.load(path, allow_pickle=False) as archive:
        meta = json.loads(str(archive["metadata"]))
        if meta.get("format") != FORMAT:
            raise ValueError("Unsupported checkpoint format")
        ext = dict(meta.get("extensions") or {})
        grid = GridSpec(tuple(meta["grid"]["shape"]), tuple(meta["grid"]["extent"]),
                        max_dim=int(ext.pop("max_dim", 128)))
        options = dict(meta["settings"])
        for name in ("velocity", "source_center"):
            if name in options:
                options[name] = tuple(options[name])
        options.update({k: v for k, v in ext.items() if k in EXTENSION_FIELDS})
        settings = PressureSettings(**options)
        if not isfinite(meta["time"]) or meta["time"] < 0 or type(meta["steps"]) is not int or meta["steps"] < 0:
            raise ValueError("Invalid checkpoint clock")
        if meta["last_dt"] is not None:
            validate_dt(meta["last_dt"])
        shapes = {k: grid.shape for k in ("density", "temperature", "pressure")}
        shapes.update({f"velocity{i}": s for i, s in enumerate(grid.face_shapes)})
        if settings.combustion_enabled:
            shapes.update(fuel=grid.shape, flame=grid.shape)
        arrays = {k: archive[k].copy() for k in shapes}
    for k, a in arrays.items():
        if a.shape != (prod(shapes[k]),) or not np.isfinite(a).all():
            raise ValueError("Invalid field: " + k)
        if k in {"density", "fuel", "flame"} and (a < 0).any():
            raise ValueError("Negative field: " + k)
    solver = DenseProjectedSmoke(grid, settings, device)
    try:
        solver.upload(arrays["density"])
        solver.upload_temperature(arrays["temperature"])
        solver.upload_velocity([arrays[f"velocity{i}"] for i in range(3)])
        solver.projector.pressure = solver.device.texture(grid.shape, arrays["pressure"], nonnegative=False)
        solver.projector.last_dt = meta["last_dt"]
        solver.projector.ready = False  # divergence diagnostics recomputed on the next step
        if solver.combustion:
            solver.combustion.fuel = solver.device.texture(grid.shape, arrays["fuel"])
            solver.combustion.flame = solver.device.texture(grid.shape, arrays["flame"])
        solver.time, solver.steps = meta["time"], meta["steps"]
        return solver
    except Exception:
        solver.close()
        raise


def advance_fixed(solver, steps, dt, sources=None):
    """Offline reproducible input schedule: `steps` steps of fixed dt."""
    validate_dt(dt)
    if type(steps) is not int or steps < 0:
        raise ValueError("Steps must be a nonnegative integer")
    if solver.solids is not None:
        raise ValueError("Fixed checkpoint runner v1 excludes colliders")
    for _ in range(steps):
        solver.step(dt, sources=sources)
