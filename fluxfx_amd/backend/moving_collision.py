"""Moving rigid colliders (backend/moving_collision.py:7-87).

``MovingSolidMask`` keeps the SolidMask interface (solid, flags, clearance)
with an owner mask (0 = fluid, k = collider k), the previous mask, and the
rigid wall velocity on solid faces; ``update`` re-rasterises the colliders and
the wall velocities, ``remap`` empties cells and faces whose solidity changed.
``MovingPressureProjector`` is the masked Jacobi projector whose closed faces
carry the wall velocity (shaders/moving_boundary.glsl in pressure_walls and
pressure_gradient).  ``MovingToleranceProjector`` (no reference counterpart:
the reference projects moving colliders with masked Jacobi only) runs the
tolerance multigrid solvers (FP32, BF16/FP16, FPCG, learned) on the moving
mask: its rediscretised levels are rebuilt on the device whenever the mask
changes, and the coarsest level is solved by FP64 conjugate gradients, as
there is no fixed pseudo-inverse to keep.
"""
import ctypes
import math

import numpy as np

from .. import _lib
from ..physics.collider_motion import validate_motion, velocity_rows
from .collision import SolidMask
from .device import ptr
from .pressure import SolidPressureProjector
from .tolerance import ToleranceProjector

__all__ = ["MovingSolidMask", "MovingPressureProjector", "MovingToleranceProjector"]


class MovingSolidMask(SolidMask):
    moving = True

    def __init__(self, grid, colliders, device=None):
        validate_motion(colliders, colliders)
        if not 1 <= len(colliders) <= 8:
            raise ValueError("Use one to eight moving colliders")
        super().__init__(grid, colliders, device)
        self.colliders = tuple(colliders)
        n = grid.cells
        self._raster(self.solid, self.colliders)
        self.previous = self.device.zeros(n, "uint8")
        self.previous.copy_(self.solid)
        self._scratch = self.device.zeros(n, "uint8")
        self._scratch2 = self.device.zeros(n, "uint8")
        self.wall_fields = [self.device.texture(s, nonnegative=False) for s in grid.face_shapes]
        # bumped by every update(): projectors with mask-derived operators rebuild them
        self.version = 0

    def _raster(self, mask, colliders):
        mask.zero_()
        for k, c in enumerate(colliders, start=1):
            rows = (ctypes.c_float * 12)(*[float(v) for r in c.inverse_rows for v in r])
            _lib.call("fx_owner_raster", self._g, 0 if c.shape == "SPHERE" else 1, rows, k, mask.data_ptr(),
                      self.device.stream)

    @property
    def allocated_bytes(self):
        faces = sum(s[0] * s[1] * s[2] for s in self.grid.face_shapes)
        return 6 * self.grid.cells + 4 * faces

    def cell_box(self, colliders):
        """Half-open cell box {x0, y0, z0, x1, y1, z1} holding every cell whose
        centre the colliders can cover: each unit sphere/box mapped forward by
        its inverse rows' inverse gives a centre and half-extents per axis
        (row norms of the forward matrix: 2-norm for the sphere, 1-norm for the
        box), widened by a cell for the rasteriser's rounding."""
        lo = [math.inf] * 3
        hi = [-math.inf] * 3
        for c in colliders:
            rows = np.array(c.inverse_rows, dtype=np.float64)
            fwd = np.linalg.inv(rows[:, :3])
            centre = -fwd @ rows[:, 3]
            half = np.linalg.norm(fwd, axis=1) if c.shape == "SPHERE" else np.abs(fwd).sum(axis=1)
            for a in range(3):
                h = self.grid.extent[a] / self.grid.shape[a]
                lo[a] = min(lo[a], math.floor((centre[a] - half[a]) / h - 0.5) - 1)
                hi[a] = max(hi[a], math.ceil((centre[a] + half[a]) / h - 0.5) + 2)
        n = self.grid.shape
        return [max(0, min(int(lo[a]), n[a])) for a in range(3)] + [max(0, min(int(hi[a]), n[a])) for a in range(3)]

    def update(self, colliders, dt):
        """New owners, flags, clearance and wall velocities for the colliders'
        new poses reached over dt (MovingSolidDevice.update), over the box of
        cells the old and new poses can cover (fx_moving_update_box)."""
        validate_motion(self.colliders, colliders)
        self.previous.copy_(self.solid)
        old_box, new_box = self.cell_box(self.colliders), self.cell_box(colliders)
        box = [min(a, b) for a, b in zip(old_box[:3], new_box[:3])] + \
              [max(a, b) for a, b in zip(old_box[3:], new_box[3:])]
        self._box = (ctypes.c_int32 * 6)(*box)
        kinds = (ctypes.c_int32 * len(colliders))(*[0 if c.shape == "SPHERE" else 1 for c in colliders])
        inv = [float(v) for c in colliders for r in c.inverse_rows for v in r]
        table = [v for old, new in zip(self.colliders, colliders) for row in velocity_rows(old, new, dt) for v in row]
        _lib.call("fx_moving_update_box", self._g, self._box, len(colliders), kinds, (ctypes.c_float * len(inv))(*inv),
                  (ctypes.c_float * len(table))(*table), self.solid.data_ptr(), self.flags.data_ptr(),
                  self.clearance.data_ptr(), self._scratch.data_ptr(), self._scratch2.data_ptr(),
                  *(f.ptr for f in self.wall_fields), self.device.stream)
        self.colliders = tuple(colliders)
        self.version += 1

    def remap(self, solver):
        """Empty the cells and rest the faces the motion covered or exposed
        (MovingSolidDevice.remap), in place over the last update's box."""
        fuel = solver.combustion.fuel.ptr if solver.combustion else None
        _lib.call("fx_remap_box", self._g, self._box, self.solid.data_ptr(), self.previous.data_ptr(),
                  *(f.ptr for f in self.wall_fields), *(f.ptr for f in solver._velocity), solver._front.ptr,
                  solver._temperature.ptr, fuel, self.device.stream)
        if solver.combustion:
            solver.combustion.flame.clear(format="FLOAT", value=(0.0,))

    def close(self):
        self.previous = self._scratch = self._scratch2 = None
        self.wall_fields = []
        self.colliders = ()
        super().close()


class MovingPressureProjector(SolidPressureProjector):
    """Masked Jacobi, impulse storage, zero start, closed faces at the wall velocity."""
    solver_name = "MOVING_JACOBI"

    def _walls(self):
        return [ptr(f) for f in self.solids.wall_fields]

    def enforce_walls(self, fields, outputs):
        _lib.call("fx_enforce_walls_moving", self._g, self.flags_ptr, *(ptr(f) for f in fields), *self._walls(),
                  *(ptr(f) for f in outputs), self.device.stream)

    def compute_divergence(self, fields, target, bound=None):
        if bound is not None:
            _lib.call("fx_divergence_bound", self._g, self.flags_ptr, *(ptr(f) for f in fields), *self._walls(),
                      ptr(target), bound.partials.data_ptr(), bound.vmax.data_ptr(), self.device.stream)
            bound.bound_ready = True
            return
        _lib.call("fx_divergence_moving", self._g, self.flags_ptr, *(ptr(f) for f in fields), *self._walls(),
                  ptr(target), self.device.stream)

    def apply_gradient(self, fields, outputs, scale):
        _lib.call("fx_gradient_moving", self._g, self.flags_ptr, ptr(self.pressure), *(ptr(f) for f in fields),
                  *self._walls(), *(ptr(f) for f in outputs), float(scale), self.device.stream)

    def metrics(self):
        return super().metrics() | {"solver": "MOVING_JACOBI", "boundary": "moving_voxel_free_slip"}


class MovingToleranceProjector(ToleranceProjector):
    """Tolerance multigrid with moving walls: closed faces carry the wall
    velocity in the walls, divergence and gradient passes, and the level
    operators follow the mask (fx_mg_rebuild_masked) before each solve after
    the colliders moved."""
    solver_name = "MOVING_TOLERANCE_MULTIGRID"

    def __init__(self, grid, settings=None, device=None, *, solids=None):
        if solids is None or not getattr(solids, "moving", False):
            raise ValueError("MovingToleranceProjector needs a MovingSolidMask")
        self._ops_version = None
        self._coarse_solid = None
        super().__init__(grid, settings, device, solids=solids)
        self.refresh_operators()

    def _masked_operators(self, solids):
        """Device arrays of every level's operator and activity bytes, filled
        by refresh_operators(); no coarsest pseudo-inverse (None)."""
        torch = self.device.torch
        ops = [(_lib.FX_OP_FLAGS, solids.flags)]
        active = [None]
        for level, shape in enumerate(self.shapes[1:], start=1):
            n = shape[0] * shape[1] * shape[2]
            if 4 ** level <= 1023:
                t = self.device.zeros(n, "int32")
                ops.append((_lib.FX_OP_PACKED, t))
            else:
                t = self.device.zeros(4 * n, "float32")
                ops.append((_lib.FX_OP_EDGES, t))
            self._edges.append(t)
            active.append(self.device.zeros(n, "uint8"))
        self._active_arrays = active
        nx, ny, nz = self.shapes[1]
        self._coarse_solid = torch.zeros(nx * ny * nz, dtype=torch.uint8, device=self.device.device)
        return ops, None

    def refresh_operators(self):
        """Rebuild levels 1.. from the current mask's open-face flags."""
        e = self.grid.extent
        _lib.call("fx_mg_rebuild_masked", ctypes.byref(self.engine.cfg), float(e[0]), float(e[1]), float(e[2]),
                  self._coarse_solid.data_ptr(), self.device.stream)
        self._ops_version = self.solids.version

    def _walls(self):
        return [ptr(f) for f in self.solids.wall_fields]

    def enforce_walls(self, fields, outputs):
        _lib.call("fx_enforce_walls_moving", self._g, self.flags_ptr, *(ptr(f) for f in fields), *self._walls(),
                  *(ptr(f) for f in outputs), self.device.stream)

    def compute_divergence(self, fields, target, bound=None):
        if bound is not None:
            _lib.call("fx_divergence_bound", self._g, self.flags_ptr, *(ptr(f) for f in fields), *self._walls(),
                      ptr(target), bound.partials.data_ptr(), bound.vmax.data_ptr(), self.device.stream)
            bound.bound_ready = True
            return
        _lib.call("fx_divergence_moving", self._g, self.flags_ptr, *(ptr(f) for f in fields), *self._walls(),
                  ptr(target), self.device.stream)

    def solve_pressure(self, dt, previous_dt):
        if self._ops_version != self.solids.version:
            self.refresh_operators()
        super().solve_pressure(dt, previous_dt)

    def apply_gradient(self, fields, outputs, scale):
        inc = self._increment_fields()
        _lib.call("fx_gradient_split_moving", self._g, self.flags_ptr, self.pressure.ptr, inc["d"].ptr,
                  *(ptr(f) for f in fields), *self._walls(), *(ptr(f) for f in outputs), self.spare.ptr,
                  float(scale), self.device.stream)
        self.pressure, self.spare = self.spare, self.pressure
        _lib.call("fx_project_mean", self._g, self.flags_ptr, self.pressure.ptr, self.ws.data_ptr(), self.device.stream)

    def _network_mask(self, level):
        if level == 0:
            return self.solids.solid.data_ptr()
        return self._coarse_solid.data_ptr()

    def metrics(self):
        return super().metrics() | {"boundary": "moving_voxel_free_slip"}
