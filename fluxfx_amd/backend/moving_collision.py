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
        return 5 * self.grid.cells + 4 * faces

    def update(self, colliders, dt):
        """New owners, flags, clearance and wall velocities for the colliders'
        new poses reached over dt (MovingSolidDevice.update)."""
        validate_motion(self.colliders, colliders)
        self.previous.copy_(self.solid)
        self._raster(self.solid, colliders)
        _lib.call("fx_build_flags", self._g, self.solid.data_ptr(), self.flags.data_ptr(), self.device.stream)
        _lib.call("fx_build_clearance", self._g, self.solid.data_ptr(), self.clearance.data_ptr(),
                  self._scratch.data_ptr(), self.device.stream)
        table = [v for old, new in zip(self.colliders, colliders) for row in velocity_rows(old, new, dt) for v in row]
        rows = (ctypes.c_float * len(table))(*table)
        _lib.call("fx_wall_velocity", self._g, self.solid.data_ptr(), len(colliders), rows,
                  *(f.ptr for f in self.wall_fields), self.device.stream)
        self.colliders = tuple(colliders)
        self.version += 1

    def remap(self, solver):
        """Empty the cells and rest the faces the motion covered or exposed
        (MovingSolidDevice.remap); swaps the solver's ping-pong fields."""
        g, s, o, st = self._g, self.solid.data_ptr(), self.previous.data_ptr(), self.device.stream
        for front, back in (("_front", "_back"), ("_temperature", "_temperature_back")):
            _lib.call("fx_remap_scalar", g, s, o, getattr(solver, front).ptr, getattr(solver, back).ptr, st)
            a, b = getattr(solver, front), getattr(solver, back)
            setattr(solver, front, b)
            setattr(solver, back, a)
        _lib.call("fx_remap_faces", g, s, o, *(f.ptr for f in self.wall_fields), *(f.ptr for f in solver._velocity),
                  *(f.ptr for f in solver._velocity_back), st)
        solver._velocity, solver._velocity_back = solver._velocity_back, solver._velocity
        if solver.combustion:
            c = solver.combustion
            _lib.call("fx_remap_scalar", g, s, o, c.fuel.ptr, c.spare.ptr, st)
            c.fuel, c.spare = c.spare, c.fuel
            c.flame.clear(format="FLOAT", value=(0.0,))

    def close(self):
        self.previous = self._scratch = None
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

    def compute_divergence(self, fields, target):
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

    def compute_divergence(self, fields, target):
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
