"""Moving rigid colliders (backend/moving_collision.py:7-87).

``MovingSolidMask`` keeps the SolidMask interface (solid, flags, clearance)
with an owner mask (0 = fluid, k = collider k), the previous mask, and the
rigid wall velocity on solid faces; ``update`` re-rasterises the colliders and
the wall velocities, ``remap`` empties cells and faces whose solidity changed.
``MovingPressureProjector`` is the masked Jacobi projector whose closed faces
carry the wall velocity (shaders/moving_boundary.glsl in pressure_walls and
pressure_gradient).
"""
import ctypes

from .. import _lib
from ..physics.collider_motion import validate_motion, velocity_rows
from .collision import SolidMask
from .device import ptr
from .pressure import SolidPressureProjector

__all__ = ["MovingSolidMask", "MovingPressureProjector"]


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
