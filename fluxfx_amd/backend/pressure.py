"""Weighted-Jacobi projection with closed free-slip walls and static solids.

Drop-in for PressureProjector (backend/pressure.py:9-120): same constructor,
attributes (pressure, spare, before, after, bounded, ready, last_dt,
last_iterations), method names, impulse/legacy storage rules, iteration
scaling and errors.  Walls are fused into the divergence and gradient passes;
``bounded`` is only materialised when ``enforce_walls`` is asked for it.
"""
from math import ceil, prod

from .. import _lib
from ..physics.config import validate_dt
from ..physics.pressure import PressureSettings, divergence_stats
from .device import HipDevice, grid_struct, ptr
from .engine import MGLevels


class PressureProjector:
    solver_name = "JACOBI"

    # ActiveBricks of the owning step: project() then writes its velocity bound
    velocity_bound = None

    def __init__(self, grid, settings=None, device=None, *, solids=None):
        self.grid = grid
        self.settings = settings or PressureSettings()
        self.device = device or HipDevice()
        self.solids = solids
        self.inv_cell = tuple(1 / h for h in grid.cell_size)
        self._g = grid_struct(grid)
        self._bounded = None
        self.pressure = self.device.texture(grid.shape)
        self.spare = self.device.texture(grid.shape)
        self.before = self.device.texture(grid.shape)
        self.after = self.device.texture(grid.shape)
        self.ready = False
        self.last_dt = None
        self.last_iterations = 0
        op = (_lib.FX_OP_FLAGS, solids.flags) if solids is not None else (_lib.FX_OP_BOX, None)
        self._sweeper = MGLevels(self.device, [grid.shape], grid.extent, _lib.FX_F32, [op], pre=0, post=1,
                                 coarse_sweeps=1, own_level0=False)

    # -- helpers -----------------------------------------------------------------
    @property
    def flags_ptr(self):
        return ptr(self.solids.flags) if self.solids is not None else None

    @property
    def bounded(self):
        if self._bounded is None:
            self._bounded = [self.device.texture(s) for s in self.grid.face_shapes]
        return self._bounded

    @bounded.setter
    def bounded(self, value):
        self._bounded = value

    @property
    def uses_impulse(self):
        return (self.settings.pressure_warm_start == "IMPULSE" or
                (self.settings.pressure_warm_start == "AUTO" and len(getattr(self, "levels", ())) > 1))

    @property
    def allocated_bytes(self):
        faces = sum(prod(s) for s in self.grid.face_shapes) if self._bounded is not None else 0
        return 4 * (4 * prod(self.grid.shape) + faces)

    def reset(self):
        self.ready = False
        self.last_dt = None
        self.last_iterations = 0
        for f in (self.pressure, self.spare, self.before, self.after, *(self._bounded or ())):
            f.clear(format="FLOAT", value=(0.0,))

    # -- passes --------------------------------------------------------------------
    def enforce_walls(self, fields, outputs):
        _lib.call("fx_enforce_walls", self._g, self.flags_ptr, *(ptr(f) for f in fields),
                  *(ptr(f) for f in outputs), self.device.stream)

    def compute_divergence(self, fields, target, bound=None):
        """bound: an ActiveBricks whose velocity bound the same pass writes."""
        if bound is not None:
            _lib.call("fx_divergence_bound", self._g, self.flags_ptr, *(ptr(f) for f in fields), None, None, None,
                      ptr(target), bound.partials.data_ptr(), bound.vmax.data_ptr(), self.device.stream)
            bound.bound_ready = True
            return
        _lib.call("fx_divergence", self._g, self.flags_ptr, *(ptr(f) for f in fields), ptr(target),
                  self.device.stream)

    def apply_gradient(self, fields, outputs, scale):
        _lib.call("fx_gradient", self._g, self.flags_ptr, ptr(self.pressure), *(ptr(f) for f in fields),
                  *(ptr(f) for f in outputs), float(scale), self.device.stream)

    def project(self, fields, outputs, dt):
        validate_dt(dt)
        if any(o is f for o in outputs for f in fields):
            raise ValueError("Advection input/output textures must be distinct")
        self.ready = False
        previous_dt, self.last_dt = self.last_dt, None
        self.compute_divergence(fields, self.before)
        impulse = self.uses_impulse
        solve_dt = 1.0 if impulse else dt
        previous_solve_dt = (1.0 if previous_dt is not None else None) if impulse else previous_dt
        if impulse and previous_dt is not None and abs(dt / previous_dt - 1.0) > 1e-6:
            # the impulse guess scales with the physical step
            _lib.call("fx_scale", self.pressure.ptr, self.spare.ptr, self.pressure.size,
                      float(dt / previous_dt), self.device.stream)
            self.pressure, self.spare = self.spare, self.pressure
        self.solve_pressure(solve_dt, previous_solve_dt)
        self.apply_gradient(fields, outputs, solve_dt)
        # the projected velocity's divergence (and, for the active-brick scalar
        # passes, its bound max |U|, |V|, |W| from the same loads)
        self.compute_divergence(outputs, self.after, bound=self.velocity_bound)
        self.last_dt = dt
        self.ready = True

    def _jacobi(self, count, dt, relaxation):
        sw = self._sweeper
        sw.bind_level0(self.pressure.tensor, self.spare.tensor, self.before.tensor)
        sw.cfg.omega = relaxation
        sw.smooth(0, count, 1.0 / dt)
        if sw.tensor(0, "p") is not self.pressure.tensor:
            self.pressure, self.spare = self.spare, self.pressure

    def solve_pressure(self, dt, previous_dt):
        cold = previous_dt != dt
        if cold:
            self.pressure.clear(format="FLOAT", value=(0.0,))
        scaled = ceil(self.settings.pressure_iterations * max(1.0, (max(self.grid.shape) / 64) ** 2))
        multiplier = self.settings.pressure_cold_multiplier if cold else 1
        self.last_iterations = min(4096, scaled * multiplier)
        self._jacobi(self.last_iterations, dt, self.settings.pressure_relaxation)

    def metrics(self):
        """Explicit synchronous readback."""
        if not self.ready:
            raise RuntimeError("Take a simulation step before measuring projection")
        before = self.device.read(self.before)
        after = self.device.read(self.after)
        return divergence_stats(before, after) | {
            "solver": self.solver_name,
            "pressure_storage": "IMPULSE" if self.uses_impulse else "LEGACY",
            "iterations": self.last_iterations,
            "base_iterations_at_64": self.settings.pressure_iterations,
            "relaxation": self.settings.pressure_relaxation,
            "boundary": "closed_free_slip", "units": "1/s",
            "before_stage": "after_wall_conditioning"}

    def close(self):
        self.ready = False
        self.last_dt = None
        self._bounded = None
        self.pressure = self.spare = self.before = self.after = None
        self._sweeper = None


class SolidPressureProjector(PressureProjector):
    """Masked Jacobi, impulse storage, zero start every step
    (backend/collision.py:75-88)."""
    solver_name = "SOLID_JACOBI"

    @property
    def uses_impulse(self):
        return True

    def solve_pressure(self, dt, previous_dt):
        self.pressure.clear(format="FLOAT", value=(0.0,))
        super().solve_pressure(dt, previous_dt)

    def metrics(self):
        return super().metrics() | {"solver": "MOVING_JACOBI" if self.settings.moving_colliders else "SOLID_JACOBI",
                                    "boundary": "static_voxel_free_slip", "warm_start": False}
