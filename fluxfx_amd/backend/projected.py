"""Closed-box smoke/fire step with pressure projection.

Drop-in for DenseProjectedSmoke (backend/projected.py:13-229): same
constructor (grid, settings, device, projector_class, colliders), reset/upload/
read, update_settings, measure_projection, step(dt, emission, sources), fault
flag.  Step order (backend/projected.py:149-202):

  1. face predictors (velocity MacCormack) with the donor bounds of their
     departure samples
  2. fused velocity transport + jet relaxation + buoyancy from the
     pre-transport density/temperature            -> provisional U, V, W
  3. optional vorticity confinement, then projection into the old velocity
     storage
  4. fused multi-channel scalar predictor (scalar MacCormack)
  5. fused scalar transport: reverse trace, correction, clamp, decay, sources
  6. fused combustion
"""
import ctypes as C

from .. import _lib
from ..physics.config import GridSpec, validate_dt
from ..physics.emission import Emission, pack_sources, MAX_SOURCES
from ..physics.interaction import validate_live_update
from ..physics.pressure import PressureSettings
from .collision import SolidMask
from .active import ActiveBricks
from .combustion import CombustionPasses
from .detail import DetailPasses
from .device import HipDevice, ptr
from .mac import DenseThermalAdvection, source_row
from .moving_collision import MovingPressureProjector, MovingSolidMask, MovingToleranceProjector
from .multigrid import MultigridPressureProjector, use_multigrid
from .pressure import PressureProjector, SolidPressureProjector
from .solid_multigrid import SolidMultigridProjector
from .tolerance import ToleranceProjector


def select_projector(grid, settings, solids):
    """Projector choice of backend/projected.py:14-43, plus the tolerance mode."""
    if settings.pressure_tolerance > 0 or settings.pressure_fixed_cycles > 0:
        return ToleranceProjector
    if solids is not None:
        return SolidMultigridProjector if use_multigrid(grid, settings) else SolidPressureProjector
    return MultigridPressureProjector if use_multigrid(grid, settings) else PressureProjector


class DenseProjectedSmoke(DenseThermalAdvection):
    """active_bricks=True runs the scalar predictor, scalar transport and
    combustion over the 8^3 bricks that can hold nonzero scalars (exact
    zero tracking, backend/active.py): the same values as the whole-grid
    passes.  Code that writes a scalar field's tensor in place outside the
    step calls invalidate_active_bricks()."""

    def __init__(self, grid=None, settings=None, device=None, projector_class=None, *, colliders=(),
                 solid_mask=None, active_bricks=False):
        self.active = None
        self.combustion = None
        self.solids = None
        self.projector = None
        self.detail = None
        self.faulted = False
        # observer(stage, fields): called inside step() with the transported
        # velocity before vorticity confinement ("advected", only when
        # confinement runs), the provisional velocity before the projection
        # ("provisional") and the transported scalars before combustion
        # ("transported"), so a caller can read the intermediate fields of a
        # step; None in normal use
        self.observer = None
        self._source_data = None
        self._multi_sources = None
        settings = settings or PressureSettings()
        grid = grid or GridSpec()
        device = device or HipDevice()
        if settings.moving_colliders and colliders:
            # moving rigid colliders: owner mask, wall velocities, masked Jacobi
            # with moving walls (backend/projected.py:17-29), or the tolerance
            # multigrid on the moving mask when a tolerance or cycle budget is set
            if solid_mask is not None:
                raise ValueError("Moving colliders are analytic shapes; a voxel mask is static")
            self.solids = MovingSolidMask(grid, colliders, device)
            tolerance = settings.pressure_tolerance > 0 or settings.pressure_fixed_cycles > 0
            projector_class = projector_class or (MovingToleranceProjector if tolerance else MovingPressureProjector)
        elif colliders or solid_mask is not None:
            self.solids = SolidMask(grid, colliders, device, mask=solid_mask)
            if settings.pressure_tolerance <= 0:
                projector_class = None
        super().__init__(grid, settings, device)
        if projector_class is None:
            projector_class = select_projector(self.grid, self.settings, self.solids)
        self.projector = projector_class(self.grid, self.settings, self.device, solids=self.solids)
        self.detail = DetailPasses(self.grid, self.device, self.solids)
        if self.settings.combustion_enabled:
            self.combustion = CombustionPasses(self.grid, self.device)
        if active_bricks:
            self.active = ActiveBricks(self.grid, self.device, self._g)
            self.projector.velocity_bound = self.active
        self.reset()

    @property
    def allocated_bytes(self):
        return (super().allocated_bytes + self.projector.allocated_bytes + self.detail.allocated_bytes +
                (self.combustion.allocated_bytes if self.combustion else 0) +
                (self.solids.allocated_bytes if self.solids else 0) +
                (self.active.allocated_bytes if self.active else 0))

    def invalidate_active_bricks(self):
        """Forget the scalar fields' live masks (rescanned before the next step)."""
        for f in (self._front, self._back, self._temperature, self._temperature_back):
            f.live = None
        if self.combustion:
            for f in (self.combustion.fuel, self.combustion.spare, self.combustion.flame):
                f.live = None

    def reset(self, seed=True):
        super().reset(seed)
        self.faulted = False
        if self.combustion:
            self.combustion.reset()
        if self.solids is not None:
            self.solids.clear_solid(self._front, self._back)
            self._front, self._back = self._back, self._front
            if self._temperature is not None:
                self.solids.clear_solid(self._temperature, self._temperature_back)
                self._temperature, self._temperature_back = self._temperature_back, self._temperature
        if self.projector is not None:
            self.projector.reset()
            self.projector.enforce_walls(self._velocity, self._velocity_back)
            self._velocity, self._velocity_back = self._velocity_back, self._velocity

    def upload_velocity(self, components):
        super().upload_velocity(components)
        if self.solids is not None:
            self.projector.enforce_walls(self._velocity, self._velocity_back)
            self._velocity, self._velocity_back = self._velocity_back, self._velocity
        self.projector.reset()

    def upload(self, values):
        super().upload(values)
        if self.solids is not None:
            self.solids.clear_solid(self._front, self._back)
            self._front, self._back = self._back, self._front
        self.projector.reset()

    def upload_temperature(self, values):
        super().upload_temperature(values)
        if self.solids is not None:
            self.solids.clear_solid(self._temperature, self._temperature_back)
            self._temperature, self._temperature_back = self._temperature_back, self._temperature
        self.projector.reset()

    def upload_fuel(self, values):
        if not self.combustion:
            raise ValueError("Enable combustion and Reset for fuel")
        self.combustion.fuel = self.device.texture(self.grid.shape, values)

    def read_fuel(self):
        if not self.combustion:
            raise ValueError("Enable combustion and Reset for fuel")
        return self.device.read(self.combustion.fuel)

    def update_settings(self, settings):
        """Apply live parameters between steps, keeping fields and pressure."""
        validate_live_update(self.settings, settings)
        self.settings = settings
        self.projector.settings = settings

    def move_colliders(self, colliders, dt):
        """Move the rigid colliders to new poses reached over dt before a step:
        new mask and wall velocities, then cells and faces whose solidity
        changed are emptied (backend/projected.py:103-115)."""
        if not self.settings.moving_colliders or not getattr(self.solids, "moving", False):
            raise ValueError("Enable Moving colliders and Reset first")
        validate_dt(dt)
        if self.faulted:
            raise RuntimeError("Reset required after a failed GPU step")
        try:
            self.solids.update(colliders, dt)
            self.solids.remap(self)
            self.projector.ready = False
        except Exception:
            self.faulted = True
            raise

    def measure_projection(self):
        return self.projector.metrics() | {"step": self.steps, "simulation_time": self.time}

    def step(self, dt=1.0 / 30.0, emission=None, sources=None):
        validate_dt(dt)
        if sources is not None:
            self._prepare_sources(sources)
        if self.faulted:
            raise RuntimeError("Reset required after a failed GPU step")
        try:
            self._step(dt, emission or Emission(self.settings.source_center), sources)
        except Exception:
            self.faulted = True
            self.projector.ready = False
            raise

    def _prepare_sources(self, sources):
        data = pack_sources(sources)
        if data != self._source_data:
            rows = [data[16 * i:16 * i + 16] for i in range(len(sources))]
            self._multi_sources = _lib.make_sources(rows, multi=True)
            self._source_data = data

    def _single_source(self, emission):
        s = self.settings
        row = source_row(s.source_center, s.source_radius, s.source_rate, s.heat_source_rate, emission.start,
                         emission.velocity, emission.coupling, s.density_mode == "TARGET",
                         s.source_profile == "SOLID", s.fuel_source_rate)
        return _lib.make_sources([row], multi=False)

    def _step(self, dt, emission, emitters=None):
        s = self.settings
        if s.turbulence_strength > 0:
            raise NotImplementedError("turbulence is not part of this build")
        table = self._multi_sources if emitters is not None else self._single_source(emission)
        solid = self.solids.solid if self.solids is not None else None
        clearance = self.solids.clearance if self.solids is not None else None
        velocity_detail = s.velocity_advection == "MACCORMACK"
        preds = bounds = None
        if velocity_detail:
            preds, bounds = self.detail.transport_velocity_bounds(self._velocity, dt)
        provisional = [self._velocity_back[0], self._velocity_back[1], self._buoyancy_scratch]
        self._advect_velocity(self._velocity, provisional, dt, maccormack=velocity_detail, predictors=preds,
                              sources=table, density=self._front, temperature=self._temperature,
                              lift=s.thermal_lift, weight=s.density_weight, solid=solid, clearance=clearance,
                              bounds=bounds)
        # the old velocity is consumed: its storage receives the projected field
        if s.vorticity_strength > 0:
            # confined field into the old velocity storage, projected into the back buffers
            # (backend/projected.py:172-175)
            if self.observer is not None:
                self.observer("advected", provisional)
            self.detail.confine(provisional, self._velocity, dt, s.vorticity_strength, s.vorticity_limit)
            if self.observer is not None:
                self.observer("provisional", list(self._velocity))
            self.projector.project(self._velocity, self._velocity_back, dt)
            self._velocity, self._velocity_back = self._velocity_back, self._velocity
        else:
            if self.observer is not None:
                self.observer("provisional", provisional)
            self.projector.project(provisional, self._velocity, dt)
        fields = [self._front, self._temperature]
        outs = [self._back, self._temperature_back]
        decays = [s.dissipation, s.cooling]
        if self.combustion:
            fields.append(self.combustion.fuel)
            outs.append(self.combustion.spare)
            decays.append(0.0)
        corrected = s.scalar_advection == "MACCORMACK"
        active = self.active
        if active is not None:
            active.transport_plan(dt, self._velocity, fields, outs, C.pointer(table))
        preds = bounds = None
        if corrected:
            preds, bounds = self.detail.scalar_predictors_bounds(fields, self._velocity, dt,
                                                                 bricks=active.predict if active else None)
        args = _lib.fx_scalar_args()
        args.g = self._g
        args.dt = dt
        args.maccormack = 1 if corrected else 0
        args.channels = len(fields)
        args.boundary = 0
        args.u, args.v, args.w = (ptr(f) for f in self._velocity)
        for k, (f, o, d) in enumerate(zip(fields, outs, decays)):
            args.field[k] = f.ptr
            args.out[k] = o.ptr
            args.pred[k] = preds[k].ptr if corrected else None
            args.decay[k] = d
            args.channel_id[k] = k
            if bounds is not None:
                args.lo[k], args.hi[k] = bounds[k][0].ptr, bounds[k][1].ptr
        args.solid = ptr(solid)
        args.clearance = ptr(clearance)
        args.sources = C.pointer(table)
        if active is not None:
            _lib.call("fx_scalar_advect_active", C.byref(args), active.write.data_ptr(), active.out_masks(outs),
                      self.device.stream)
        else:
            _lib.call("fx_scalar_advect", C.byref(args), self.device.stream)
        self._front, self._back = self._back, self._front
        self._temperature, self._temperature_back = self._temperature_back, self._temperature
        if self.combustion:
            self.combustion.fuel, self.combustion.spare = self.combustion.spare, self.combustion.fuel
        if self.observer is not None:
            self.observer("transported", [self._front, self._temperature] +
                          ([self.combustion.fuel] if self.combustion else []))
        if self.combustion:
            self.combustion.react(self, dt, active)
        self.time += dt
        self.steps += 1

    def preview_field(self, channel):
        if channel == "COLLISION":
            if self.solids is None:
                raise ValueError("Add a collider and Reset for collision display")
            return self.solids.mask
        if channel in {"FUEL", "FLAME"}:
            if not self.combustion:
                raise ValueError("Enable combustion and Reset for fuel/flame display")
            return self.combustion.fuel if channel == "FUEL" else self.combustion.flame
        return self.temperature if channel == "TEMPERATURE" else self.density

    def close(self):
        if self.active is not None:
            self.active.close()
            self.active = None
        if self.combustion:
            self.combustion.close()
        if self.projector is not None:
            self.projector.close()
        if self.detail is not None:
            self.detail.close()
        self.projector = self.detail = None
        super().close()
        if self.solids is not None:
            self.solids.close()
        self.combustion = self.solids = None
