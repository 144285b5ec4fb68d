"""Dense ping-pong fields: density, MAC velocity, temperature.

Drop-ins for DenseAdvection (backend/dense.py:12-65), DenseMACAdvection
(backend/mac.py:12-81) and DenseThermalAdvection (backend/thermal.py:9-80):
same constructors, reset/seed, upload/read, step order and ping-pong commits.
The analytic-flow DenseAdvection.step samples its prescribed flow from a MAC
field seeded with that flow (exact inside the box for the affine base+rotation
field; face sampling clamps at the walls).
"""
import ctypes as C
from math import prod

from .. import _lib
from ..physics.config import AdvectionSettings, GridSpec, validate_dt
from ..physics.thermal import ThermalSettings
from .device import HipDevice, grid_struct, ptr

FACE_OFFSETS = ((0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0))
AXIS_MASKS = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))
VELOCITY_NAMES = ("velocityU", "velocityV", "velocityW")


def seed_params(center=(0, 0, 0), radius=1.0, scale=0.0, base=(0, 0, 0), omega=0.0):
    return (C.c_float * 9)(*center, radius, scale, *base, omega)


def source_row(center, radius, density_rate=0.0, heat_rate=0.0, start=None, velocity=(0., 0., 0.),
               coupling=0.0, target=False, solid=False, fuel_rate=0.0):
    start = center if start is None else start
    return (*center, radius, *start, coupling, *velocity, density_rate, heat_rate, float(target), float(solid),
            fuel_rate)


class DenseAdvection:
    def __init__(self, grid=None, settings=None, device=None):
        self.grid = grid or GridSpec()
        self.settings = settings or AdvectionSettings()
        self.device = device or HipDevice()
        self._g = grid_struct(self.grid)
        self._front = self.device.texture(self.grid.shape)
        self._back = self.device.texture(self.grid.shape)
        self._flow = None
        self.time = 0.0
        self.steps = 0
        self.reset()

    @property
    def density(self):
        return self._front

    def _seed_scalar(self, target, scale):
        s = self.settings
        _lib.call("fx_seed", self._g, 0, seed_params(s.source_center, s.source_radius, scale), target.ptr,
                  self.device.stream)
        target.live = None

    def reset(self, seed=True):
        self._front.clear()
        self._back.clear()
        if seed:
            self._seed_scalar(self._front, 1.0)
        self.time = 0.0
        self.steps = 0

    def upload(self, values):
        """Test/debug entry point: replaces density, clears the spare field."""
        self._front = self.device.texture(self.grid.shape, values)
        self._back.clear()
        self.time = 0.0
        self.steps = 0

    def _seeded_velocity(self, fields, base, omega):
        for a, f in enumerate(fields):
            _lib.call("fx_seed", self._g, a + 1, seed_params(base=base, omega=omega), f.ptr, self.device.stream)

    def _density_transport(self, velocity, dt, rate, decay, front, back):
        """Zero-extension SL transport with the stationary soft source
        (shaders/mac_advect_density.glsl, shaders/advect.glsl)."""
        s = self.settings
        src = _lib.make_sources([source_row(s.source_center, s.source_radius, rate)], multi=False)
        args = _lib.fx_scalar_args()
        args.g = self._g
        args.dt = dt
        args.maccormack = 0
        args.channels = 1
        args.boundary = 1
        args.u, args.v, args.w = (ptr(f) for f in velocity)
        args.field[0] = front.ptr
        args.out[0] = back.ptr
        args.decay[0] = decay
        args.channel_id[0] = 0
        args.solid = None
        args.sources = C.pointer(src)
        _lib.call("fx_scalar_advect", C.byref(args), self.device.stream)

    def step(self, dt=1.0 / 30.0):
        validate_dt(dt)
        s = self.settings
        if self._flow is None:
            self._flow = [self.device.texture(sh) for sh in self.grid.face_shapes]
        self._seeded_velocity(self._flow, s.velocity, s.angular_speed)
        self._density_transport(self._flow, dt, s.source_rate, s.dissipation, self._front, self._back)
        self._front, self._back = self._back, self._front
        self.time += dt
        self.steps += 1

    def read_density(self):
        return self.device.read(self._front, self.grid.shape)

    def close(self):
        self._front = self._back = self._flow = None


class DenseMACAdvection(DenseAdvection):
    def __init__(self, grid=None, settings=None, device=None):
        self._velocity = self._velocity_back = None
        grid = grid or GridSpec()
        device = device or HipDevice()
        self._velocity = [device.texture(s) for s in grid.face_shapes]
        self._velocity_back = [device.texture(s) for s in grid.face_shapes]
        self._empty_sources = _lib.make_sources([], multi=False)
        super().__init__(grid, settings, device)

    @property
    def allocated_bytes(self):
        return self.grid.mac_bytes

    def reset(self, seed=True):
        super().reset(seed)
        if self._velocity is None:
            return
        self._seeded_velocity(self._velocity, self.settings.velocity, self.settings.angular_speed)
        for f in self._velocity_back:
            f.clear()

    def upload_velocity(self, components):
        """Signed face data; validated and allocated before state changes."""
        if len(components) != 3:
            raise ValueError("MAC velocity requires U, V, W components")
        fields = [self.device.texture(shape, values, nonnegative=False)
                  for shape, values in zip(self.grid.face_shapes, components)]
        self._velocity = fields

    def read_velocity(self):
        return tuple(self.device.read(f, s) for f, s in zip(self._velocity, self.grid.face_shapes))

    def _advect_velocity(self, old, out, dt, *, maccormack=False, predictors=None, sources=None,
                         density=None, temperature=None, lift=0.0, weight=0.0, solid=None, clearance=None,
                         bounds=None):
        args = _lib.fx_velocity_args()
        args.g = self._g
        args.dt = dt
        args.maccormack = 1 if maccormack else 0
        args.u, args.v, args.w = (ptr(f) for f in old)
        if predictors is not None:
            args.pu, args.pv, args.pw = (ptr(f) for f in predictors)
        args.ou, args.ov, args.ow = (ptr(f) for f in out)
        args.density = ptr(density)
        args.temperature = ptr(temperature)
        args.thermal_lift = lift
        args.density_weight = weight
        args.solid = ptr(solid)
        args.clearance = ptr(clearance)
        if bounds is not None:
            (args.lu, args.lv, args.lw), (args.hu, args.hv, args.hw) = ([ptr(f) for f in b] for b in bounds)
        src = sources if sources is not None else self._empty_sources
        args.sources = C.pointer(src)
        _lib.call("fx_velocity_advect", C.byref(args), self.device.stream)

    def step(self, dt=1.0 / 30.0):
        validate_dt(dt)
        # all components gather from the same old velocity
        self._advect_velocity(self._velocity, self._velocity_back, dt)
        s = self.settings
        self._density_transport(self._velocity_back, dt, s.source_rate, s.dissipation, self._front, self._back)
        self._velocity, self._velocity_back = self._velocity_back, self._velocity
        self._front, self._back = self._back, self._front
        self.time += dt
        self.steps += 1

    def close(self):
        self._velocity = self._velocity_back = None
        super().close()


class DenseThermalAdvection(DenseMACAdvection):
    def __init__(self, grid=None, settings=None, device=None):
        self._temperature = self._temperature_back = self._buoyancy_scratch = None
        grid = grid or GridSpec()
        device = device or HipDevice()
        self._temperature = device.texture(grid.shape)
        self._temperature_back = device.texture(grid.shape)
        self._buoyancy_scratch = device.texture(grid.face_shapes[2])
        super().__init__(grid, settings or ThermalSettings(), device)

    @property
    def temperature(self):
        return self._temperature

    @property
    def allocated_bytes(self):
        return self.grid.mac_bytes + 8 * prod(self.grid.shape) + 4 * prod(self.grid.face_shapes[2])

    def reset(self, seed=True):
        super().reset(seed)
        if self._temperature is None:
            return
        for f in (self._temperature, self._temperature_back, self._buoyancy_scratch):
            f.clear()
        if seed:
            self._seed_scalar(self._temperature, self.settings.initial_temperature)

    def upload_temperature(self, values):
        self._temperature = self.device.texture(self.grid.shape, values, nonnegative=False)

    def read_temperature(self):
        return self.device.read(self._temperature, self.grid.shape)

    def step(self, dt=1.0 / 30.0):
        """SL velocity, buoyancy from the previous scalars, then SL scalars with
        the new velocity (backend/thermal.py:113-140)."""
        validate_dt(dt)
        s = self.settings
        back = [self._velocity_back[0], self._velocity_back[1], self._buoyancy_scratch]
        self._advect_velocity(self._velocity, back, dt, density=self._front, temperature=self._temperature,
                              lift=s.thermal_lift, weight=s.density_weight)
        self._density_transport(back, dt, s.source_rate, s.dissipation, self._front, self._back)
        self._density_transport(back, dt, s.heat_source_rate, s.cooling, self._temperature,
                                self._temperature_back)
        spare_w = self._velocity_back[2]
        self._velocity_back, self._velocity = self._velocity, back
        self._buoyancy_scratch = spare_w
        self._front, self._back = self._back, self._front
        self._temperature, self._temperature_back = self._temperature_back, self._temperature
        self.time += dt
        self.steps += 1

    def close(self):
        self._temperature = self._temperature_back = self._buoyancy_scratch = None
        super().close()
