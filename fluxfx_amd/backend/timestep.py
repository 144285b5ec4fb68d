"""Adaptive CFL timestep from device max-abs reductions.

Drop-in for AdaptiveTimestep (backend/timestep.py:6-57): ``bounds(solver)``
returns max |u|, |v|, |w|, |density|, |temperature| (one fused reduction
launch instead of the reference's 4^3-block chains; five floats read back),
``select(solver, max_dt, cfl, emission=None, sources=None)`` the report dict
with the reference's keys.

The step must keep every point's displacement within ``cfl`` cells.  Over a
step dt a point moves at most ``rate*dt + acc*dt**2`` cells with

* rate = sum over axes of (largest speed along the axis) / (cell size): the
  field's maxima, raised to the jet speed of any source that couples one in;
* acc = sum of the accelerations the step can add, in cells per second^2:
  buoyancy (lift*T + weight*rho) / h_z on the vertical axis, and the force
  caps of vorticity confinement and turbulence spread over all three axes;

``physics.timestep.choose_timestep`` then picks dt.  The projectors that
store the impulse take continuous steps under the AUTO policy, the legacy
Jacobi projector power-of-two steps.
"""
import ctypes as C
from math import isfinite

from .. import _lib
from ..physics.timestep import choose_timestep

FIELDS = 5  # u, v, w, density, temperature


class AdaptiveTimestep:
    def __init__(self, grid, device):
        self.grid = grid
        self.device = device
        self._out = device.zeros(8)

    @property
    def allocated_bytes(self):
        return 32

    def bounds(self, solver):
        fields = (*solver._velocity, solver.density, solver.temperature)
        arrays = (C.c_void_p * FIELDS)(*[f.ptr for f in fields])
        sizes = (C.c_int64 * FIELDS)(*[f.size for f in fields])
        _lib.call("fx_max_abs", FIELDS, arrays, sizes, self._out.data_ptr(), self.device.stream)
        maxima = [float(v) for v in self._out[:FIELDS].cpu().tolist()]
        if not all(isfinite(v) and v < 1e30 for v in maxima):
            raise RuntimeError("Nonfinite or excessive field values; reset simulation")
        return maxima

    @staticmethod
    def _speeds(maxima, motions):
        """Largest speed per axis: the field's, or a coupled source's jet."""
        speeds = list(maxima[:3])
        for m in motions:
            if m.coupling > 0:
                speeds = [max(s, abs(j)) for s, j in zip(speeds, m.velocity)]
        return speeds

    def _acceleration(self, settings, density, temperature):
        inv = [1.0 / h for h in self.grid.cell_size]
        spread = sum(inv)  # a force cap along any direction, in cells/s^2
        terms = [(settings.thermal_lift * temperature + settings.density_weight * density) * inv[2]]
        if settings.vorticity_strength > 0:
            terms.append(settings.vorticity_limit * spread)
        if settings.turbulence_strength > 0:
            terms.append(min(settings.turbulence_strength, settings.turbulence_limit) * spread)
        return sum(terms)

    def select(self, solver, max_dt, cfl, emission=None, sources=None):
        maxima = self.bounds(solver)
        if sources is not None:
            motions = [s.motion for s in sources]
        else:
            motions = [] if emission is None else [emission]
        speeds = self._speeds(maxima, motions)
        rate = sum(v / h for v, h in zip(speeds, self.grid.cell_size))
        s = solver.settings
        acc = self._acceleration(s, maxima[3], maxima[4])
        policy = s.timestep_policy
        if policy == "AUTO":
            policy = "CONTINUOUS" if solver.projector.uses_impulse else "QUANTIZED"
        dt = choose_timestep(max_dt, cfl, rate, acc, quantized=policy == "QUANTIZED")
        return {"dt": dt, "max_dt": max_dt, "target": cfl, "rate_bound": rate, "acceleration_rate_bound": acc,
                "estimated_courant": rate * dt + acc * dt * dt, "policy": policy, "max_abs_velocity": speeds,
                "mode": "ADAPTIVE"}

    def close(self):
        self._out = None
