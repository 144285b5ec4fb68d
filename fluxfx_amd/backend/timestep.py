"""Adaptive CFL timestep from device max-abs reductions
(backend/timestep.py:6-57).  One fused reduction launch replaces the
reference's 4^3-block chains; five floats are read back."""
import ctypes as C
from math import isfinite

from .. import _lib
from ..physics.timestep import choose_timestep


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
        arrays = (C.c_void_p * 5)(*[f.ptr for f in fields])
        sizes = (C.c_int64 * 5)(*[f.size for f in fields])
        _lib.call("fx_max_abs", 5, arrays, sizes, self._out.data_ptr(), self.device.stream)
        maxima = [float(v) for v in self._out[:5].cpu().tolist()]
        if any(not isfinite(v) or v >= 1e30 for v in maxima):
            raise RuntimeError("Nonfinite or excessive field values; reset simulation")
        return maxima

    def select(self, solver, max_dt, cfl, emission=None, sources=None):
        u, v, w, density, temperature = self.bounds(solver)
        motions = [s.motion for s in sources] if sources is not None else ([emission] if emission is not None else [])
        for motion in motions:
            if motion.coupling > 0:
                u, v, w = (max(cur, abs(t)) for cur, t in zip((u, v, w), motion.velocity))
        h = self.grid.cell_size
        rate = sum(speed / spacing for speed, spacing in zip((u, v, w), h))
        s = solver.settings
        acceleration = (s.thermal_lift * temperature + s.density_weight * density) / h[2]
        if s.vorticity_strength > 0:
            acceleration += s.vorticity_limit * sum(1 / sp for sp in h)
        if s.turbulence_strength > 0:
            acceleration += min(s.turbulence_strength, s.turbulence_limit) * sum(1 / sp for sp in h)
        quantized = (s.timestep_policy == "QUANTIZED" or
                     (s.timestep_policy == "AUTO" and not solver.projector.uses_impulse))
        dt = choose_timestep(max_dt, cfl, rate, acceleration, quantized=quantized)
        return {"dt": dt, "max_dt": max_dt, "target": cfl, "rate_bound": rate,
                "acceleration_rate_bound": acceleration, "estimated_courant": rate * dt + acceleration * dt * dt,
                "policy": "QUANTIZED" if quantized else "CONTINUOUS", "max_abs_velocity": [u, v, w],
                "mode": "ADAPTIVE"}

    def close(self):
        self._out = None
