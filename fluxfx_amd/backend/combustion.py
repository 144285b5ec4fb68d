"""Fuel storage and the fused reaction pass (backend/combustion.py:4-33).

One kernel replaces the reference's burn + three apply dispatches; the
ping-pong swaps of fuel, density and temperature are the same.
"""
from math import prod

from .. import _lib


class CombustionPasses:
    def __init__(self, grid, device):
        self.grid, self.device = grid, device
        self.fuel = device.texture(grid.shape)
        self.spare = device.texture(grid.shape)
        self.flame = device.texture(grid.shape)  # consumed fuel per second in the last step

    @property
    def allocated_bytes(self):
        return 12 * prod(self.grid.shape)

    def reset(self):
        for f in (self.fuel, self.spare, self.flame):
            f.clear()

    def react(self, solver, dt, active=None):
        """active: the solver's ActiveBricks (backend/active.py), or None for
        the whole-grid pass."""
        s = solver.settings
        rates = (float(dt), float(s.ignition_temperature), float(s.burn_rate), float(s.heat_yield),
                 float(s.smoke_yield))
        ins = (self.fuel, solver._temperature, solver._front)
        outs = (self.spare, solver._temperature_back, solver._back, self.flame)
        if active is None:
            _lib.call("fx_combustion", self.fuel.size, *rates, *(f.ptr for f in ins), *(f.ptr for f in outs),
                      self.device.stream)
        else:
            active.pointwise_plan(ins, outs)
            _lib.call("fx_combustion_active", solver._g, *rates, *(f.ptr for f in ins), *(f.ptr for f in outs),
                      active.react.data_ptr(), active.out_masks(outs), self.device.stream)
        self.fuel, self.spare = self.spare, self.fuel
        solver._front, solver._back = solver._back, solver._front
        solver._temperature, solver._temperature_back = solver._temperature_back, solver._temperature

    def close(self):
        self.fuel = self.spare = self.flame = None
