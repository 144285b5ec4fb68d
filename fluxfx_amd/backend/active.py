"""Exact active bricks of the scalar passes (fx_active_* in libfluxhip).

No reference counterpart: the reference's sparse engine keeps density on
bricks above a 1e-4 threshold (native/projection.mm), which changes values.
Here each scalar field carries a live mask (Field.live: one byte per 8^3
brick, 0 = every cell exactly 0) and the scalar predictor, the scalar
transport and the combustion pass run under plans built from those masks:
bricks that can only produce zeros are skipped (or stored 0 where the output
buffer holds stale values), so the step's values are those of the whole-grid
passes.  Masks of fields written by other code are unknown (None) and are
rebuilt by a scan before the next plan.
"""
import ctypes as C
from math import prod

from .. import _lib
from .device import ptr

SIDE = 8


class ActiveBricks:
    def __init__(self, grid, device, g):
        self.grid, self.device, self._g = grid, device, g
        self.shape = tuple(-(-n // SIDE) for n in grid.shape)  # bricks along x, y, z
        self.count = prod(self.shape)
        self.predict = device.zeros(self.count, "uint8")
        self.write = device.zeros(self.count, "uint8")
        self.react = device.zeros(self.count, "uint8")
        self.vmax = device.zeros(3)
        # the projector's divergence of the projected velocity also writes
        # vmax (fx_divergence_bound) and sets `bound_ready`; else fx_max_abs
        self.partials = device.zeros(3 * -(-prod(grid.shape) // 1024))
        self.bound_ready = False
        self._sizes = (C.c_int64 * 3)(*[prod(s) for s in grid.face_shapes])

    @property
    def allocated_bytes(self):
        # three plans, the velocity bound and its partials, and the live masks
        # of up to seven scalar fields (density, temperature and fuel pairs, flame)
        return 10 * self.count + 12 + 4 * self.partials.numel()

    def live(self, field):
        """The field's live mask, scanned from its values when unknown."""
        if field.live is None:
            mask = self.device.zeros(self.count, "uint8")
            _lib.call("fx_active_scan", self._g, field.ptr, mask.data_ptr(), self.device.stream)
            field.live = mask
        return field.live

    def _masks(self, fields):
        return (C.c_void_p * len(fields))(*[self.live(f).data_ptr() for f in fields])

    def transport_plan(self, dt, velocity, inputs, outputs, sources):
        """Plan of the scalar predictor (self.predict) and transport
        (self.write) of `inputs` into `outputs` by `velocity` over dt."""
        if not self.bound_ready:
            arrays = (C.c_void_p * 3)(*[ptr(f) for f in velocity])
            _lib.call("fx_max_abs", 3, arrays, self._sizes, self.vmax.data_ptr(), self.device.stream)
        self.bound_ready = False
        _lib.call("fx_active_plan", self._g, float(dt), self.vmax.data_ptr(), 1, len(inputs), self._masks(inputs),
                  len(outputs), self._masks(outputs), sources, self.predict.data_ptr(), self.write.data_ptr(),
                  self.device.stream)

    def pointwise_plan(self, inputs, outputs):
        """Plan of a pointwise pass (self.react)."""
        _lib.call("fx_active_plan", self._g, 0.0, None, 0, len(inputs), self._masks(inputs), len(outputs),
                  self._masks(outputs), None, None, self.react.data_ptr(), self.device.stream)

    def out_masks(self, fields):
        return self._masks(fields)

    def fractions(self):
        """Shares of bricks the last plans compute or store 0 in (host read,
        diagnostics)."""
        share = lambda mask, state: float((mask == state).float().mean())
        return {"predict": share(self.predict, 1), "transport": share(self.write, 1),
                "transport_zero": share(self.write, 2), "combustion": share(self.react, 1)}

    def close(self):
        self.predict = self.write = self.react = self.vmax = self.partials = None
