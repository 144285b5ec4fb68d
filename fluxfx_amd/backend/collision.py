"""Static voxel solids: mask build, open-face flags, solid clear.

Plays the role of SolidDevice (backend/collision.py:12-72): the reference
decorates the device so masked shader variants receive the mask; here the
masked kernels take the mask/flags pointers directly, so this object owns the
uint8 mask (1 = solid) and the derived level-0 flags.
"""
import ctypes

import numpy as np

from .. import _lib
from .device import HipDevice, grid_struct
from .pressure import SolidPressureProjector

__all__ = ["SolidMask", "SolidPressureProjector"]


class SolidMask:
    def __init__(self, grid, colliders=(), device=None, *, mask=None):
        if mask is None and not 1 <= len(colliders) <= 8:
            raise ValueError("Use one to eight static colliders")
        self.device = device or HipDevice()
        self.grid = grid
        self._g = grid_struct(grid)
        n = grid.cells
        if mask is not None:
            m = np.asarray(mask).reshape(-1)
            if m.size != n:
                raise ValueError("Mask size must match grid")
            self.solid = self.device.torch.from_numpy((m > 0.5).astype(np.uint8)).to(self.device.device)
        else:
            self.solid = self.device.zeros(n, "uint8")
            for c in colliders:
                if c.shape not in ("SPHERE", "BOX"):
                    raise ValueError("Collider must be a sphere or box")
                rows = (ctypes.c_float * 12)(*[float(v) for r in c.inverse_rows for v in r])
                _lib.call("fx_solid_raster", self._g, 0 if c.shape == "SPHERE" else 1, rows,
                          self.solid.data_ptr(), self.device.stream)
        self.flags = self.device.zeros(n, "uint8")
        _lib.call("fx_build_flags", self._g, self.solid.data_ptr(), self.flags.data_ptr(), self.device.stream)
        # Chebyshev clearance to the nearest solid/wall cell: lets the transport
        # kernels skip trace clipping far from solids
        self.clearance = self.device.zeros(n, "uint8")
        scratch = self.device.zeros(n, "uint8")
        _lib.call("fx_build_clearance", self._g, self.solid.data_ptr(), self.clearance.data_ptr(),
                  scratch.data_ptr(), self.device.stream)
        self.device.synchronize()
        del scratch

    @property
    def mask(self):
        return self.solid

    @property
    def allocated_bytes(self):
        return 3 * self.grid.cells

    def numpy(self):
        nx, ny, nz = self.grid.shape
        return self.solid.cpu().numpy().reshape(nz, ny, nx).astype(bool)

    def clear_solid(self, source, target):
        """target = solid ? 0 : source (shaders/solid_clear.glsl)."""
        _lib.call("fx_solid_clear", self._g, self.solid.data_ptr(), source.ptr, target.ptr, self.device.stream)
        target.live = None

    def close(self):
        self.solid = self.flags = self.clearance = None
