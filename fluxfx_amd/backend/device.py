"""GPU boundary: device arrays standing in for the reference's 3D textures.

Mirrors BlenderGPUDevice.texture/read/probe_3d (backend/device.py:30-95):
``texture(shape, values=None, *, nonnegative=True, channels=1)`` returns an owned
field with ``clear``/``read``; invalid uploads raise ValueError before any state
changes.  Memory is owned by torch (HIP allocations on cuda:N); every kernel
runs through libfluxhip's C-ABI on torch's current stream.
"""
from math import isfinite, prod
import ctypes as C

import numpy as np

from .. import _lib
from .._lib import GPUUnavailable

__all__ = ["HipDevice", "Field", "GPUUnavailable"]


class Field:
    """One device array in the reference's flat x-fast layout.

    ``live``: the field's exact active-brick mask (a uint8 tensor, one byte
    per 8^3 brick, 0 = every cell of the brick is exactly 0) kept by the
    active-brick scalar passes, or None when unknown.  A new field starts
    unknown and clear() keeps it exact; code that writes the tensor in place
    by other means sets it to None."""

    __slots__ = ("tensor", "shape", "channels", "device", "live")

    def __init__(self, tensor, shape, channels, device):
        self.tensor = tensor
        self.shape = tuple(shape)
        self.channels = channels
        self.device = device
        self.live = None

    @property
    def ptr(self):
        return self.tensor.data_ptr()

    @property
    def size(self):
        return self.tensor.numel()

    def clear(self, format="FLOAT", value=(0.0,)):
        if self.tensor.dtype == self.device.torch.float32 and len(set(value)) == 1:
            _lib.call("fx_fill", self.ptr, self.size, float(value[0]), self.device.stream)
        else:
            self.tensor.fill_(value[0])
        if self.live is not None and all(v == 0 for v in value):
            self.live.zero_()
        else:
            self.live = None

    def read(self):
        return self.device.read(self, self.shape)

    def numpy(self):
        """[z, y, x] host copy (synchronous)."""
        nx, ny, nz = self.shape
        a = self.tensor.detach().cpu().numpy()
        return a.reshape((nz, ny, nx) if self.channels == 1 else (nz, ny, nx, self.channels))


class HipDevice:
    """Owns the torch device, the stream handle and libfluxhip."""

    def __init__(self, index=0):
        try:
            import torch
        except ImportError as exc:  # pragma: no cover
            raise GPUUnavailable("PyTorch-ROCm is required for device memory") from exc
        if not torch.cuda.is_available():
            raise GPUUnavailable("No HIP device is visible")
        self.torch = torch
        self.index = index
        self.device = torch.device("cuda", index)
        _lib.load()

    # -- streams ---------------------------------------------------------------
    @property
    def stream(self):
        return C.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def synchronize(self):
        self.torch.cuda.synchronize(self.device)

    # -- memory ----------------------------------------------------------------
    def empty(self, count, dtype="float32"):
        return self.torch.empty(int(count), dtype=getattr(self.torch, dtype), device=self.device)

    def zeros(self, count, dtype="float32"):
        return self.torch.zeros(int(count), dtype=getattr(self.torch, dtype), device=self.device)

    def texture(self, shape, values=None, *, nonnegative=True, channels=1, dtype="float32"):
        if channels not in (1, 4):
            raise ValueError("Only scalar or four-channel fields are supported")
        n = prod(shape) * channels
        if values is not None:
            arr = np.asarray(values, dtype=np.float64).ravel()
            if arr.size != n or not np.isfinite(arr).all() or (nonnegative and (arr < 0).any()):
                raise ValueError("Field upload must match grid and satisfy finite/sign constraints")
            t = self.torch.from_numpy(arr.astype(np.float32 if dtype == "float32" else dtype)).to(self.device)
        else:
            t = self.zeros(n, dtype)
        return Field(t, shape, channels, self)

    def field_like(self, other):
        return Field(self.zeros(other.size, str(other.tensor.dtype).split(".")[-1]), other.shape,
                     other.channels, self)

    @staticmethod
    def read(field, shape=None):
        """Synchronous readback as a flat x-fast list."""
        if shape is not None and prod(shape) * field.channels != field.size:
            raise ValueError("read shape does not match the field")
        return field.tensor.detach().cpu().double().tolist()

    def probe_3d(self):
        """Exact 7x5x3 pattern through a device fill/scale round trip
        (backend/device.py:83-95)."""
        shape = (7, 5, 3)
        expected = [float(x + 10 * y + 100 * z) for z in range(3) for y in range(5) for x in range(7)]
        f = self.texture(shape, expected)
        g = self.texture(shape)
        _lib.call("fx_scale", f.ptr, g.ptr, g.size, 1.0, self.stream)
        actual = g.read()
        err = max(abs(a - b) for a, b in zip(actual, expected))
        if not all(isfinite(v) for v in actual) or err > 1e-6:
            raise GPUUnavailable(f"3D store/readback mismatch: {err}")
        return {"status": "PASS", "format": "R32F", "shape": shape, "texels_checked": len(actual),
                "max_abs_error": err}


def grid_struct(grid):
    return _lib.make_grid(grid.shape, grid.extent)


def ptr(x):
    """Device address of a Field / tensor / None."""
    if x is None:
        return None
    if isinstance(x, Field):
        return x.ptr
    return x.data_ptr()
