"""Grid descriptor and base advection settings.

Mirrors physics/config.py:6-62 of the reference (same names, validation and
errors).  One addition: ``max_dim`` lifts the reference's 128-cell cap for the
256^3/384^3 workloads; its default keeps the reference behaviour.
"""
from dataclasses import dataclass
from math import ceil, isfinite, prod


@dataclass(frozen=True)
class GridSpec:
    shape: tuple = (64, 64, 64)
    extent: tuple = (1.0, 1.0, 1.0)
    max_dim: int = 128

    def __post_init__(self):
        if type(self.max_dim) is not int or self.max_dim < 2:
            raise ValueError("max_dim must be an integer >= 2")
        if len(self.shape) != 3 or any(type(n) is not int or not 2 <= n <= self.max_dim for n in self.shape):
            raise ValueError(f"P0 grid dimensions must be integers in [2, {self.max_dim}]")
        if len(self.extent) != 3 or any(not isfinite(x) or x <= 0 for x in self.extent):
            raise ValueError("Domain extents must be finite and positive")

    @property
    def cell_size(self):
        return tuple(e / n for e, n in zip(self.extent, self.shape))

    @property
    def face_shapes(self):
        return tuple(tuple(n + (1 if i == a else 0) for i, n in enumerate(self.shape)) for a in range(3))

    @property
    def cells(self):
        return prod(self.shape)

    @property
    def mac_bytes(self):
        return self.density_bytes + 2 * 4 * sum(prod(s) for s in self.face_shapes)

    @property
    def density_bytes(self):
        return 2 * 4 * prod(self.shape)


@dataclass(frozen=True)
class AdvectionSettings:
    velocity: tuple = (0.0, 0.0, 0.25)
    angular_speed: float = 0.8
    source_center: tuple = (0.5, 0.5, 0.16)
    source_radius: float = 0.12
    source_rate: float = 2.0
    dissipation: float = 0.1

    def __post_init__(self):
        if len(self.velocity) != 3 or len(self.source_center) != 3:
            raise ValueError("Velocity and source center must have three components")
        values = (*self.velocity, self.angular_speed, *self.source_center, self.source_radius,
                  self.source_rate, self.dissipation)
        if not all(isfinite(v) for v in values):
            raise ValueError("Simulation parameters must be finite")
        if self.source_radius <= 0 or self.source_rate < 0 or self.dissipation < 0:
            raise ValueError("Radius must be positive; source rate and dissipation nonnegative")


def workgroups(shape, local_size=(4, 4, 4)):
    if len(shape) != 3 or len(local_size) != 3 or any(n <= 0 for n in (*shape, *local_size)):
        raise ValueError("Dispatch requires three positive dimensions")
    return tuple(ceil(n / l) for n, l in zip(shape, local_size))


def validate_dt(dt):
    if not isfinite(dt) or not 0 < dt <= 0.1:
        raise ValueError("P0 timestep must be finite and in (0, 0.1] seconds")
