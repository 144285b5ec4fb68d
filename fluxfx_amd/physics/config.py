"""Grid descriptor and base advection settings.

Interface of physics/config.py:6-62 of the reference: the names ``GridSpec``,
``AdvectionSettings``, ``workgroups`` and ``validate_dt``, their fields,
defaults, properties and ``ValueError`` messages.  One addition: ``max_dim``
lifts the reference's 128-cell cap for the 256^3/384^3 workloads; its default
keeps the reference behaviour.

Validation is table driven: each settings class lists (check, message) rules
evaluated in order, and the first failing rule raises its message.
"""
from dataclasses import dataclass
from math import isfinite, prod

# one float32 value per cell or face on the device
FLOAT_BYTES = 4


def _finite(*values):
    return all(isfinite(v) for v in values)


def _raise_first(rules, obj):
    for check, message in rules:
        if not check(obj):
            raise ValueError(message)


def _grid_rules(max_dim):
    def dims_ok(g):
        return len(g.shape) == 3 and all(type(n) is int and 2 <= n <= max_dim for n in g.shape)

    return ((dims_ok, f"P0 grid dimensions must be integers in [2, {max_dim}]"),
            (lambda g: len(g.extent) == 3 and all(isfinite(e) and e > 0 for e in g.extent),
             "Domain extents must be finite and positive"))


@dataclass(frozen=True)
class GridSpec:
    """Cell counts (nx, ny, nz) and the domain size in metres."""
    shape: tuple = (64, 64, 64)
    extent: tuple = (1.0, 1.0, 1.0)
    max_dim: int = 128

    def __post_init__(self):
        if type(self.max_dim) is not int or self.max_dim < 2:
            raise ValueError("max_dim must be an integer >= 2")
        _raise_first(_grid_rules(self.max_dim), self)

    @property
    def cell_size(self):
        return tuple(length / count for length, count in zip(self.extent, self.shape))

    @property
    def face_shapes(self):
        """Staggered U, V, W arrays: one extra face along their own axis."""
        out = []
        for axis in range(3):
            shape = list(self.shape)
            shape[axis] += 1
            out.append(tuple(shape))
        return tuple(out)

    @property
    def cells(self):
        return prod(self.shape)

    @property
    def density_bytes(self):
        """Front and back density buffers."""
        return 2 * FLOAT_BYTES * self.cells

    @property
    def mac_bytes(self):
        """Density buffers plus front and back U, V, W."""
        faces = sum(prod(s) for s in self.face_shapes)
        return self.density_bytes + 2 * FLOAT_BYTES * faces


_ADVECTION_RULES = (
    (lambda s: len(s.velocity) == 3 and len(s.source_center) == 3,
     "Velocity and source center must have three components"),
    (lambda s: _finite(*s.velocity, *s.source_center, s.angular_speed, s.source_radius, s.source_rate,
                       s.dissipation),
     "Simulation parameters must be finite"),
    (lambda s: s.source_radius > 0 and s.source_rate >= 0 and s.dissipation >= 0,
     "Radius must be positive; source rate and dissipation nonnegative"),
)


@dataclass(frozen=True)
class AdvectionSettings:
    velocity: tuple = (0.0, 0.0, 0.25)
    angular_speed: float = 0.8
    source_center: tuple = (0.5, 0.5, 0.16)
    source_radius: float = 0.12
    source_rate: float = 2.0
    dissipation: float = 0.1

    def __post_init__(self):
        _raise_first(_ADVECTION_RULES, self)


def workgroups(shape, local_size=(4, 4, 4)):
    """Groups per axis that cover ``shape`` with ``local_size`` blocks."""
    if len(shape) != 3 or len(local_size) != 3 or min(*shape, *local_size) <= 0:
        raise ValueError("Dispatch requires three positive dimensions")
    return tuple(-(-count // block) for count, block in zip(shape, local_size))


def validate_dt(dt):
    if not (isfinite(dt) and 0 < dt <= 0.1):
        raise ValueError("P0 timestep must be finite and in (0, 0.1] seconds")
