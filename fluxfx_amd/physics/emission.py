"""Source descriptors and the packed source table.

Interface of physics/emission.py:7-58: ``Emission``, ``motion_velocity``,
``Source``, ``pack_sources``, ``MAX_SOURCES``/``SOURCE_TABLE_SHAPE``, their
fields, errors and messages, and the table layout (one 16-float row per
source, zero padded to eight rows), which the HIP source table reads.
"""
from dataclasses import dataclass
from itertools import chain
from math import isfinite, sqrt

MAX_SOURCES = 8
ROW_FLOATS = 16
SOURCE_TABLE_SHAPE = (4, 4, MAX_SOURCES)
DENSITY_MODES = frozenset({"RATE", "TARGET"})
PROFILES = frozenset({"SOFT", "SOLID"})


def _vec3(v):
    return len(v) == 3


def _all_finite(values):
    return all(isfinite(x) for x in values)


@dataclass(frozen=True)
class Emission:
    """Where a source swept from this step, the jet it imposes and the
    coupling rate of that jet (1/s)."""
    start: tuple
    velocity: tuple = (0., 0., 0.)
    coupling: float = 0.

    def __post_init__(self):
        shapes_ok = _vec3(self.start) and _vec3(self.velocity)
        if not (shapes_ok and _all_finite((*self.start, *self.velocity, self.coupling)) and self.coupling >= 0):
            raise ValueError("Invalid emission parameters")


def motion_velocity(previous, current, interval, inheritance, jet, speed_limit):
    """Velocity a moving source hands to the fluid: its displacement over
    ``interval`` times ``inheritance``, shortened to at most ``speed_limit``
    along the same direction, plus the ``jet``."""
    vectors_ok = all(_vec3(v) for v in (previous, current, jet))
    scalars = (interval, inheritance, speed_limit)
    if not (vectors_ok and _all_finite((*previous, *current, *jet, *scalars))
            and interval > 0 and inheritance >= 0 and speed_limit > 0):
        raise ValueError("Invalid motion interval or limits")
    scale = inheritance / interval
    inherited = [(b - a) * scale for a, b in zip(previous, current)]
    speed = sqrt(sum(c * c for c in inherited))
    shrink = 1. if speed <= speed_limit else speed_limit / max(speed, 1e-20)
    return tuple(c * shrink + j for c, j in zip(inherited, jet))


@dataclass(frozen=True)
class Source:
    center: tuple
    radius: float
    density_rate: float
    heat_rate: float
    motion: Emission
    density_mode: str = "RATE"
    profile: str = "SOFT"
    fuel_rate: float = 0.0

    def __post_init__(self):
        checks = (
            (isfinite(self.fuel_rate) and self.fuel_rate >= 0, "Fuel rate must be finite and nonnegative"),
            (self.density_mode in DENSITY_MODES and self.profile in PROFILES, "Unknown source mode/profile"),
            (_vec3(self.center) and _all_finite((*self.center, self.radius, self.density_rate, self.heat_rate))
             and self.radius > 0 and self.density_rate >= 0 and isinstance(self.motion, Emission),
             "Invalid spherical source"),
        )
        for ok, message in checks:
            if not ok:
                raise ValueError(message)

    def row(self):
        """16 floats: centre, radius, start, coupling, velocity, density rate,
        heat rate, target flag, solid flag, fuel rate."""
        m = self.motion
        flags = (float(self.density_mode == "TARGET"), float(self.profile == "SOLID"))
        return (*self.center, self.radius, *m.start, m.coupling, *m.velocity, self.density_rate, self.heat_rate,
                *flags, self.fuel_rate)


def pack_sources(sources):
    """The source table as one flat tuple of MAX_SOURCES rows."""
    sources = list(sources)
    if len(sources) > MAX_SOURCES:
        raise ValueError("At most eight sources are supported")
    if not all(isinstance(s, Source) for s in sources):
        raise TypeError("Expected Source")
    padding = (0.,) * (ROW_FLOATS * (MAX_SOURCES - len(sources)))
    return tuple(chain.from_iterable(s.row() for s in sources)) + padding
