"""Source descriptors and the packed source table (physics/emission.py:7-58)."""
from dataclasses import dataclass
from math import isfinite, sqrt

MAX_SOURCES = 8
SOURCE_TABLE_SHAPE = (4, 4, MAX_SOURCES)


@dataclass(frozen=True)
class Emission:
    start: tuple
    velocity: tuple = (0., 0., 0.)
    coupling: float = 0.

    def __post_init__(self):
        if len(self.start) != 3 or len(self.velocity) != 3 or \
                not all(isfinite(x) for x in (*self.start, *self.velocity, self.coupling)) or self.coupling < 0:
            raise ValueError("Invalid emission parameters")


def motion_velocity(previous, current, interval, inheritance, jet, speed_limit):
    """Inherited translation capped at speed_limit, plus the jet."""
    if any(len(v) != 3 for v in (previous, current, jet)) or \
            not all(isfinite(v) for v in (*previous, *current, *jet, interval, inheritance, speed_limit)) or \
            interval <= 0 or inheritance < 0 or speed_limit <= 0:
        raise ValueError("Invalid motion interval or limits")
    vel = tuple((b - a) / interval * inheritance for a, b in zip(previous, current))
    speed = sqrt(sum(v * v for v in vel))
    factor = min(1., speed_limit / max(speed, 1e-20))
    return tuple(v * factor + j for v, j in zip(vel, jet))


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
        if not isfinite(self.fuel_rate) or self.fuel_rate < 0:
            raise ValueError("Fuel rate must be finite and nonnegative")
        if self.density_mode not in {"RATE", "TARGET"} or self.profile not in {"SOFT", "SOLID"}:
            raise ValueError("Unknown source mode/profile")
        if len(self.center) != 3 or not all(isfinite(v) for v in (*self.center, self.radius, self.density_rate,
                                                                   self.heat_rate)) \
                or self.radius <= 0 or self.density_rate < 0 or not isinstance(self.motion, Emission):
            raise ValueError("Invalid spherical source")

    def row(self):
        """16 floats: centre, radius, start, coupling, velocity, density rate,
        heat rate, target flag, solid flag, fuel rate."""
        return (*self.center, self.radius, *self.motion.start, self.motion.coupling, *self.motion.velocity,
                self.density_rate, self.heat_rate, float(self.density_mode == "TARGET"),
                float(self.profile == "SOLID"), self.fuel_rate)


def pack_sources(sources):
    """One 16-float row per source, zero-padded to MAX_SOURCES rows."""
    if len(sources) > MAX_SOURCES:
        raise ValueError("At most eight sources are supported")
    data = []
    for source in sources:
        if not isinstance(source, Source):
            raise TypeError("Expected Source")
        data.extend(source.row())
    return tuple(data + [0.] * (16 * MAX_SOURCES - len(data)))
