"""CFL timestep selection.

Interface of physics/timestep.py:6-23 (``choose_timestep(max_dt, cfl, rate,
acceleration_rate=0.0, *, quantized=True)``, its errors and messages).

A step dt moves material by at most ``rate*dt + acceleration_rate*dt**2``
cells, and the bound must stay at or below the CFL target c.  The largest
such dt is the positive root of ``a*dt**2 + r*dt - c = 0``:

    dt* = (-r + sqrt(r**2 + 4*a*c)) / (2*a) = 2*c / (r + sqrt(r**2 + 4*a*c)),

written in the second (rationalised) form, which has no cancellation when
r**2 >> a*c and stays finite for a = 0.  The continuous policy takes 99 % of
dt*; the quantized policy takes max_dt / 2**k for the smallest k that meets
the bound (power-of-two steps keep timestep-scaled pressure comparable across
steps).  Either way a step below one microsecond is refused.
"""
from math import floor, isfinite, log2, sqrt

from .config import validate_dt

MIN_DT = 1e-6
MARGIN = 0.99


def _cfl_root(cfl, rate, acceleration):
    return 2.0 * cfl / (rate + sqrt(rate * rate + 4.0 * acceleration * cfl))


def _too_small():
    return RuntimeError("Required timestep below 1 microsecond; reduce forces or reset")


def choose_timestep(max_dt, cfl, rate, acceleration_rate=0.0, *, quantized=True):
    validate_dt(max_dt)
    valid = all(isfinite(v) for v in (cfl, rate, acceleration_rate)) and 0 < cfl <= 2 and \
        rate >= 0 and acceleration_rate >= 0
    if not valid:
        raise ValueError("CFL and rate bounds must be finite and valid")

    def courant(dt):
        return rate * dt + acceleration_rate * dt * dt

    if courant(max_dt) <= cfl:
        return max_dt
    root = _cfl_root(cfl, rate, acceleration_rate)
    if not quantized:
        dt = MARGIN * root
        if dt < MIN_DT:
            raise _too_small()
        return min(max_dt, dt)
    # smallest k with courant(max_dt / 2**k) <= cfl: start from the root's
    # estimate one below and step up (at most a couple of checks, rounding)
    k = max(1, floor(log2(max_dt / root)))
    while courant(max_dt / 2.0 ** k) > cfl:
        k += 1
    dt = max_dt / 2.0 ** k
    if dt < MIN_DT:
        raise _too_small()
    return dt
