"""Scalar one-step fuel kinetics (physics/combustion.py:6-11)."""
from math import expm1, isfinite
from .config import validate_dt


def react(fuel, temperature, dt, ignition, burn_rate):
    """Consumed fuel over dt: first-order decay above ignition."""
    validate_dt(dt)
    if not all(isfinite(v) for v in (fuel, temperature, ignition, burn_rate)) or min(fuel, ignition, burn_rate) < 0:
        raise ValueError("Invalid combustion state")
    return fuel * -expm1(-burn_rate * dt) if temperature >= ignition else 0.
