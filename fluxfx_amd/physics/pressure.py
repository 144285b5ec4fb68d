"""Pressure/solver settings and divergence statistics.

Mirrors physics/pressure.py:7-80 of the reference (field names, defaults,
validation messages).  Added fields (no reference counterpart) configure the
tolerance-driven solvers:

* ``pressure_tolerance``: 0 keeps the reference's fixed cycle budget; > 0 runs
  to ||b - A q|| <= tol*||b|| (b mean-projected), true FP32 residual, FP64 norms.
* ``pressure_precision``: storage of V-cycle levels, FP32 / BF16 / FP16; BF16 and
  FP16 run inside FP32 iterative refinement.
* ``pressure_krylov``: NONE (stationary V-cycles), PCG (flexible CG with a
  V-cycle preconditioner) or LEARNED (flexible CG with the U-Net
  preconditioner, MG fallback).
* ``learned_weights``/``learned_iterations``: the U-Net's trained parameters
  (.npz written by UNetPreconditioner.save) and how many U-Net preconditioned
  iterations run before the V-cycle preconditioned fallback takes over;
  ``learned_level`` 1 runs the network on the 8-child mean of the residual one
  multigrid level down (a 128^3-trained network on a 256^3 grid sees its own
  resolution) and prolongs its output before the smoothing sweeps.
* ``pressure_fixed_cycles``: > 0 runs exactly that many stationary V-cycles of
  the config-mode hierarchy below (the configs' fixed budget, e.g. 8 per step)
  instead of solving to a tolerance.
* ``mg_pre``/``mg_post``/``mg_coarsest``/``mg_restriction``/``mg_levels``:
  V-cycle shape.  ``mg_coarsest`` = 4 with 3/3 sweeps is the reference
  (backend/multigrid.py:57-81); the configs use 2/2 and an 8^3 FP64 direct solve.
"""
from dataclasses import dataclass
from math import isfinite, fsum, sqrt
from .thermal import ThermalSettings

_CHOICES = {
    "turbulence_mask": {"ALL", "DENSITY", "HEAT"},
    "coarse_solver": {"DIRECT", "SMOOTH"},
    "timestep_policy": {"AUTO", "CONTINUOUS", "QUANTIZED"},
    "velocity_advection": {"SEMI_LAGRANGIAN", "MACCORMACK"},
    "scalar_advection": {"SEMI_LAGRANGIAN", "MACCORMACK"},
    "pressure_warm_start": {"AUTO", "IMPULSE", "LEGACY"},
    "density_mode": {"RATE", "TARGET"},
    "source_profile": {"SOFT", "SOLID"},
    "pressure_solver": {"AUTO", "JACOBI", "MULTIGRID"},
    "pressure_precision": {"FP32", "BF16", "FP16"},
    "pressure_krylov": {"NONE", "PCG", "LEARNED"},
    "mg_restriction": {"MEAN", "FULL_WEIGHTING"},
}

_MESSAGES = {
    "turbulence_mask": "Invalid turbulence mask",
    "coarse_solver": "Unknown coarse pressure solver",
    "timestep_policy": "Unknown timestep policy",
    "velocity_advection": "Unknown velocity advection",
    "scalar_advection": "Unknown scalar advection",
    "pressure_warm_start": "Unknown pressure warm start",
    "density_mode": "Unknown emission mode/profile",
    "source_profile": "Unknown emission mode/profile",
    "pressure_solver": "Unknown pressure solver",
    "pressure_precision": "Unknown pressure precision",
    "pressure_krylov": "Unknown pressure Krylov method",
    "mg_restriction": "Unknown multigrid restriction",
}


# fields with no reference counterpart (kept apart in checkpoints so that the
# reference's loader reads the settings it knows)
EXTENSION_FIELDS = ("pressure_tolerance", "pressure_max_cycles", "pressure_fixed_cycles", "pressure_precision",
                    "pressure_krylov", "mg_pre", "mg_post", "mg_coarsest", "mg_restriction", "learned_weights",
                    "learned_iterations", "learned_sweeps", "learned_level")


def _int_in(value, low, high):
    return type(value) is int and low <= value <= high


@dataclass(frozen=True)
class PressureSettings(ThermalSettings):
    turbulence_strength: float = 0.0
    turbulence_scale: float = 0.5
    turbulence_speed: float = 0.5
    turbulence_seed: int = 0
    turbulence_octaves: int = 3
    turbulence_limit: float = 2.0
    turbulence_mask: str = "DENSITY"
    turbulence_threshold: float = 0.2
    combustion_enabled: bool = False
    fuel_source_rate: float = 1.0
    ignition_temperature: float = 150.0
    burn_rate: float = 4.0
    heat_yield: float = 600.0
    smoke_yield: float = 1.0
    moving_colliders: bool = False
    velocity_advection: str = "SEMI_LAGRANGIAN"
    timestep_policy: str = "AUTO"
    pressure_warm_start: str = "AUTO"
    coarse_solver: str = "DIRECT"
    density_mode: str = "RATE"
    source_profile: str = "SOFT"
    scalar_advection: str = "SEMI_LAGRANGIAN"
    vorticity_strength: float = 0.0
    vorticity_limit: float = 2.0
    pressure_solver: str = "AUTO"
    pressure_cycles: int = 2
    pressure_iterations: int = 80
    pressure_relaxation: float = 2.0 / 3.0
    pressure_cold_multiplier: int = 4
    # tolerance-driven solvers (no reference counterpart)
    pressure_tolerance: float = 0.0
    pressure_max_cycles: int = 40
    # > 0: exactly this many V-cycles of the config-mode hierarchy per step, no tolerance
    pressure_fixed_cycles: int = 0
    pressure_precision: str = "FP32"
    pressure_krylov: str = "NONE"
    mg_pre: int = 3
    mg_post: int = 3
    mg_coarsest: int = 4
    mg_restriction: str = "MEAN"
    # learned preconditioner: trained weights (.npz) and the U-Net iteration budget before MG fallback
    learned_weights: str = ""
    learned_iterations: int = 2
    learned_sweeps: int = 2
    # grid the U-Net runs on: 0 = the pressure grid, 1 = one multigrid level coarser
    learned_level: int = 0

    def __post_init__(self):
        super().__post_init__()
        for value, low, high in ((self.turbulence_strength, 0, 20), (self.turbulence_scale, .01, 4),
                                 (self.turbulence_speed, 0, 10), (self.turbulence_limit, .01, 20),
                                 (self.turbulence_threshold, .001, 2000)):
            if not isfinite(value) or not low <= value <= high:
                raise ValueError("Invalid turbulence parameter")
        if not _int_in(self.turbulence_seed, 0, 65535):
            raise ValueError("Invalid turbulence seed")
        if not _int_in(self.turbulence_octaves, 1, 4):
            raise ValueError("Use one to four turbulence scales")
        for name, allowed in _CHOICES.items():
            if getattr(self, name) not in allowed:
                raise ValueError(_MESSAGES[name])
        if type(self.combustion_enabled) is not bool:
            raise ValueError("Combustion enabled must be boolean")
        if not all(isfinite(v) and v >= 0 for v in (self.fuel_source_rate, self.ignition_temperature,
                                                    self.burn_rate, self.heat_yield, self.smoke_yield)):
            raise ValueError("Combustion parameters must be finite and nonnegative")
        if not isfinite(self.vorticity_strength) or not 0 <= self.vorticity_strength <= 20:
            raise ValueError("Curl strength must lie in [0,20]")
        if not isfinite(self.vorticity_limit) or not 0 < self.vorticity_limit <= 20:
            raise ValueError("Curl acceleration limit must lie in (0,20]")
        if not _int_in(self.pressure_cycles, 1, 16):
            raise ValueError("Multigrid cycles must be an integer in [1, 16]")
        if not _int_in(self.pressure_iterations, 1, 4096):
            raise ValueError("Pressure iterations must be an integer in [1, 4096]")
        if not _int_in(self.pressure_cold_multiplier, 1, 8):
            raise ValueError("Cold-start multiplier must be an integer in [1, 8]")
        if not isfinite(self.pressure_relaxation) or not 0 < self.pressure_relaxation < 1:
            raise ValueError("Jacobi relaxation must lie strictly between zero and one")
        if not isfinite(self.pressure_tolerance) or not 0 <= self.pressure_tolerance < 1:
            raise ValueError("Pressure tolerance must lie in [0, 1)")
        if not _int_in(self.pressure_max_cycles, 1, 1000):
            raise ValueError("Maximum pressure cycles must be an integer in [1, 1000]")
        if not _int_in(self.pressure_fixed_cycles, 0, 64):
            raise ValueError("Fixed pressure cycles must be an integer in [0, 64]")
        if self.pressure_fixed_cycles and (self.pressure_tolerance > 0 or self.pressure_krylov != "NONE"):
            raise ValueError("Fixed pressure cycles run stationary V-cycles without a tolerance")
        if not _int_in(self.mg_pre, 0, 16) or not _int_in(self.mg_post, 0, 16) or self.mg_pre + self.mg_post == 0:
            raise ValueError("Smoothing sweeps must be integers in [0, 16], not both zero")
        if self.mg_coarsest not in (2, 4, 8, 16):
            raise ValueError("Coarsest multigrid side must be 2, 4, 8 or 16")
        if self.learned_level not in (0, 1):
            raise ValueError("The learned preconditioner runs on level 0 or 1")
        if not _int_in(self.learned_iterations, 0, 1000):
            raise ValueError("Learned iterations must be an integer in [0, 1000]")
        if not _int_in(self.learned_sweeps, 0, 16) or self.learned_sweeps % 2:
            raise ValueError("Learned smoothing sweeps must be an even integer in [0, 16]")


def divergence_stats(before, after):
    """RMS/max/mean before and after projection (physics/pressure.py:70-80)."""
    before = list(before)
    after = list(after)
    if not before or len(before) != len(after):
        raise ValueError("Divergence arrays must be nonempty and equal in length")
    if not all(isfinite(v) for v in (*before, *after)):
        raise ValueError("Non-finite divergence")
    rms_before = sqrt(fsum(v * v for v in before) / len(before))
    rms_after = sqrt(fsum(v * v for v in after) / len(after))
    ratio = rms_after / rms_before if rms_before > 0 else (0.0 if rms_after == 0 else None)
    return {"rms_before": rms_before, "rms_after": rms_after, "ratio": ratio,
            "max_before": max(abs(v) for v in before), "max_after": max(abs(v) for v in after),
            "mean_before": fsum(before) / len(before), "mean_after": fsum(after) / len(after)}
