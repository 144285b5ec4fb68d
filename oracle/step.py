"""One dense smoke/fire step in FP64 (test infrastructure only).

Order follows DenseProjectedSmoke._step (backend/projected.py:149-202):
velocity transport + emission -> buoyancy on W (pre-transport rho, T) ->
projection (backend/pressure.py:57-98) -> scalar transport with the projected
velocity -> combustion (backend/combustion.py:21-31).
"""
import math
from dataclasses import dataclass, field
from typing import Optional
import numpy as np

from . import mac, multigrid, transport


def react(fuel, temperature, density, dt, ignition, burn_rate, heat_yield, smoke_yield):
    """burned = T>=T_ign ? max(fuel,0)*(1-exp(-k dt)) : 0; flame = burned/dt;
    fuel = max(fuel - dt*flame, 0); rho += dt*smoke*flame; T += dt*heat*flame.

    shaders/combustion_burn.glsl:1-8, shaders/combustion_apply.glsl:1-7,
    physics/combustion.py:6-11 (expm1 form).
    """
    burned = np.where(temperature >= ignition, np.maximum(fuel, 0.0) * -math.expm1(-burn_rate * dt), 0.0)
    flame = burned / dt
    fuel_n = np.maximum(fuel - dt * flame, 0.0)
    return fuel_n, temperature + dt * heat_yield * flame, density + dt * smoke_yield * flame, flame


def choose_timestep(max_dt, cfl, rate, acceleration_rate=0.0, quantized=True):
    """physics/timestep.py:6-23."""
    if not (0 < max_dt <= 0.1) or not math.isfinite(max_dt):
        raise ValueError("P0 timestep must be finite and in (0, 0.1] seconds")
    if not all(math.isfinite(v) for v in (cfl, rate, acceleration_rate)) or not 0 < cfl <= 2 \
            or min(rate, acceleration_rate) < 0:
        raise ValueError("CFL and rate bounds must be finite and valid")
    if not quantized:
        if rate * max_dt + acceleration_rate * max_dt * max_dt <= cfl:
            return max_dt
        den = rate + math.hypot(rate, 2 * math.sqrt(acceleration_rate) * math.sqrt(cfl))
        dt = .99 * (2 * cfl / den)
        if dt < 1e-6:
            raise RuntimeError("Required timestep below 1 microsecond; reduce forces or reset")
        return min(max_dt, dt)
    dt = max_dt
    while rate * dt + acceleration_rate * dt * dt > cfl:
        dt *= 0.5
        if dt < 1e-6:
            raise RuntimeError("Required timestep below 1 microsecond; reduce forces or reset")
    return dt


@dataclass
class PressureConfig:
    """Which projector the step uses.

    kind: 'JACOBI' (legacy units), 'MULTIGRID' (reference plain MG, impulse),
    'SOLID_JACOBI', 'SOLID_MULTIGRID' (reference aggregation MG),
    'TOLERANCE' (masked geometric MG to a relative residual, impulse).
    """
    kind: str = "JACOBI"
    iterations: int = 80
    relaxation: float = 2 / 3
    cold_multiplier: int = 4
    cycles: int = 2
    coarse_solver: str = "DIRECT"
    tol: float = 1e-5
    pre: int = 2
    post: int = 2
    coarsest: int = 8
    restriction: str = "MEAN"
    max_cycles: int = 50
    warm_start: bool = True
    # TOLERANCE only: run exactly this many V-cycles instead of stopping at tol
    # (parity tests give both sides the same amount of work)
    fixed_cycles: Optional[int] = None


@dataclass
class State:
    grid: mac.Grid
    velocity: tuple
    density: np.ndarray
    temperature: np.ndarray
    fuel: Optional[np.ndarray] = None
    flame: Optional[np.ndarray] = None
    solid: Optional[np.ndarray] = None
    pressure: Optional[np.ndarray] = None
    last_dt: Optional[float] = None
    info: dict = field(default_factory=dict)
    # moving colliders: per-cell owner (0 fluid) and the wall velocity per face axis
    owner: Optional[np.ndarray] = None
    boundary: Optional[tuple] = None


def move_colliders(state, previous, colliders, dt, rows_fn):
    """DenseProjectedSmoke.move_colliders (backend/projected.py:103-115):
    rasterise the new owners, wall velocities from the rigid motion, then
    remap the fields (backend/moving_collision.py:53-83).  ``previous`` and
    ``colliders``: (kind, rows) pairs; ``rows_fn(old, new, dt)`` gives a
    collider's velocity rows."""
    grid = state.grid
    old_solid = state.solid if state.solid is not None else np.zeros(grid.array_shape, bool)
    owner = mac.raster_owners(grid, colliders)
    solid = owner > 0
    boundary = mac.wall_velocity(grid, owner, [rows_fn(o, n, dt) for o, n in zip(previous, colliders)])
    state.density = mac.remap_scalar(state.density, solid, old_solid)
    state.temperature = mac.remap_scalar(state.temperature, solid, old_solid)
    if state.fuel is not None:
        state.fuel = mac.remap_scalar(state.fuel, solid, old_solid)
        state.flame = np.zeros(grid.array_shape)
    state.velocity = mac.remap_faces(state.velocity, grid, solid, old_solid, boundary)
    state.owner, state.solid, state.boundary = owner, solid, boundary


def project(state, provisional, dt, cfg, mg_cache=None):
    """PressureProjector.project (backend/pressure.py:57-83) for every projector kind."""
    grid, solid = state.grid, state.solid
    moving = state.boundary is not None
    bounded = mac.walls_moving(provisional, grid, solid, state.boundary) if moving else \
        mac.walls(provisional, grid, solid)
    before = mac.divergence(bounded, grid, solid)
    impulse = cfg.kind != "JACOBI"
    prev = state.last_dt
    p = state.pressure if state.pressure is not None else np.zeros(grid.array_shape)
    info = {}
    if cfg.kind == "JACOBI" or cfg.kind == "SOLID_JACOBI":
        solve_dt = 1.0 if impulse else dt
        prev_solve = (1.0 if prev is not None else None) if impulse else prev
        if impulse and prev is not None and abs(dt / prev - 1.0) > 1e-6:
            p = p * (dt / prev)
        cold = prev_solve != solve_dt
        if cold or cfg.kind == "SOLID_JACOBI":
            p = np.zeros(grid.array_shape)
        scaled = math.ceil(cfg.iterations * max(1.0, (max(grid.shape) / 64) ** 2))
        its = min(4096, scaled * (cfg.cold_multiplier if cold else 1))
        p = mac.jacobi_solve(before, grid, solve_dt, its, cfg.relaxation, solid, p0=p)
        info["iterations"] = its
    elif cfg.kind == "MULTIGRID":
        if prev is not None and abs(dt / prev - 1.0) > 1e-6:
            p = p * (dt / prev)
        if prev is None:
            p = np.zeros(grid.array_shape)
        solve_dt = 1.0
        if len(multigrid.hierarchy_shapes(grid.shape)) == 1:
            raise ValueError("grid does not coarsen")
        p = multigrid.reference_mg_solve(before, grid, 1.0, cfg.cycles, p0=p,
                                         coarse_solver=cfg.coarse_solver)
        info["cycles"] = cfg.cycles
    elif cfg.kind == "SOLID_MULTIGRID":
        solve_dt = 1.0
        cycles = cfg.cycles * (cfg.cold_multiplier if prev is None else 1)
        p, _ = multigrid.solid_mg_solve(before, grid, solid, cycles, cfg.coarse_solver)
        info["cycles"] = cycles
    elif cfg.kind == "TOLERANCE":
        solve_dt = 1.0
        key = "tol_mg"
        mg = (mg_cache or {}).get(key)
        if mg is None:
            mg = multigrid.MaskedMG(grid, solid, cfg.pre, cfg.post, cfg.coarsest, cfg.restriction)
            if mg_cache is not None:
                mg_cache[key] = mg
        x0 = None
        if cfg.warm_start and prev is not None and state.pressure is not None:
            x0 = state.pressure * (dt / prev)
            # a start no better than zero is dropped (||b - A x0|| >= ||b||)
            if mg.relative_residual(x0, mg.project_mean(before)) >= 1.0:
                x0 = None
        if cfg.fixed_cycles is not None:
            p, hist = mg.solve(before, x0, 0.0, cfg.fixed_cycles)
        else:
            p, hist = mg.solve(before, x0, cfg.tol, cfg.max_cycles)
        info["cycles"] = len(hist) - 1
        info["residual_history"] = hist
    else:
        raise ValueError(cfg.kind)
    if moving:
        corrected = mac.gradient_correct_moving(bounded, p, grid, solve_dt, solid, state.boundary)
    else:
        corrected = mac.gradient_correct(bounded, p, grid, solve_dt, solid)
    after = mac.divergence(corrected, grid, solid)
    info["before"] = before
    info["after"] = after
    state.pressure = p
    state.last_dt = dt
    return corrected, info


def step(state, settings, dt, sources, cfg, multi=False, mg_cache=None):
    """One step: returns nothing, mutates ``state``.

    ``settings`` carries velocity_advection, scalar_advection, thermal_lift,
    density_weight, dissipation, cooling, combustion_enabled, ignition_temperature,
    burn_rate, heat_yield, smoke_yield, vorticity_strength, vorticity_limit
    (physics/pressure.py:7-36 names).
    ``sources`` is a list of ``transport.Src``.
    """
    grid, solid = state.grid, state.solid
    vel = state.velocity
    vmc = settings.velocity_advection == "MACCORMACK"
    pred = transport.face_predictor(vel, grid, dt, solid) if vmc else None
    adv = transport.advect_velocity(vel, grid, dt, vmc, pred, sources, solid)
    w = transport.buoyancy(adv[2], state.density, state.temperature, dt,
                           settings.thermal_lift, settings.density_weight)
    provisional = (adv[0], adv[1], w)
    if getattr(settings, "vorticity_strength", 0.0) > 0:
        # confinement of the provisional field before the projection (backend/projected.py:172-175)
        provisional, _ = transport.confine(provisional, grid, dt, settings.vorticity_strength,
                                           settings.vorticity_limit)
    projected, info = project(state, provisional, dt, cfg, mg_cache)
    smc = settings.scalar_advection == "MACCORMACK"
    fields = [state.density, state.temperature]
    decays = [settings.dissipation, settings.cooling, 0.0]
    if settings.combustion_enabled:
        fields.append(state.fuel)
    new = transport.advect_scalars(fields, projected, grid, dt, smc, decays, sources, multi, solid)
    state.velocity = projected
    state.density, state.temperature = new[0], new[1]
    if settings.combustion_enabled:
        fuel, t, rho, flame = react(new[2], new[1], new[0], dt, settings.ignition_temperature,
                                    settings.burn_rate, settings.heat_yield, settings.smoke_yield)
        state.fuel, state.temperature, state.density, state.flame = fuel, t, rho, flame
    state.info = info
    return info
