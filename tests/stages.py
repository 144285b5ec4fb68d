"""Stage-by-stage parity of a HIP smoke step against the FP64 oracle.

Each step of ``DenseProjectedSmoke`` is checked in four stages, each fed the
same inputs on both sides: the state the HIP step starts from is downloaded
and given to the oracle, and the HIP step's intermediate fields (read through
``DenseProjectedSmoke.observer``) are the oracle's inputs for the next stage.
So every comparison measures one stage's FP32 error, not errors carried over
from earlier steps or stages:

1. velocity transport + emission + buoyancy (+ confinement): HIP provisional
   velocity vs the oracle's from the same start state;
2. projection: HIP projected velocity and pressure vs the oracle's projection
   of the HIP provisional velocity, from the same warm start and with the
   same number of sweeps or V-cycles (the pressure, a potential, to the FP32
   forward-error bound of ``tests.parity.potential_tol``);
3. scalar transport: HIP transported density/temperature/fuel vs the oracle's
   transport of the same start scalars by the HIP projected velocity;
4. combustion: HIP end state vs the oracle's reaction of the HIP transported
   scalars.

With vorticity confinement, stage 1 is split at the confinement: the HIP
transported velocity against the oracle's transport, then the HIP confined
velocity against the oracle's confinement of the HIP transported velocity
(with the allowance of ``oracle.decisions.confinement_allowance`` where the
force direction is set by rounding).

The bar of ``tests.parity.pointwise`` applies to every field: 1e-5 of the
field's largest magnitude, plus, at a point within rounding distance of a
transport decision, the largest change that decision could make
(``oracle.decisions``).
"""
import dataclasses

import numpy as np

from oracle import decisions as dec, step as ostep, transport as otr
from tests.parity import EPS, TOL, gauge_free, pointwise, potential_tol, to_np


def _faces(og, fields):
    return tuple(to_np(f, og.face_array_shape(a)) for a, f in enumerate(fields))


def _cells(og, f):
    return to_np(f, og.array_shape)


class Recorder:
    def __init__(self, og):
        self.og = og
        self.stages = {}

    def __call__(self, stage, fields):
        if stage in ("advected", "provisional"):
            self.stages[stage] = _faces(self.og, fields)
        else:
            self.stages[stage] = [_cells(self.og, f) for f in fields]


def step_stagewise(sim, og, settings, dt, srcs, cfg, *, solid=None, multi_sources=None, eps=EPS,
                   pressure_tol=TOL, converged=False, boundary=None):
    """Run one HIP step and check its four stages; returns the reports.

    ``converged``: the HIP projection runs a solver the oracle does not
    restate step by step (16-bit V-cycles, the learned preconditioner); its
    result is compared with the FP64 oracle's converged projection of the
    same provisional velocity, to ``pressure_tol``.  ``boundary``: the wall
    velocity of moving colliders per face axis (the HIP step's own)."""
    s = settings
    vel0 = _faces(og, sim._velocity)
    rho0, t0 = _cells(og, sim.density), _cells(og, sim.temperature)
    fuel0 = _cells(og, sim.combustion.fuel) if sim.combustion else None
    p0 = _cells(og, sim.projector.pressure)
    last_dt = sim.projector.last_dt
    rec = Recorder(og)
    sim.observer = rec
    try:
        sim.step(dt, sources=multi_sources)
    finally:
        sim.observer = None
    reports = []
    # 1. velocity transport, emission, buoyancy, confinement
    vmc = s.velocity_advection == "MACCORMACK"
    pred = otr.face_predictor(vel0, og, dt, solid) if vmc else None
    adv = otr.advect_velocity(vel0, og, dt, vmc, pred, srcs, solid)
    prov = (adv[0], adv[1], otr.buoyancy(adv[2], rho0, t0, dt, s.thermal_lift, s.density_weight))
    confined = getattr(s, "vorticity_strength", 0.0) > 0
    allow = dec.velocity_allowance(vel0, og, dt, solid, vmc, eps)
    advected = rec.stages["advected" if confined else "provisional"]
    for a in range(3):
        reports.append(pointwise(advected[a], prov[a], allow[a], label=f"advected {'uvw'[a]}"))
    if confined:
        # 1b. confinement of the HIP transported field (a continuous map)
        confined_o, _ = otr.confine(advected, og, dt, s.vorticity_strength, s.vorticity_limit)
        callow = dec.confinement_allowance(advected, og, dt, s.vorticity_strength, s.vorticity_limit)
        for a in range(3):
            reports.append(pointwise(rec.stages["provisional"][a], confined_o[a], callow[a],
                                     label=f"confined {'uvw'[a]}"))
    # 2. projection of the HIP provisional velocity from the same warm start
    if cfg.kind == "TOLERANCE" and cfg.fixed_cycles is None and not converged:
        # the same number of V-cycles as the HIP solve took to reach tolerance
        cfg = dataclasses.replace(cfg, fixed_cycles=int(sim.measure_projection()["v_cycles"]))
    if converged:
        # a solver of another precision: against the FP64 solution itself
        cfg = dataclasses.replace(cfg, kind="TOLERANCE", tol=1e-11, max_cycles=300, fixed_cycles=None)
    state = ostep.State(og, vel0, rho0, t0, solid=solid, pressure=p0, last_dt=last_dt, boundary=boundary)
    projected, info = ostep.project(state, rec.stages["provisional"], dt, cfg)
    vel1 = _faces(og, sim._velocity)
    for a in range(3):
        reports.append(pointwise(vel1[a], projected[a], tol=pressure_tol, label=f"projected {'uvw'[a]}"))
    fluid = None if solid is None else ~solid
    reports.append(pointwise(gauge_free(_cells(og, sim.projector.pressure), fluid), gauge_free(state.pressure, fluid),
                             tol=pressure_tol if converged else potential_tol(og.shape), label="pressure"))
    # 3. scalar transport by the HIP projected velocity
    smc = s.scalar_advection == "MACCORMACK"
    fields = [rho0, t0] + ([fuel0] if fuel0 is not None else [])
    decays = [s.dissipation, s.cooling, 0.0]
    new = otr.advect_scalars(fields, vel1, og, dt, smc, decays, srcs, multi_sources is not None, solid)
    allow = dec.scalar_allowance(fields, vel1, og, dt, solid, smc, eps)
    names = ["density", "temperature", "fuel"]
    for k in range(len(fields)):
        reports.append(pointwise(rec.stages["transported"][k], new[k], allow[k], label=f"transported {names[k]}"))
    # 4. combustion of the HIP transported scalars
    if sim.combustion:
        tr_rho, tr_t, tr_fuel = rec.stages["transported"]
        fuel, temp, rho, flame = ostep.react(tr_fuel, tr_t, tr_rho, dt, s.ignition_temperature, s.burn_rate,
                                             s.heat_yield, s.smoke_yield)
        # the ignition test compares the same fp32 temperature on both sides; only
        # a temperature within fp32 rounding of the threshold can branch apart
        edge = np.abs(tr_t - s.ignition_temperature) <= 1e-7 * max(1.0, abs(s.ignition_temperature))
        for name, got, ref in (("burnt fuel", _cells(og, sim.combustion.fuel), fuel),
                               ("burnt temperature", _cells(og, sim.temperature), temp),
                               ("burnt density", _cells(og, sim.density), rho),
                               ("flame", _cells(og, sim.combustion.flame), flame)):
            reports.append(pointwise(got, ref, np.where(edge, np.abs(ref) + np.abs(got), 0.0), label=name))
    return reports, info
