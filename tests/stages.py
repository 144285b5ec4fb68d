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
   same number of sweeps or V-cycles;
3. scalar transport: HIP transported density/temperature/fuel vs the oracle's
   transport of the same start scalars by the HIP projected velocity;
4. combustion: HIP end state vs the oracle's reaction of the HIP transported
   scalars.

The bar of ``tests.parity.pointwise`` applies to every field.  Fragile points
(within ``eps`` cells of a transport decision) may differ by at most the
spread of the donors they could have chosen; with vorticity confinement the
fragile set is widened by the confinement stencil (two cells), since the
force at a face reads the curl of its neighbours.
"""
import dataclasses

import numpy as np

from oracle import decisions as dec, step as ostep, transport as otr
from tests.parity import TOL, gauge_free, pointwise, to_np


def _faces(og, fields):
    return tuple(to_np(f, og.face_array_shape(a)) for a, f in enumerate(fields))


def _cells(og, f):
    return to_np(f, og.array_shape)


def _dilate(mask, r):
    out = mask.copy()
    for ax in range(3):
        acc = out.copy()
        for s in range(1, r + 1):
            acc |= np.roll(out, s, ax) | np.roll(out, -s, ax)
        out = acc
    return out


def _spread(field, q):
    """Width of the donor hull around departure q (face/cell index space)."""
    sz, sy, sx = field.shape
    lo, hi = dec.donor_hull(field, q, (sx, sy, sz))
    return hi - lo


def decisive(margin, spread):
    """Margins of the points whose decision can change the result: where all
    donors a point could choose hold the same value (spread 0) the donor set
    and the correction's range test give the same value either way."""
    if margin is None:
        return None
    return np.where(spread > 0, margin, np.inf)


def _fragile_check(got, ref, margin, spread, eps, label, tol=TOL, max_fragile=5e-3):
    """pointwise() on firm points; fragile points within the donor spread."""
    rep = pointwise(got, ref, margin, eps, tol, max_fragile=max_fragile, label=label)
    if margin is not None:
        fragile = np.asarray(margin) <= eps
        scale = max(float(np.abs(ref).max()), 1e-30)
        excess = fragile & (np.abs(got - ref) > spread + tol * scale)
        rep["beyond_spread"] = int(excess.sum())
        assert not excess.any(), rep
    return rep


class Recorder:
    def __init__(self, og):
        self.og = og
        self.stages = {}

    def __call__(self, stage, fields):
        if stage == "provisional":
            self.stages[stage] = _faces(self.og, fields)
        else:
            self.stages[stage] = [_cells(self.og, f) for f in fields]


def step_stagewise(sim, og, settings, dt, srcs, cfg, *, solid=None, multi_sources=None, eps=1e-4,
                   pressure_tol=TOL, max_fragile=5e-3):
    """Run one HIP step and check its four stages; returns the reports."""
    s = settings
    vel0 = _faces(og, sim._velocity)
    rho0, t0 = _cells(og, sim.density), _cells(og, sim.temperature)
    fuel0 = _cells(og, sim.combustion.fuel) if sim.combustion else None
    p0 = _cells(og, sim.projector.pressure)
    last_dt = sim.projector.last_dt if sim.projector.ready else None
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
    if confined:
        prov, _ = otr.confine(prov, og, dt, s.vorticity_strength, s.vorticity_limit)
    decided = vmc or solid is not None
    margins = dec.velocity_margins(vel0, og, dt, solid, vmc) if decided else (None,) * 3
    deps = dec.face_departures(vel0, og, dt, solid) if decided else (None,) * 3
    for a in range(3):
        m = margins[a]
        spread = None
        if m is not None:
            spread = _spread(vel0[a], deps[a])
            m = decisive(m, spread)
            if confined:
                frag = _dilate(m <= eps, 2)
                m = np.where(frag, 0.0, m)
                spread = spread + 2 * dt * s.vorticity_limit
        reports.append(_fragile_check(rec.stages["provisional"][a], prov[a], m, spread, eps,
                                      f"provisional {'uvw'[a]}", max_fragile=max_fragile * (25 if confined else 1)))
    # 2. projection of the HIP provisional velocity from the same warm start
    if cfg.kind == "TOLERANCE" and cfg.fixed_cycles is None:
        # the same number of V-cycles as the HIP solve took to reach tolerance
        cfg = dataclasses.replace(cfg, fixed_cycles=int(sim.measure_projection()["v_cycles"]))
    state = ostep.State(og, vel0, rho0, t0, solid=solid, pressure=p0, last_dt=last_dt)
    projected, info = ostep.project(state, rec.stages["provisional"], dt, cfg)
    vel1 = _faces(og, sim._velocity)
    for a in range(3):
        reports.append(pointwise(vel1[a], projected[a], tol=pressure_tol, label=f"projected {'uvw'[a]}"))
    fluid = None if solid is None else ~solid
    reports.append(pointwise(gauge_free(_cells(og, sim.projector.pressure), fluid), gauge_free(state.pressure, fluid),
                             tol=pressure_tol, label="pressure"))
    # 3. scalar transport by the HIP projected velocity
    smc = s.scalar_advection == "MACCORMACK"
    fields = [rho0, t0] + ([fuel0] if fuel0 is not None else [])
    decays = [s.dissipation, s.cooling, 0.0]
    new = otr.advect_scalars(fields, vel1, og, dt, smc, decays, srcs, multi_sources is not None, solid)
    sdec = smc or solid is not None
    sm = dec.scalar_margins(vel1, og, dt, solid, smc) if sdec else None
    cq = dec.cell_departures(vel1, og, dt, solid) if sdec else None
    names = ["density", "temperature", "fuel"]
    for k, f in enumerate(fields):
        spread = _spread(f, cq) if sdec else None
        reports.append(_fragile_check(rec.stages["transported"][k], new[k], decisive(sm, spread), spread, eps,
                                      f"transported {names[k]}", max_fragile=max_fragile))
    # 4. combustion of the HIP transported scalars
    if sim.combustion:
        tr_rho, tr_t, tr_fuel = rec.stages["transported"]
        fuel, temp, rho, flame = ostep.react(tr_fuel, tr_t, tr_rho, dt, s.ignition_temperature, s.burn_rate,
                                             s.heat_yield, s.smoke_yield)
        # the ignition test compares the same fp32 temperature on both sides; only
        # a temperature within fp32 rounding of the threshold can branch apart
        edge = np.abs(tr_t - s.ignition_temperature) <= 1e-7 * max(1.0, abs(s.ignition_temperature))
        margin = np.where(edge, 0.0, np.inf)
        for name, got, ref in (("burnt fuel", _cells(og, sim.combustion.fuel), fuel),
                               ("burnt temperature", _cells(og, sim.temperature), temp),
                               ("burnt density", _cells(og, sim.density), rho),
                               ("flame", _cells(og, sim.combustion.flame), flame)):
            reports.append(pointwise(got, ref, margin, eps=1.0, label=name))
    return reports, info
