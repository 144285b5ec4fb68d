"""Whole-step parity: DenseProjectedSmoke on libfluxhip against the FP64
oracle, stage by stage on identical inputs (tests/stages.py), every field to
1e-5 of its largest magnitude except the faces/cells within 1e-4 cells of a
MacCormack or solid decision (tests/parity.py)."""
import numpy as np
import pytest

from oracle import mac, transport as otr, step as ostep
from tests import stages
from tests.helpers import faces_np, cells_np, sphere_mask
from tests.parity import pointwise

pytestmark = pytest.mark.gpu


def _setup(device, shape, settings, solid=None, seed=0, multi_sources=None):
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    grid = GridSpec(shape, (1., 1., 1.), max_dim=512)
    og = mac.Grid(shape, (1., 1., 1.))
    sim = DenseProjectedSmoke(grid, settings, device, solid_mask=None if solid is None else solid.ravel())
    rng = np.random.default_rng(seed)
    vel = tuple(rng.uniform(-.05, .05, og.face_array_shape(a)).astype(np.float32).astype(np.float64)
                for a in range(3))
    sim.upload_velocity([v.ravel() for v in vel])
    sim.upload(np.zeros(og.array_shape).ravel())
    sim.upload_temperature(np.zeros(og.array_shape).ravel())
    state = ostep.State(og, mac.walls(vel, og, solid) if solid is not None else vel,
                        np.zeros(og.array_shape), np.zeros(og.array_shape), solid=solid)
    s = settings
    if s.combustion_enabled:
        fuel = rng.uniform(0, 1, og.array_shape).astype(np.float32).astype(np.float64)
        sim.upload_fuel(fuel.ravel())
        state.fuel = fuel
    if multi_sources is None:
        srcs = [otr.Src(s.source_center, s.source_radius, s.source_rate, s.heat_source_rate,
                        target=s.density_mode == "TARGET", solid_profile=s.source_profile == "SOLID",
                        fuel_rate=s.fuel_source_rate if s.combustion_enabled else 0.0)]
    else:
        srcs = multi_sources[1]
    return sim, state, og, srcs


def _config(s, solid):
    if s.pressure_tolerance > 0:
        return ostep.PressureConfig("TOLERANCE", tol=s.pressure_tolerance, pre=s.mg_pre, post=s.mg_post,
                                    coarsest=s.mg_coarsest)
    if s.pressure_solver == "JACOBI":
        return ostep.PressureConfig("SOLID_JACOBI" if solid is not None else "JACOBI",
                                    iterations=s.pressure_iterations, cold_multiplier=s.pressure_cold_multiplier)
    return ostep.PressureConfig("MULTIGRID", cycles=s.pressure_cycles)


def _stagewise(device, shape, settings, solid=None, steps=3, dt=1 / 128, seed=0, multi_sources=None, **kw):
    """Every step of the HIP run checked stage by stage (tests/stages.py)."""
    sim, _, og, srcs = _setup(device, shape, settings, solid, seed, multi_sources)
    cfg = _config(settings, solid)
    for _ in range(steps):
        stages.step_stagewise(sim, og, settings, dt, srcs, cfg, solid=solid,
                              multi_sources=multi_sources[0] if multi_sources else None, **kw)
    return sim, og


def _trajectory(device, shape, settings, solid=None, steps=3, dt=1 / 128, seed=0):
    """HIP and oracle each run the whole trajectory from the same start."""
    sim, state, og, srcs = _setup(device, shape, settings, solid, seed)
    cfg = _config(settings, solid)
    for _ in range(steps):
        sim.step(dt)
        ostep.step(state, settings, dt, srcs, cfg)
    pairs = {"density": (cells_np(og, sim.density), state.density),
             "temperature": (cells_np(og, sim.temperature), state.temperature)}
    for a, (g, o) in enumerate(zip(faces_np(og, sim._velocity), state.velocity)):
        pairs["uvw"[a]] = (g, o)
    reports = {k: pointwise(g, o, label=k) for k, (g, o) in pairs.items()}
    # the pressure is the iterate of a Jacobi solve that stops short of
    # convergence and starts from the previous step's iterate, so the FP32
    # rounding of every earlier sweep is still in it (measured 4.3e-5 after 3
    # steps); each step's own solve matches to 1e-7 (the stage-wise tests)
    reports["pressure"] = pointwise(cells_np(og, sim.projector.pressure), state.pressure, tol=1e-4, label="pressure")
    return reports


def _settings(**kw):
    from fluxfx_amd.physics.pressure import PressureSettings
    base = dict(source_center=(.5, .5, .16), source_radius=.1, source_rate=1.0, heat_source_rate=1.0,
                dissipation=0.0, cooling=0.0, thermal_lift=1.0, density_weight=0.0, source_profile="SOLID",
                initial_temperature=0.0)
    base.update(kw)
    return PressureSettings(**base)


MC = dict(velocity_advection="MACCORMACK", scalar_advection="MACCORMACK")
TOL_MG = dict(pressure_tolerance=1e-5, mg_pre=2, mg_post=2, mg_coarsest=8)


def test_step_jacobi_semi_lagrangian_trajectory(device):
    """Semi-Lagrangian transport and Jacobi make no discrete choice, so the
    whole 3-step trajectory stays within 1e-5 of the oracle's."""
    s = _settings(pressure_solver="JACOBI", pressure_iterations=64, pressure_cold_multiplier=1)
    _trajectory(device, (24, 24, 24), s, steps=3)


def test_step_jacobi_stagewise(device):
    s = _settings(pressure_solver="JACOBI", pressure_iterations=64, pressure_cold_multiplier=1)
    _stagewise(device, (24, 24, 24), s)


def test_step_jacobi_maccormack(device):
    s = _settings(pressure_solver="JACOBI", pressure_iterations=64, pressure_cold_multiplier=1, **MC)
    _stagewise(device, (24, 24, 24), s)


def test_step_reference_multigrid(device):
    s = _settings(pressure_solver="MULTIGRID", pressure_cycles=2, **MC)
    sim, _ = _stagewise(device, (32, 32, 32), s)
    assert sim.measure_projection()["v_cycles"] == 2


def test_step_tolerance_with_solids_and_combustion(device):
    """Tolerance multigrid (the oracle runs the V-cycles the HIP solve took),
    a sphere, MacCormack, emitter fuel and the fused combustion pass."""
    og = mac.Grid((32, 32, 32))
    solid = sphere_mask(og, (16, 16, 18), 5)
    s = _settings(combustion_enabled=True, ignition_temperature=.3, burn_rate=15., heat_yield=1., smoke_yield=.5,
                  fuel_source_rate=1.0, **TOL_MG, **MC)
    sim, _ = _stagewise(device, (32, 32, 32), s, solid=solid)
    m = sim.measure_projection()
    assert m["converged"] and m["relative_residual"] <= 1e-5


def test_step_vorticity_confinement(device):
    """Confinement of the provisional field before the projection
    (backend/projected.py:172-175), with a solid and MacCormack transport."""
    og = mac.Grid((32, 32, 32))
    solid = sphere_mask(og, (16, 15, 18), 5)
    s = _settings(vorticity_strength=4.0, vorticity_limit=.5, **TOL_MG, **MC)
    sim, _ = _stagewise(device, (32, 32, 32), s, solid=solid, seed=7)
    assert sim.detail.curl is not None and sim.measure_projection()["converged"]


# 16-bit V-cycles and FPCG are other solvers than the oracle's FP32 V-cycle, so
# their projection is held against the FP64 solution of the same provisional
# velocity (converged to 1e-11).  The HIP solve stops at ||b - A q|| <= 1e-5
# ||b||, and the error that residual leaves is what this compares: measured
# worst 8.7e-5 for the pressure and 2.1e-5 for the corrected faces (32^3 steps
# with a sphere, every precision and Krylov choice), 6.9e-5 / 3.7e-5 in the
# 64x48x40 solver tests, relative to each field's largest magnitude.
MIXED_TOL = 2e-4


def test_step_fire_combustion_vorticity_solid(device):
    """The fire workload's step on a small grid: fuel and heat emitter,
    combustion, vorticity confinement and a sphere, BF16 V-cycle FPCG."""
    og = mac.Grid((32, 32, 32))
    solid = sphere_mask(og, (15, 17, 20), 4)
    dt = 1 / 64
    s = _settings(combustion_enabled=True, ignition_temperature=.3, burn_rate=-np.log(.75) / dt, heat_yield=1.,
                  smoke_yield=.5, source_rate=0.0, fuel_source_rate=1.0 / dt, heat_source_rate=1.0 / dt,
                  vorticity_strength=.3, pressure_precision="BF16", pressure_krylov="PCG", **TOL_MG, **MC)
    sim, _ = _stagewise(device, (32, 32, 32), s, solid=solid, dt=dt, seed=11, converged=True,
                        pressure_tol=MIXED_TOL)
    assert sim.measure_projection()["converged"]


@pytest.mark.parametrize("precision,krylov", [("BF16", "NONE"), ("BF16", "PCG"), ("FP16", "NONE"),
                                              ("FP16", "PCG"), ("FP32", "PCG")])
def test_step_mixed_precision_matches_fp64_oracle(device, precision, krylov):
    """Mixed-precision projection inside full MacCormack steps with a solid:
    transport stages to the FP32 bar, the projection against the FP64
    solution to MIXED_TOL, and the projected divergence at the solver's
    tolerance."""
    og = mac.Grid((32, 32, 32))
    solid = sphere_mask(og, (16, 16, 18), 5)
    s = _settings(pressure_precision=precision, pressure_krylov=krylov, **TOL_MG, **MC)
    sim, _ = _stagewise(device, (32, 32, 32), s, solid=solid, converged=True, pressure_tol=MIXED_TOL)
    m = sim.measure_projection()
    assert m["converged"] and m["relative_residual"] <= 1e-5
    after = cells_np(og, sim.projector.after)
    before = cells_np(og, sim.projector.before)
    assert np.sqrt((after ** 2).sum()) <= 2e-5 * np.sqrt((before ** 2).sum())


def test_config1_plume_64_jacobi_16_steps(device):
    """The 64^3 closed-box plume (emitter radius 6 cells at cells (32, 32, 10),
    density and temperature +1 per step, seeded U(-0.05, 0.05) velocity inside
    the emitter, buoyancy 1.0*T, MacCormack everywhere, Jacobi 128 iterations,
    dt = 1/128) over 16 steps; steps 1, 8 and 16 are checked stage by stage
    against the FP64 oracle."""
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    n, dt, steps = 64, 1 / 128, 16
    og = mac.Grid((n, n, n))
    c = (32 / n, 32 / n, 10 / n)
    s = _settings(source_center=c, source_radius=6 / n, source_rate=1 / dt, heat_source_rate=1 / dt,
                  pressure_solver="JACOBI", pressure_iterations=128, pressure_cold_multiplier=1, **MC)
    rng = np.random.default_rng(1)
    vel = []
    for a in range(3):
        sz, sy, sx = og.face_array_shape(a)
        z, y, x = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
        off = [.5, .5, .5]
        off[2 - a] = 0.0
        d2 = ((x + off[2]) / n - c[0]) ** 2 + ((y + off[1]) / n - c[1]) ** 2 + ((z + off[0]) / n - c[2]) ** 2
        vel.append(np.where(d2 < (6 / n) ** 2, rng.uniform(-.05, .05, (sz, sy, sx)), 0.0).astype(np.float32)
                   .astype(np.float64))
    sim = DenseProjectedSmoke(GridSpec((n, n, n)), s, device)
    sim.upload_velocity([v.ravel() for v in vel])
    sim.upload(np.zeros(n ** 3))
    sim.upload_temperature(np.zeros(n ** 3))
    srcs = [otr.Src(c, 6 / n, 1 / dt, 1 / dt, solid_profile=True)]
    cfg = ostep.PressureConfig("JACOBI", iterations=128, cold_multiplier=1)
    for k in range(1, steps + 1):
        if k in (1, 8, 16):
            stages.step_stagewise(sim, og, s, dt, srcs, cfg)
        else:
            sim.step(dt)
    assert cells_np(og, sim.density).max() > .5


def test_step_multi_sources_order_independent(device):
    from fluxfx_amd.physics.emission import Source, Emission
    s = _settings(pressure_solver="JACOBI", pressure_iterations=32, pressure_cold_multiplier=1)
    recs = [Source((.5, .5, .3), .15, 1, .5, Emission((.5, .5, .3)), "TARGET", "SOLID"),
            Source((.45, .5, .3), .12, 2, 1.0, Emission((.4, .5, .25), (0, 0, 1), 4.0), "RATE", "SOFT")]
    osrc = [otr.Src(r.center, r.radius, r.density_rate, r.heat_rate, r.motion.start, r.motion.velocity,
                    r.motion.coupling, r.density_mode == "TARGET", r.profile == "SOLID", r.fuel_rate) for r in recs]
    a, og = _stagewise(device, (20, 20, 20), s, steps=2, multi_sources=(recs, osrc))
    b, _ = _stagewise(device, (20, 20, 20), s, steps=2, multi_sources=(recs[::-1], osrc[::-1]))
    assert np.abs(cells_np(og, a.density) - cells_np(og, b.density)).max() < 1e-6


def test_fault_flag_and_errors(device):
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    sim = DenseProjectedSmoke(GridSpec((16, 16, 16)), _settings(pressure_solver="JACOBI"), device)
    with pytest.raises(ValueError):
        sim.step(0.2)
    with pytest.raises(ValueError):
        sim.upload([-1.0] * 16 ** 3)
    with pytest.raises(RuntimeError):
        sim.projector.metrics()
    sim.step(1 / 60)
    assert sim.projector.ready and sim.steps == 1


@pytest.mark.parametrize("solver", ["JACOBI", "MULTIGRID"])
def test_checkpoint_resume_is_bit_exact(device, tmp_path, solver):
    """fluxfx-checkpoint-1: save after some steps, reload, and the resumed run
    matches the uninterrupted one bit for bit (deterministic kernels)."""
    from fluxfx_amd.backend.checkpoint import save_checkpoint, load_checkpoint, advance_fixed
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    from fluxfx_amd.physics.config import GridSpec
    s = _settings(pressure_solver=solver, pressure_iterations=40, velocity_advection="MACCORMACK",
                  scalar_advection="MACCORMACK", combustion_enabled=True, ignition_temperature=.3, burn_rate=15.,
                  heat_yield=1., smoke_yield=.5, fuel_source_rate=1.0)
    grid = GridSpec((16, 16, 16))
    sim = DenseProjectedSmoke(grid, s, device)
    advance_fixed(sim, 3, 1 / 64)
    path = tmp_path / "state.npz"
    save_checkpoint(sim, path)
    advance_fixed(sim, 2, 1 / 64)
    resumed = load_checkpoint(path, device)
    assert resumed.steps == 3 and resumed.time == pytest.approx(3 / 64)
    advance_fixed(resumed, 2, 1 / 64)
    pairs = [(sim.density, resumed.density), (sim.temperature, resumed.temperature),
             (sim.combustion.fuel, resumed.combustion.fuel), *zip(sim._velocity, resumed._velocity)]
    for a, b in pairs:
        assert np.array_equal(a.tensor.cpu().numpy(), b.tensor.cpu().numpy())
    sim.close()
    resumed.close()
