"""The workload configurations at sizes the FP64 oracle checks in seconds,
stage by stage (tests/stages.py) on the HIP run's own states.

* 128^3 static sphere with fixed-budget multigrid, at 64^3: three seeded
  emitters (centres x, y in [8, 56], z in [4, 16] cells, radius 2-4 cells,
  heat rate U[0.5, 1.5]/dt), a sphere of radius 8 at the centre, exactly 8
  V(2,2) cycles per step of the 4-level hierarchy 64 -> 8 with an FP64 8^3
  direct solve, full-weighting and 8-child-mean restriction, 32 steps;
* the 256^3 mixed-precision scene at 64^3 (the benchmark's emitter and its 8
  seeded boxes scaled to 64^3): BF16 and FP16 V-cycles inside FP32 iterative
  refinement, and BF16 V-cycle FPCG, 32 steps to 1e-5 relative residual;
  pressure and corrected faces against the FP64 solution, divergence after
  the projection, every advected field to the FP32 bar;
* the 384^3 fire at 32^3: fuel and heat emitter, combustion (a quarter of
  the fuel above ignition 0.3 burns per step, heat 1.0, smoke 0.5),
  vorticity confinement 0.3 and a sphere moving 0.5 cells per step on a
  seeded straight path (masked Jacobi projector with moving walls).
"""
import numpy as np
import pytest

from oracle import mac, step as ostep, transport as otr
from tests import stages
from tests.helpers import cells_np, faces_np, sphere_mask
from tests.test_gpu_step import MIXED_TOL

pytestmark = pytest.mark.gpu


def _still(sim, n):
    """Zero velocity, density, temperature (and fuel) to start from."""
    sim.upload_velocity([np.zeros(int(np.prod(sh))) for sh in sim.grid.face_shapes])
    sim.upload(np.zeros(n ** 3))
    sim.upload_temperature(np.zeros(n ** 3))
    if sim.combustion:
        sim.upload_fuel(np.zeros(n ** 3))


def _config2_scene(n, seed=2):
    from fluxfx_amd.physics.emission import Emission, Source
    rng = np.random.default_rng(seed)
    dt = 0.5 / n
    s = n / 128
    sources, osrc = [], []
    for _ in range(3):
        c = (rng.uniform(16, 112) * s / n, rng.uniform(16, 112) * s / n, rng.uniform(8, 32) * s / n)
        r = int(rng.integers(4, 9)) * s / n
        heat = rng.uniform(.5, 1.5) / dt
        sources.append(Source(c, r, 1 / dt, heat, Emission(c), "RATE", "SOLID"))
        osrc.append(otr.Src(c, r, 1 / dt, heat, solid_profile=True))
    return sources, osrc, dt


@pytest.mark.parametrize("restriction", ["FULL_WEIGHTING", "MEAN"])
def test_config2_fixed_cycle_multigrid_with_sphere(device, restriction):
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.physics.pressure import PressureSettings
    n, steps = 64, 32
    print(f"config 2 at {n}^3: {restriction}", flush=True)
    og = mac.Grid((n,) * 3)
    solid = sphere_mask(og, (n / 2,) * 3, n / 8)
    sources, osrc, dt = _config2_scene(n)
    s = PressureSettings(dissipation=0.0, cooling=0.0, thermal_lift=1.0, density_weight=0.0,
                         initial_temperature=0.0, velocity_advection="MACCORMACK", scalar_advection="MACCORMACK",
                         pressure_fixed_cycles=8, mg_pre=2, mg_post=2, mg_coarsest=8, mg_restriction=restriction)
    sim = DenseProjectedSmoke(GridSpec((n,) * 3), s, device, solid_mask=solid.ravel())
    _still(sim, n)
    assert [lv["shape"] for lv in sim.projector.levels] == [(64,) * 3, (32,) * 3, (16,) * 3, (8,) * 3]
    cfg = ostep.PressureConfig("TOLERANCE", pre=2, post=2, coarsest=8, restriction=restriction, fixed_cycles=8)
    for k in range(1, steps + 1):
        if k in (1, steps):
            stages.step_stagewise(sim, og, s, dt, osrc, cfg, solid=solid, multi_sources=sources)
        else:
            sim.step(dt, sources=sources)
        m = sim.measure_projection()
        assert m["v_cycles"] == 8 and m["fixed_cycles"] == 8
    assert m["rms_after"] < 1e-2 * m["rms_before"]
    assert cells_np(og, sim.density).max() > .5


@pytest.mark.parametrize("precision,krylov", [("BF16", "NONE"), ("FP16", "NONE"), ("BF16", "PCG")])
def test_config3_mixed_precision_scene(device, precision, krylov):
    import dataclasses
    import bench
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    from fluxfx_amd.physics.config import GridSpec
    n, steps = 64, 32
    print(f"config 3 at {n}^3: {precision} {krylov}", flush=True)
    og = mac.Grid((n,) * 3)
    center, radius, boxes = bench.scene(n)
    solid = bench.box_mask(n, boxes)
    s, dt = bench.settings_for(n, "plain")
    s = dataclasses.replace(s, pressure_precision=precision, pressure_krylov=krylov)
    sim = DenseProjectedSmoke(GridSpec((n,) * 3), s, device, solid_mask=solid.ravel())
    sim.upload_velocity([v.ravel() for v in bench.initial_velocity(n, center, radius)])
    sim.upload(np.zeros(n ** 3))
    sim.upload_temperature(np.zeros(n ** 3))
    srcs = [otr.Src(s.source_center, s.source_radius, s.source_rate, s.heat_source_rate, solid_profile=True)]
    cfg = ostep.PressureConfig("TOLERANCE", tol=1e-5, pre=2, post=2, coarsest=8)
    for k in range(1, steps + 1):
        if k in (1, steps):
            stages.step_stagewise(sim, og, s, dt, srcs, cfg, solid=solid, converged=True, pressure_tol=MIXED_TOL)
        else:
            sim.step(dt)
        m = sim.measure_projection()
        assert m["converged"] and m["relative_residual"] <= 1e-5, (k, m)
        after = cells_np(og, sim.projector.after)
        before = cells_np(og, sim.projector.before)
        assert np.sqrt((after ** 2).sum()) <= 2e-5 * np.sqrt((before ** 2).sum()), k


def test_config5_fire_with_moving_sphere(device):
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    from fluxfx_amd.physics.collider_motion import velocity_rows
    from fluxfx_amd.physics.collision import Collider, primitive
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.physics.pressure import PressureSettings
    n, steps = 32, 6
    print(f"config 5 at {n}^3", flush=True)
    og = mac.Grid((n,) * 3)
    h = 1.0 / n
    dt = 0.5 / n
    rng = np.random.default_rng(5)
    start = np.array([rng.uniform(.3, .4), rng.uniform(.4, .6), rng.uniform(.55, .65)])
    direction = rng.normal(size=3)
    direction /= np.linalg.norm(direction)
    radius = 4 * h
    sphere = lambda k: primitive("SPHERE", tuple(start + direction * 0.5 * h * k), (radius,) * 3)
    s = PressureSettings(source_center=(.5, .5, .2), source_radius=.12, source_rate=0.0, heat_source_rate=1 / dt,
                         fuel_source_rate=1 / dt, combustion_enabled=True, ignition_temperature=.3,
                         burn_rate=-np.log(.75) / dt, heat_yield=1.0, smoke_yield=.5, vorticity_strength=.3,
                         dissipation=0.0, cooling=0.0, thermal_lift=1.0, density_weight=0.0,
                         initial_temperature=0.0, source_profile="SOLID", moving_colliders=True,
                         velocity_advection="MACCORMACK", scalar_advection="MACCORMACK", pressure_iterations=80,
                         pressure_cold_multiplier=1)
    sim = DenseProjectedSmoke(GridSpec((n,) * 3), s, device, colliders=(sphere(0),))
    _still(sim, n)
    srcs = [otr.Src(s.source_center, s.source_radius, 0.0, s.heat_source_rate, solid_profile=True,
                    fuel_rate=s.fuel_source_rate)]
    cfg = ostep.PressureConfig("SOLID_JACOBI", iterations=s.pressure_iterations, cold_multiplier=1)
    pair = lambda c: (c.shape, c.inverse_rows)
    for k in range(1, steps + 1):
        new = sphere(k)
        # the move on both sides from the HIP state: mask, wall velocity, remap
        st = ostep.State(og, faces_np(og, sim._velocity), cells_np(og, sim.density), cells_np(og, sim.temperature),
                         fuel=cells_np(og, sim.combustion.fuel),
                         solid=sim.solids.solid.cpu().numpy().reshape(og.array_shape) > 0)
        prev = pair(sim.solids.colliders[0])
        sim.move_colliders((new,), dt)
        ostep.move_colliders(st, [prev], [pair(new)], dt,
                             lambda a, b, d: velocity_rows(Collider(*a), Collider(*b), d))
        mask = sim.solids.solid.cpu().numpy().reshape(og.array_shape) > 0
        assert np.array_equal(mask, st.solid)
        walls = faces_np(og, sim.solids.wall_fields)
        for g, o in zip(walls, st.boundary):
            assert np.abs(g - o).max() <= 1e-5 * max(1.0, np.abs(o).max())
        for g, o in zip(faces_np(og, sim._velocity), st.velocity):
            assert np.abs(g - o).max() <= 1e-5 * max(np.abs(o).max(), 1e-3)
        for g, o in ((sim.density, st.density), (sim.temperature, st.temperature), (sim.combustion.fuel, st.fuel)):
            assert np.array_equal(cells_np(og, g), o)
        stages.step_stagewise(sim, og, s, dt, srcs, cfg, solid=mask, boundary=walls)
    assert cells_np(og, sim.combustion.flame).max() > 0
