"""Moving rigid colliders on the GPU (backend/moving_collision.py): whole
steps with a collider sweeping through the plume against the FP64 oracle, and
the gates of scripts/moving_collider_validate.py:20-66."""
import numpy as np
import pytest

from oracle import mac, step as ostep
from tests.helpers import faces_np, cells_np, rel_err

pytestmark = pytest.mark.gpu

N = 32


def _settings():
    from fluxfx_amd.physics.pressure import PressureSettings
    return PressureSettings(moving_colliders=True, source_rate=0, heat_source_rate=0, initial_temperature=0,
                            thermal_lift=0, density_weight=0, dissipation=0, cooling=0,
                            scalar_advection="MACCORMACK", velocity_advection="MACCORMACK", pressure_iterations=120)


def _rows_fn(old, new, dt):
    from fluxfx_amd.physics.collision import Collider
    from fluxfx_amd.physics.collider_motion import velocity_rows
    return velocity_rows(Collider(*old), Collider(*new), dt)


@pytest.mark.parametrize("kind", ["SPHERE", "BOX"])
def test_moving_collider_matches_oracle_and_reference_gates(device, kind):
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.physics.collision import primitive
    from fluxfx_amd.physics.collider_motion import MotionPath
    print(f"moving collider: {kind}", flush=True)
    s = _settings()
    grid = GridSpec((N,) * 3)
    og = mac.Grid((N,) * 3)
    initial = primitive(kind, (.35, .5, .5), (.12, .16, .12))
    target = primitive(kind, (.65, .5, .5), (.12, .16, .12))
    sim = DenseProjectedSmoke(grid, s, device, colliders=(initial,))
    sim.upload(np.ones(N ** 3))
    sim.upload_temperature(np.full(N ** 3, 10.0))
    path = MotionPath((initial,), (target,), .3)
    dt = path.limit_dt(1 / 60, grid.cell_size)
    pair = lambda cs: [(c.shape, c.inverse_rows) for c in cs]
    solid0 = mac.raster_colliders(og, pair([initial]))
    st = ostep.State(og, og.zeros_faces(), np.where(solid0, 0.0, 1.0), np.where(solid0, 0.0, 10.0), solid=solid0)
    cfg = ostep.PressureConfig("SOLID_JACOBI", iterations=s.pressure_iterations,
                               cold_multiplier=s.pressure_cold_multiplier)
    prev = pair([initial])
    for k in range(3):
        cur_c = path.at((k + 1) * dt)
        old_mask = sim.solids.solid.cpu().numpy().reshape(og.array_shape) > 0
        sim.move_colliders(cur_c, dt)
        ostep.move_colliders(st, prev, pair(cur_c), dt, _rows_fn)
        mask = sim.solids.solid.cpu().numpy().reshape(og.array_shape) > 0
        assert np.array_equal(mask, st.solid)
        if k == 0:
            blocked = mask | old_mask
            assert np.all(cells_np(og, sim.density)[blocked] == 0) and np.all(cells_np(og, sim.temperature)[blocked] == 0)
            assert (old_mask & ~mask).any()
        for g, o in zip(faces_np(og, sim.solids.wall_fields), st.boundary):
            assert np.abs(g - o).max() < 1e-5 * max(1.0, np.abs(o).max())
        sim.step(dt)
        ostep.step(st, s, dt, [], cfg)
        prev = pair(cur_c)
    # fields against the FP64 oracle
    errs = {"density": rel_err(cells_np(og, sim.density), st.density),
            "temperature": rel_err(cells_np(og, sim.temperature), st.temperature)}
    for a, (g, o) in enumerate(zip(faces_np(og, sim._velocity), st.velocity)):
        errs["uvw"[a]] = rel_err(g, o)
    assert all(v < 5e-3 for v in errs.values()), errs
    # closed faces carry the wall velocity exactly; the fluid is pushed
    opens = mac.face_open(og, mask)
    walls = faces_np(og, sim.solids.wall_fields)
    vel = faces_np(og, sim._velocity)
    err = max(np.abs(v[~o] - b[~o]).max() for v, o, b in zip(vel, opens, walls))
    peak = max(np.abs(v[o]).max() for v, o in zip(vel, opens))
    assert err < 1e-6 and peak > .02, (err, peak)
    m = sim.measure_projection()
    assert m["solver"] == "MOVING_JACOBI" and m["rms_after"] < m["rms_before"]
    # a collider that stops has zero wall velocity
    sim.move_colliders(sim.solids.colliders, dt)
    assert max(np.abs(w).max() for w in faces_np(og, sim.solids.wall_fields)) < 1e-6
    sim.close()


def test_moving_colliders_select_tolerance_projector_and_reject_static_use(device):
    from fluxfx_amd.backend.moving_collision import MovingToleranceProjector
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.physics.collision import primitive
    from fluxfx_amd.physics.pressure import PressureSettings
    grid = GridSpec((16,) * 3)
    c = primitive("SPHERE", (.5, .5, .5), (.2,) * 3)
    tol = DenseProjectedSmoke(grid, PressureSettings(moving_colliders=True, pressure_tolerance=1e-5), device,
                              colliders=(c,))
    assert isinstance(tol.projector, MovingToleranceProjector)
    tol.close()
    sim = DenseProjectedSmoke(grid, PressureSettings(), device, colliders=(c,))
    with pytest.raises(ValueError):
        sim.move_colliders((c,), .01)
