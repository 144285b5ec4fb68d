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


def test_box_update_and_remap_equal_the_whole_grid_passes(device):
    """The per-step update and remap run over the box the colliders' old and
    new poses cover (fx_moving_update_box, fx_remap_box); the whole-grid
    passes (fx_owner_raster, fx_build_flags, fx_build_clearance,
    fx_wall_velocity, fx_remap_scalar, fx_remap_faces) on the same inputs give
    the same bits everywhere, for a sphere and a turning box that approach each
    other, one of them leaving the grid."""
    import ctypes as C
    from fluxfx_amd import _lib
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    from fluxfx_amd.physics.collision import Collider, primitive
    from fluxfx_amd.physics.collider_motion import velocity_rows
    from fluxfx_amd.physics.config import GridSpec
    torch = device.torch
    n = (40, 48, 36)
    grid = GridSpec(n, (1.0, 1.2, 0.9))
    s = _settings()
    e = np.array(grid.extent)

    def turned_box(center, half, angle):
        c, sn = np.cos(angle), np.sin(angle)
        rows = []
        for a, h in zip([(c, sn, 0.0), (-sn, c, 0.0), (0.0, 0.0, 1.0)], half):
            r = [v / h for v in a]
            rows.append((r[0], r[1], r[2], -(r[0] * center[0] + r[1] * center[1] + r[2] * center[2])))
        return Collider("BOX", tuple(rows))

    pose = lambda k: (primitive("SPHERE", tuple(e * (.3 + .03 * k, .5, .45)), (.16,) * 3),
                      turned_box(tuple(e * (.75, .5 - .02 * k, .1 - .02 * k)), (.12, .2, .1), .2 * k))
    sim = DenseProjectedSmoke(grid, s, device, colliders=pose(0))
    rng = np.random.default_rng(4)
    sim.upload(rng.uniform(0, 1, grid.cells))
    sim.upload_temperature(rng.uniform(0, 1, grid.cells))
    sim.upload_velocity([rng.uniform(-1, 1, int(np.prod(sh))) for sh in grid.face_shapes])
    g = _lib.make_grid(grid.shape, grid.extent)
    st = device.stream
    dt = .01
    for k in range(1, 5):
        print(f"box update: move {k}", flush=True)
        old = sim.solids.colliders
        new = pose(k)
        prev = sim.solids.solid.clone()
        vel = [f.tensor.clone() for f in sim._velocity]
        rho, temp = sim.density.tensor.clone(), sim.temperature.tensor.clone()
        sim.move_colliders(new, dt)
        # the whole-grid passes on the same inputs
        owner = torch.zeros_like(prev)
        for j, c in enumerate(new, start=1):
            rows = (C.c_float * 12)(*[float(v) for r in c.inverse_rows for v in r])
            _lib.call("fx_owner_raster", g, 0 if c.shape == "SPHERE" else 1, rows, j, owner.data_ptr(), st)
        flags = torch.zeros_like(prev)
        _lib.call("fx_build_flags", g, owner.data_ptr(), flags.data_ptr(), st)
        clear, scratch = torch.zeros_like(prev), torch.zeros_like(prev)
        _lib.call("fx_build_clearance", g, owner.data_ptr(), clear.data_ptr(), scratch.data_ptr(), st)
        walls = [torch.zeros_like(v) for v in vel]
        table = [v for a, b in zip(old, new) for row in velocity_rows(a, b, dt) for v in row]
        _lib.call("fx_wall_velocity", g, owner.data_ptr(), len(new), (C.c_float * len(table))(*table),
                  *(w.data_ptr() for w in walls), st)
        outs = [torch.empty_like(v) for v in vel]
        _lib.call("fx_remap_faces", g, owner.data_ptr(), prev.data_ptr(), *(w.data_ptr() for w in walls),
                  *(v.data_ptr() for v in vel), *(o.data_ptr() for o in outs), st)
        rho_o, temp_o = torch.empty_like(rho), torch.empty_like(temp)
        _lib.call("fx_remap_scalar", g, owner.data_ptr(), prev.data_ptr(), rho.data_ptr(), rho_o.data_ptr(), st)
        _lib.call("fx_remap_scalar", g, owner.data_ptr(), prev.data_ptr(), temp.data_ptr(), temp_o.data_ptr(), st)
        torch.cuda.synchronize()
        sol = sim.solids
        assert torch.equal(sol.solid, owner) and torch.equal(sol.flags, flags) and torch.equal(sol.clearance, clear)
        for a in range(3):
            assert torch.equal(sol.wall_fields[a].tensor, walls[a]), a
            assert torch.equal(sim._velocity[a].tensor, outs[a]), a
        assert torch.equal(sim.density.tensor, rho_o) and torch.equal(sim.temperature.tensor, temp_o)
        assert torch.equal(sol.previous, prev)
        sim.step(dt)
    sim.close()
