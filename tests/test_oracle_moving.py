"""Moving-collider oracle (backend/moving_collision.py, shaders/moving_*.glsl):
the gates of scripts/moving_collider_validate.py:20-66 on the FP64
restatement, plus closed forms of the wall velocity."""
import numpy as np

from fluxfx_amd.physics.collision import primitive
from fluxfx_amd.physics.collider_motion import MotionPath, velocity_rows
from fluxfx_amd.physics.pressure import PressureSettings
from oracle import mac, step as ostep, transport as otr


def _pairs(colliders):
    return [(c.shape, c.inverse_rows) for c in colliders]


def _rows(old, new, dt):
    from fluxfx_amd.physics.collision import Collider
    return velocity_rows(Collider(*old), Collider(*new), dt)


def test_translating_box_wall_velocity_closed_form():
    """A box translating at 1 m/s along x: every interior face next to it
    carries u = 1 and v = w = 0; faces away from it carry 0."""
    g = mac.Grid((16, 16, 16))
    a = primitive("BOX", (.45, .5, .5), (.15, .2, .1))
    b = primitive("BOX", (.46, .5, .5), (.15, .2, .1))
    owner = mac.raster_owners(g, _pairs([b]))
    bu, bv, bw = mac.wall_velocity(g, owner, [velocity_rows(a, b, .01)])
    solid = owner > 0
    open_u = mac.face_open(g, solid)[0] | ~mac.face_open(g, None)[0]
    assert np.allclose(bu[~open_u], 1.0) and np.all(bu[open_u] == 0.0)
    assert np.abs(bv).max() < 1e-9 and np.abs(bw).max() < 1e-9


def test_owner_is_first_collider_and_matches_union_raster():
    g = mac.Grid((12, 12, 12))
    cs = [primitive("SPHERE", (.4, .5, .5), (.2,) * 3), primitive("BOX", (.6, .5, .5), (.2, .1, .1))]
    owner = mac.raster_owners(g, _pairs(cs))
    assert np.array_equal(owner > 0, mac.raster_colliders(g, _pairs(cs)))
    first = mac.raster_colliders(g, _pairs(cs[:1]))
    assert np.all(owner[first] == 1) and np.all(owner[~first & (owner > 0)] == 2)


def _moving_case(shape_kind, steps=3, n=32):
    g = mac.Grid((n, n, n))
    s = PressureSettings(moving_colliders=True, source_rate=0, heat_source_rate=0, initial_temperature=0,
                         thermal_lift=0, density_weight=0, dissipation=0, cooling=0,
                         scalar_advection="MACCORMACK", velocity_advection="MACCORMACK", pressure_iterations=120)
    initial = primitive(shape_kind, (.35, .5, .5), (.12, .16, .12))
    target = primitive(shape_kind, (.65, .5, .5), (.12, .16, .12))
    path = MotionPath((initial,), (target,), .3)
    dt = path.limit_dt(1 / 60, g.cell_size)
    solid0 = mac.raster_colliders(g, _pairs([initial]))
    st = ostep.State(g, g.zeros_faces(), np.where(solid0, 0.0, 1.0), np.where(solid0, 0.0, 10.0), solid=solid0)
    cfg = ostep.PressureConfig("SOLID_JACOBI", iterations=120, cold_multiplier=1)
    prev = _pairs([initial])
    hist = []
    for k in range(steps):
        cur = _pairs(path.at((k + 1) * dt))
        old_solid = st.solid.copy()
        ostep.move_colliders(st, prev, cur, dt, _rows)
        hist.append((old_solid, st.solid.copy(), st.density.copy(), st.temperature.copy()))
        info = ostep.step(st, s, dt, [], cfg)
        prev = cur
    return g, st, info, hist, path, dt


def test_moving_box_and_sphere_reference_gates():
    for kind in ("SPHERE", "BOX"):
        g, st, info, hist, path, dt = _moving_case(kind, steps=2)
        old, new, rho, heat = hist[0]
        assert np.array_equal(new, mac.raster_colliders(g, _pairs(path.at(dt))))
        blocked = old | new
        assert np.all(rho[blocked] == 0) and np.all(heat[blocked] == 0)
        assert (old & ~new).any()  # the motion exposes cells
        opens = mac.face_open(g, st.solid)
        err = max(np.abs(v[~o] - b[~o]).max() for v, o, b in zip(st.velocity, opens, st.boundary))
        peak = max(np.abs(v[o]).max() for v, o in zip(st.velocity, opens))
        assert err < 1e-12 and peak > .02, (err, peak)
        stats = mac.divergence_stats(info["before"], info["after"])
        assert stats["rms_after"] < stats["rms_before"]


def test_stationary_wall_stops():
    g, st, info, hist, path, dt = _moving_case("BOX", steps=1)
    cur = _pairs(path.at(dt))
    ostep.move_colliders(st, cur, cur, dt, _rows)
    assert max(np.abs(b).max() for b in st.boundary) < 1e-12
