"""Pins the transport/emission/combustion/timestep oracles.

Reference known answers: FluxFX_Source/tests/test_mac.py:13-64,
test_thermal.py:26-37, test_physics.py:20-33, test_combustion.py:11-21,
test_timestep.py:5-34, test_sources.py:7-14, and the CPU-checkable gates of
FluxFX_Source/scripts/detail_validate.py:19-59 (MacCormack beats SL on a
Gaussian, bounded transport, solid target = sphere indicator and no
accumulation, signed heat T = -15 rho, source-order independence).
Independent checks: numerical quadrature of the swept emitter profile and a
fine-stepping brute force for the solid trace clip.
"""
import math

import numpy as np
import pytest

from oracle import mac, transport as tr, step as st


def test_face_shapes_and_positions():
    g = mac.Grid((6, 5, 4), (2.0, 3.0, 4.0))
    assert g.face_shapes == ((7, 5, 4), (6, 6, 4), (6, 5, 5))
    (px, py, pz), _ = tr.face_positions(g, 0)
    assert (px[0, 0, 0], py[0, 0, 0], pz[0, 0, 0]) == (0, .3, .5)
    (px, py, pz), _ = tr.face_positions(g, 2)
    assert (px[0, 0, 0], py[0, 0, 0], pz[0, 0, 0]) == (1 / 6, .3, 0) and pz.max() == 4


def _seed(g, base):
    return tuple(np.full(g.face_array_shape(a), base[a]) for a in range(3))


def test_constant_flow_exact_including_outside_points():
    g = mac.Grid((6, 5, 4), (2.0, 3.0, 4.0))
    vel = _seed(g, (.2, -.3, .4))
    for p in ((.7, 1.3, 2.2), (-10, 20, -1), (2, 3, 4)):
        v = tr.velocity_at(vel, g, *(np.array(c) for c in p))
        assert np.allclose([float(x) for x in v], (.2, -.3, .4), atol=1e-12)
    new = tr.advect_velocity(vel, g, .1)
    for f, v in zip(new, (.2, -.3, .4)):
        assert np.abs(f - v).max() < 1e-12


def test_affine_staggered_reconstruction():
    g = mac.Grid((6, 5, 4), (2.0, 3.0, 4.0))
    fields = []
    for a in range(3):
        (px, py, pz), _ = tr.face_positions(g, a)
        fields.append(px + 2 * py + 3 * pz + a)
    v = tr.velocity_at(tuple(fields), g, np.array(.8), np.array(1.4), np.array(2.1))
    for a in range(3):
        assert abs(float(v[a]) - (.8 + 2 * 1.4 + 3 * 2.1 + a)) < 1e-12


def test_shear_is_stationary_and_zero_stays_zero():
    g = mac.Grid((6, 5, 4), (2.0, 3.0, 4.0))
    (px, py, pz), _ = tr.face_positions(g, 0)
    fields = (py - 1.5, np.zeros(g.face_array_shape(1)), np.zeros(g.face_array_shape(2)))
    for a, b in zip(fields, tr.advect_velocity(fields, g, .1)):
        assert np.abs(a - b).max() < 1e-12
    z = g.zeros_faces()
    assert all(not f.any() for f in tr.advect_velocity(z, g, .1))


def test_component_extrema_do_not_grow():
    g = mac.Grid((6, 5, 4), (2.0, 3.0, 4.0))
    fields = tuple(np.array([((i * 7) % 19 - 9) / 10 for i in range(int(np.prod(g.face_array_shape(a))))])
                   .reshape(g.face_array_shape(a)) for a in range(3))
    for before, after in zip(fields, tr.advect_velocity(fields, g, .07)):
        assert after.min() >= before.min() - 1e-12 and after.max() <= before.max() + 1e-12


def test_one_cell_translation_of_a_scalar():
    g = mac.Grid((8, 4, 4))
    vel = _seed(g, (1.25, 0, 0))
    d = np.zeros(g.array_shape)
    d[1, 1, 2] = 1.0
    out = tr.advect_scalars([d], vel, g, .1)[0]
    assert abs(out[1, 1, 3] - 1) < 1e-12 and abs(out.sum() - 1) < 1e-12


def test_buoyancy_reference_vectors():
    g = mac.Grid((4, 3, 2))
    w = np.zeros(g.face_array_shape(2))
    res = tr.buoyancy(w, np.zeros(g.array_shape), np.full(g.array_shape, 100.), .1, .005, .05)
    assert np.allclose(res, .05)
    temp = np.concatenate([np.zeros((1, 3, 4)), np.full((1, 3, 4), 100.)])
    res = tr.buoyancy(w, np.zeros(g.array_shape), temp, .1, .005, .05)
    assert np.allclose(res[0], 0) and np.allclose(res[1], .025) and np.allclose(res[2], .05)


def test_decay_and_source_scaling():
    g = mac.Grid((8, 6, 4))
    field = np.array([float((i * 7) % 13) for i in range(192)]).reshape(g.array_shape)
    z = g.zeros_faces()
    out = tr.advect_scalars([field], z, g, .1, decays=(.7,))[0]
    assert np.abs(out - field * math.exp(-.07)).max() < 1e-12
    c = tuple(.5 * h for h in g.cell_size)
    out = tr.advect_scalars([np.zeros(g.array_shape)], z, g, .05, sources=[tr.Src(c, .12, 4.0)])[0]
    assert abs(out[0, 0, 0] - .2) < 1e-12


def test_emission_weight_matches_quadrature():
    rng = np.random.default_rng(0)
    for profile in (False, True):
        c, s, R = (.55, .5, .45), (.35, .42, .3), .12
        pts = rng.uniform(.2, .7, (400, 3))
        got = tr.emission_weight(tuple(pts.T), c, s, R, profile)
        t = (np.arange(20000) + .5) / 20000
        path = np.array(s)[None] + t[:, None] * (np.array(c) - np.array(s))[None]
        for k in range(0, 400, 7):
            d2 = ((pts[k][None] - path) ** 2).sum(1) / R ** 2
            vals = (d2 < 1).astype(float) if profile else np.maximum(1 - d2, 0) ** 2
            assert abs(got[k] - vals.mean()) < 2e-4
    # stationary soft profile closed form
    w = tr.emission_weight((np.array(.5), np.array(.5), np.array(.55)), (.5, .5, .5), (.5, .5, .5), .1, False)
    assert abs(float(w) - (1 - .25) ** 2) < 1e-12


def _gaussian_case(mode):
    n = 32
    g = mac.Grid((n,) * 3)
    vel = _seed(g, (.2, 0, 0))
    (px, py, pz), _ = tr.cell_positions(g)
    gauss = lambda c: np.exp(-((px - c[0]) ** 2 + (py - c[1]) ** 2 + (pz - c[2]) ** 2) / (2 * .05 ** 2))
    f = gauss((.3, .5, .5))
    for _ in range(60):
        f = tr.advect_scalars([f], vel, g, 1 / 60, maccormack=(mode == "MC"))[0]
    return f, gauss((.5, .5, .5)), gauss((.3, .5, .5))


def test_maccormack_beats_semi_lagrangian_and_stays_bounded():
    errs = {}
    for mode in ("SL", "MC"):
        f, ref, init = _gaussian_case(mode)
        assert f.min() >= -1e-6 and f.max() <= init.max() + 1e-6
        errs[mode] = np.abs(f - ref).mean()
    assert errs["MC"] < .7 * errs["SL"]


def test_solid_target_equals_sphere_and_does_not_accumulate():
    n = 32
    g = mac.Grid((n,) * 3)
    z = g.zeros_faces()
    src = [tr.Src((.5, .5, .5), .2, 1.0, target=True, solid_profile=True)]
    a = tr.advect_scalars([np.zeros(g.array_shape)], z, g, .04, maccormack=True, sources=src)[0]
    (px, py, pz), _ = tr.cell_positions(g)
    exp = (((px - .5) ** 2 + (py - .5) ** 2 + (pz - .5) ** 2) < .04).astype(float)
    assert np.abs(a - exp).max() < 1e-12
    b = tr.advect_scalars([a], z, g, .04, maccormack=True, sources=src)[0]
    assert np.abs(a - b).max() < 1e-12


def test_signed_heat_single_injection():
    n = 16
    g = mac.Grid((n,) * 3)
    z = g.zeros_faces()
    src = [tr.Src((.5, .5, .16), .12, 2.0, -30.0)]
    rho, temp = tr.advect_scalars([np.zeros(g.array_shape)] * 2, z, g, .04, maccormack=True, sources=src)
    assert rho.min() >= 0 and temp.min() < 0 and np.abs(temp + 15 * rho).max() < 1e-12


def test_mixed_sources_order_independent():
    n = 16
    g = mac.Grid((n,) * 3)
    z = g.zeros_faces()
    recs = [tr.Src((.5,) * 3, .2, 1.0, target=True, solid_profile=True), tr.Src((.5,) * 3, .2, 2.0)]
    a = tr.advect_scalars([np.zeros(g.array_shape)], z, g, .04, sources=recs, multi=True)[0]
    b = tr.advect_scalars([np.zeros(g.array_shape)], z, g, .04, sources=recs[::-1], multi=True)[0]
    assert np.abs(a - b).max() < 1e-15


def test_clip_trace_brute_force():
    rng = np.random.default_rng(3)
    g = mac.Grid((12, 10, 8))
    solid = rng.uniform(size=g.array_shape) < .15
    o = rng.uniform(.05, .95, (200, 3))
    o = o[~tr.solid_at(solid, *(np.floor(o[:, a] * g.shape[a]).astype(int) for a in range(3)))]
    t = o + rng.uniform(-.3, .3, o.shape)
    res = np.stack(tr.clip_solid_trace(solid, g, tuple(o.T), tuple(t.T)), 1)
    for k in range(len(o)):
        # march finely along the segment (domain-clamped as the DDA does)
        a = np.clip(o[k] * g.shape, 1e-4, np.array(g.shape) - 1e-4)
        b = np.clip(t[k] * g.shape, 1e-4, np.array(g.shape) - 1e-4)
        hit = None
        for s in np.linspace(0, 1, 20001):
            c = np.floor(a + s * (b - a)).astype(int)
            if tr.solid_at(solid, c[0], c[1], c[2]):
                hit = s
                break
        if hit is None:
            assert np.abs(res[k] * g.shape - b).max() < 1e-9
        else:
            stop = res[k] * g.shape
            frac = np.linalg.norm(stop - a) / max(np.linalg.norm(b - a), 1e-30)
            assert frac <= hit + 1e-3 and frac >= hit - 2e-3
            cell = np.floor(stop).astype(int)
            assert not tr.solid_at(solid, *cell)


def test_sealed_slab_blocks_every_trace():
    g = mac.Grid((8, 8, 8))
    solid = np.zeros(g.array_shape, bool)
    solid[4] = True
    rng = np.random.default_rng(4)
    o = np.column_stack([rng.uniform(.05, .95, 300), rng.uniform(.05, .95, 300), rng.uniform(.05, .45, 300)])
    t = np.column_stack([o[:, 0], o[:, 1], rng.uniform(.65, .95, 300)])
    res = tr.clip_solid_trace(solid, g, tuple(o.T), tuple(t.T))
    assert (res[2] < .5).all()


def test_combustion_reference_vectors():
    def react(fuel, temp, dt, ign, k):
        f, _, _, flame = st.react(np.array(float(fuel)), np.array(float(temp)), np.array(0.), dt, ign, k, 0, 0)
        return float(flame) * dt
    assert react(1, 149, .1, 150, 4) == 0 and react(0, 200, .1, 150, 4) == 0 and react(1, 200, .1, 150, 0) == 0
    assert abs(react(1, 150, .1, 150, 4) - (1 - math.exp(-.4))) < 1e-15
    f = 1.
    for _ in range(10):
        f -= react(f, 200, .01, 150, 4)
    assert abs(f - (1 - react(1, 200, .1, 150, 4))) < 1e-12
    for k in (0, 4, 10000):
        assert 0 <= react(1, 200, .1, 150, k) <= 1


def test_timestep_reference_vectors():
    assert st.choose_timestep(.03, .75, 0) == .03
    assert st.choose_timestep(.03, .75, 100) == .0075
    dt = st.choose_timestep(.1, .75, 0, 10000)
    assert dt * dt * 10000 <= .75 < (2 * dt) ** 2 * 10000
    for rate in (0, 1, 100, 1000):
        for acc in (0, 1, 100, 10000):
            q = st.choose_timestep(.1, .75, rate, acc)
            c = st.choose_timestep(.1, .75, rate, acc, quantized=False)
            assert rate * q + acc * q * q <= .75 and rate * c + acc * c * c <= .75 and c >= q * .99
    for args in ((0, .75, 0), (.03, 0, 0), (.03, .75, -1), (.03, .75, float("nan"))):
        with pytest.raises(ValueError):
            st.choose_timestep(*args)
    with pytest.raises(RuntimeError):
        st.choose_timestep(.03, .75, 1e9)


def test_max_abs_handles_nonfinite():
    assert tr.max_abs([1.0, -3.0, 2.0]) == 3.0
    assert tr.max_abs([1.0, float("nan")]) == 3.4e38


def _rotation_faces(grid, omega=1.0, modulate=None):
    """Solid-body rotation about the vertical axis through the domain centre,
    sampled at face positions (scripts/detail_validate.py:60-69), optionally
    scaled by 1 + 0.5 sin(0.1 i) over the flat face index (:80)."""
    n = grid.shape
    out = []
    for a in range(3):
        sz, sy, sx = grid.face_array_shape(a)
        z, y, x = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
        off = tr.FACE_OFFSETS[a]
        px, py = (x + off[0]) / n[0], (y + off[1]) / n[1]
        v = (-omega * (py - .5), omega * (px - .5), 0.0 * px)[a]
        if modulate:
            v = v * (1 + .5 * np.sin(np.arange(v.size) * .1)).reshape(v.shape)
        out.append(v)
    return tuple(out)


def test_solid_rotation_curl_and_no_interior_force():
    """detail_validate.py:70-78: curl = (0, 0, 2) in the interior, and a
    constant curl leaves interior faces untouched."""
    n = 12
    g = mac.Grid((n, n, n))
    vel = _rotation_faces(g)
    out, cu = tr.confine(vel, g, .04, 4, 2)
    inner = cu[3:n - 3, 3:n - 3, 3:n - 3]
    assert np.abs(inner[..., 2] - 2).max() < 1e-12 and np.abs(inner[..., :2]).max() < 1e-12
    assert np.allclose(inner[..., 3], 2.0)
    for o, v in zip(out, vel):
        assert np.abs(o - v)[4:-4, 4:-4, 4:-4].max() < 1e-12


def test_affine_flow_curl_closed_form():
    """An affine velocity u = A p has the constant curl of A away from the
    clamped boundary cells (central differences are exact on linear fields)."""
    n = (10, 8, 6)
    g = mac.Grid(n, (1.3, .8, 1.7))
    A = np.array([[.3, -1.1, .4], [.7, .2, -.5], [1.3, .9, -.5]])
    vel = []
    for a in range(3):
        sz, sy, sx = g.face_array_shape(a)
        z, y, x = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
        off = tr.FACE_OFFSETS[a]
        p = [(x + off[0]) * g.cell_size[0], (y + off[1]) * g.cell_size[1], (z + off[2]) * g.cell_size[2]]
        vel.append(A[a, 0] * p[0] + A[a, 1] * p[1] + A[a, 2] * p[2])
    cu = tr.curl(vel, g)
    expect = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    inner = cu[1:-1, 1:-1, 1:-1, :3]
    assert np.abs(inner - expect).max() < 1e-12


def test_confinement_spins_up_a_vortex():
    """The force is strength*h*(N x w) with N pointing up the |w| gradient: on
    the flank of a counter-clockwise vortex whose |w| peaks at its core the
    force points along the flow (a swapped cross product would brake it)."""
    n = 24
    g = mac.Grid((n, n, n))
    vel = []
    for a in range(3):
        sz, sy, sx = g.face_array_shape(a)
        z, y, x = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
        off = tr.FACE_OFFSETS[a]
        px, py = (x + off[0]) / n - .5, (y + off[1]) / n - .5
        s = np.exp(-(px * px + py * py) / .02)
        vel.append((-py * s, px * s, 0.0 * px)[a])
    out, cu = tr.confine(vel, g, .05, 2.0, 10.0)
    k, j = n // 2, n // 2
    i = n // 2 + 1  # on the +x flank of the core: flow +y, w_z > 0 and falling outwards
    assert vel[1][k, j, i] > 0 and cu[k, j, i, 2] > 0
    assert out[1][k, j, i] - vel[1][k, j, i] > 1e-4


def test_confinement_force_is_bounded():
    """detail_validate.py:79-84: a nonuniform rotation gets a finite force,
    each face changes by at most dt * limit."""
    n = 12
    g = mac.Grid((n, n, n))
    vel = _rotation_faces(g, modulate=True)
    out, _ = tr.confine(vel, g, .04, 20, .5)
    delta = max(np.abs(o - v).max() for o, v in zip(out, vel))
    assert 1e-4 < delta <= .04 * .5 + 1e-12
