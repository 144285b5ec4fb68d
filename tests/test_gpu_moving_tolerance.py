"""Tolerance multigrid with moving colliders (no reference counterpart: the
reference projects moving colliders with masked Jacobi only).

* the level operators rebuilt on the device from the moving mask
  (fx_mg_rebuild_masked) equal the host construction of the same mask bit for
  bit: packed words, float4 conductances, activity bytes, level 1's mask;
* the coarsest level's FP64 conjugate gradients (fx_mg_coarse_cg) give the
  pseudo-inverse's solution;
* whole steps with a sphere moving through the plume, FP32 V-cycles stage by
  stage against the FP64 oracle, and the 16-bit and learned solvers of the
  fire workload against the oracle's converged projection.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import mac, step as ostep, transport as otr
from tests import stages
from tests.helpers import cells_np, faces_np
from tests.test_gpu_step import MIXED_TOL

pytestmark = pytest.mark.gpu


def _rows_fn(old, new, dt):
    from fluxfx_amd.physics.collision import Collider
    from fluxfx_amd.physics.collider_motion import velocity_rows
    return velocity_rows(Collider(*old), Collider(*new), dt)


def _turned_box(center, half, angle):
    """Box collider turned by `angle` about z: inverse rows q = R^T (p - c) / half."""
    from fluxfx_amd.physics.collision import Collider
    c, sn = np.cos(angle), np.sin(angle)
    axes = [(c, sn, 0.0), (-sn, c, 0.0), (0.0, 0.0, 1.0)]
    rows = []
    for a, h in zip(axes, half):
        r = [v / h for v in a]
        rows.append((r[0], r[1], r[2], -(r[0] * center[0] + r[1] * center[1] + r[2] * center[2])))
    return Collider("BOX", tuple(rows))


def _host_levels(proj):
    """(op, words or float32 edges, activity bytes) per level >= 1 from the
    host construction of the projector's current mask, and level 1's mask."""
    from fluxfx_amd import _lib
    from fluxfx_amd.backend.engine import activity_bytes, conductance_hierarchy, pack_conductances
    shapes = proj.shapes
    nx, ny, nz = proj.grid.shape
    solid = proj.solids.solid.cpu().numpy().astype(bool).reshape(nz, ny, nx)
    edges = conductance_hierarchy(shapes, proj.grid.extent, solid)
    out = []
    for level, (e, shape) in enumerate(zip(edges, shapes[1:]), start=1):
        words = pack_conductances(e, shape, proj.grid.extent, level)
        act = activity_bytes(e[..., 3] > 0).ravel()
        if words is not None:
            out.append((_lib.FX_OP_PACKED, words.ravel().view(np.uint32), act))
        else:
            out.append((_lib.FX_OP_EDGES, np.ascontiguousarray(e, np.float32).ravel(), act))
    coarse_solid = solid.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).all(axis=(1, 3, 5)).ravel()
    return out, coarse_solid


def _check_operators(proj):
    host, coarse_solid = _host_levels(proj)
    cfg = proj.engine.cfg
    for level, (op, data, act) in enumerate(host, start=1):
        assert cfg.level[level].op == op, level
        dev = proj.ops[level][1].cpu().numpy()
        if op == 3:  # FX_OP_PACKED
            assert np.array_equal(dev.view(np.uint32), data), level
        else:
            assert np.array_equal(dev.view(np.uint32), data.view(np.uint32)), level
        assert np.array_equal(proj._active_arrays[level].cpu().numpy(), act), level
    assert np.array_equal(proj._coarse_solid.cpu().numpy().astype(bool), coarse_solid)


@pytest.mark.parametrize("shape,extent", [((64, 64, 64), (1.0, 1.0, 1.0)), ((48, 32, 64), (1.5, 1.0, 2.0))])
def test_rebuilt_operators_equal_host_construction(device, shape, extent):
    from fluxfx_amd.backend.moving_collision import MovingToleranceProjector
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    from fluxfx_amd.physics.collision import primitive
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.physics.pressure import PressureSettings
    print(f"rebuilt operators: {shape} {extent}", flush=True)
    e = np.array(extent)
    first = primitive("SPHERE", tuple(e * (.4, .45, .5)), tuple(e * .17))
    box = primitive("BOX", tuple(e * (.7, .6, .3)), tuple(e * (.1, .2, .12)))
    s = PressureSettings(moving_colliders=True, pressure_tolerance=1e-5, mg_coarsest=2)
    sim = DenseProjectedSmoke(GridSpec(shape, extent), s, device, colliders=(first, box))
    proj = sim.projector
    assert isinstance(proj, MovingToleranceProjector)
    # the hierarchy reaches a float4 level where 4^level > 1023 on the cube
    assert any(o[0] == 2 for o in proj.ops) or len(proj.shapes) <= 5
    _check_operators(proj)
    # the sphere moves and the box turns a little: rebuilt before the next solve
    moved = primitive("SPHERE", tuple(e * (.47, .43, .52)), tuple(e * .17))
    turned = _turned_box(tuple(e * (.7, .6, .3)), tuple(e * (.1, .2, .12)), 0.3)
    sim.move_colliders((moved, turned), .01)
    sim.step(.01)
    _check_operators(proj)
    sim.close()


def test_coarse_conjugate_gradients_give_the_pseudo_inverse_solution(device):
    from fluxfx_amd import _lib
    from fluxfx_amd.backend.engine import conductance_hierarchy, edges_matrix
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    from fluxfx_amd.physics.collision import primitive
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.physics.pressure import PressureSettings
    torch = device.torch
    s = PressureSettings(moving_colliders=True, pressure_tolerance=1e-5, mg_coarsest=8)
    c = primitive("SPHERE", (.45, .5, .5), (.3,) * 3)
    sim = DenseProjectedSmoke(GridSpec((64,) * 3), s, device, colliders=(c,))
    proj = sim.projector
    eng = proj.engine
    inc = proj._increment_fields()
    eng.bind_level0(inc["d"].tensor, inc["d2"].tensor, inc["r0"].tensor)
    solid = proj.solids.solid.cpu().numpy().astype(bool).reshape(64, 64, 64)
    last = conductance_hierarchy(proj.shapes, proj.grid.extent, solid)[-1]
    A = edges_matrix(last)
    m = A.shape[0]
    assert m == 512 and proj.ops[-1][0] == _lib.FX_OP_PACKED
    rng = np.random.default_rng(7)
    x = rng.normal(size=m) * (last[..., 3].ravel() > 0)
    rhs = (A @ x).astype(np.float32)
    want = np.linalg.pinv(A, rcond=1e-12, hermitian=True) @ rhs.astype(np.float64)
    eng.tensor(len(proj.shapes) - 1, "rhs").copy_(torch.from_numpy(rhs))
    info = torch.zeros(2, dtype=torch.float64, device=device.device)
    _lib.call("fx_mg_coarse_cg", C.byref(eng.cfg), info.data_ptr(), device.stream)
    got = eng.tensor(len(proj.shapes) - 1, "p").cpu().numpy().astype(np.float64)
    its, res = info.cpu().numpy()
    print(f"coarse CG: {int(its)} iterations, relative residual {res:.2e}", flush=True)
    assert 0 < its < m and res <= 1e-7
    assert np.abs(got - want).max() <= 1e-4 * np.abs(want).max()
    sim.close()


def _fire_scene(n, s_kw):
    from fluxfx_amd.physics.collision import primitive
    from fluxfx_amd.physics.pressure import PressureSettings
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
                         velocity_advection="MACCORMACK", scalar_advection="MACCORMACK", pressure_tolerance=1e-5,
                         mg_pre=2, mg_post=2, mg_coarsest=8, **s_kw)
    srcs = [otr.Src(s.source_center, s.source_radius, 0.0, s.heat_source_rate, solid_profile=True,
                    fuel_rate=s.fuel_source_rate)]
    return s, srcs, sphere, dt


MODES = {
    "fp32": {},
    "bf16_pcg": {"pressure_precision": "BF16", "pressure_krylov": "PCG"},
    "fp16_refinement": {"pressure_precision": "FP16", "pressure_krylov": "NONE"},
    "learned": {"pressure_precision": "BF16", "pressure_krylov": "LEARNED", "learned_iterations": 1,
                "learned_sweeps": 0, "learned_level": 1},
}


@pytest.mark.parametrize("mode", list(MODES))
def test_fire_with_moving_sphere_and_tolerance_multigrid(device, mode):
    """The fire workload at 32^3 with a sphere moving 0.5 cells per step,
    projected to 1e-5 by each solver: the move checked exactly, then every
    step stage by stage (FP32: the oracle runs the same V-cycles; 16-bit and
    learned: against the oracle's converged projection)."""
    import bench
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    from fluxfx_amd.physics.config import GridSpec
    n, steps = 32, 5
    print(f"moving sphere + tolerance MG at {n}^3: {mode}", flush=True)
    kw = dict(MODES[mode])
    if mode == "learned":
        kw["learned_weights"] = str(bench.WEIGHTS)
    s, srcs, sphere, dt = _fire_scene(n, kw)
    og = mac.Grid((n,) * 3)
    sim = DenseProjectedSmoke(GridSpec((n,) * 3), s, device, colliders=(sphere(0),))
    assert sim.projector.solver_name == "MOVING_TOLERANCE_MULTIGRID"
    sim.upload_velocity([np.zeros(int(np.prod(sh))) for sh in sim.grid.face_shapes])
    sim.upload(np.zeros(n ** 3))
    sim.upload_temperature(np.zeros(n ** 3))
    sim.upload_fuel(np.zeros(n ** 3))
    cfg = ostep.PressureConfig("TOLERANCE", tol=1e-5, pre=2, post=2, coarsest=8)
    pair = lambda c: (c.shape, c.inverse_rows)
    for k in range(1, steps + 1):
        new = sphere(k)
        st = ostep.State(og, faces_np(og, sim._velocity), cells_np(og, sim.density), cells_np(og, sim.temperature),
                         fuel=cells_np(og, sim.combustion.fuel),
                         solid=sim.solids.solid.cpu().numpy().reshape(og.array_shape) > 0)
        prev = pair(sim.solids.colliders[0])
        sim.move_colliders((new,), dt)
        ostep.move_colliders(st, [prev], [pair(new)], dt, _rows_fn)
        mask = sim.solids.solid.cpu().numpy().reshape(og.array_shape) > 0
        assert np.array_equal(mask, st.solid)
        walls = faces_np(og, sim.solids.wall_fields)
        stages.step_stagewise(sim, og, s, dt, srcs, cfg, solid=mask, boundary=walls,
                              converged=mode != "fp32", pressure_tol=MIXED_TOL if mode != "fp32" else 1e-5)
        m = sim.measure_projection()
        assert m["converged"] and m["relative_residual"] <= 1e-5, (k, m)
        assert m["solver"] == "MOVING_TOLERANCE_MULTIGRID" and m["boundary"] == "moving_voxel_free_slip"
        after = cells_np(og, sim.projector.after)
        before = cells_np(og, sim.projector.before)
        assert np.sqrt((after ** 2).sum()) <= 2e-5 * np.sqrt((before ** 2).sum()), k
    assert cells_np(og, sim.combustion.flame).max() > 0
    sim.close()
