"""Exact active bricks of the scalar passes (no reference counterpart).

A step with active_bricks=True runs the scalar predictor, transport and
combustion only on the 8^3 bricks a plan selects from the fields' exact-zero
masks.  Its fields must equal those of the same step over the whole grid, value
for value, after every step: plumes with box solids and tolerance multigrid,
the fire workload with a moving sphere, semi-Lagrangian scalars from several
emitters on a grid that is not a multiple of 8, a velocity too fast for a
plan (every brick computed), and a field replaced between steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(sim):
    out = {"density": sim.density.tensor, "temperature": sim.temperature.tensor,
           **{f"velocity{a}": f.tensor for a, f in enumerate(sim._velocity)},
           "pressure": sim.projector.pressure.tensor}
    if sim.combustion:
        out.update(fuel=sim.combustion.fuel.tensor, flame=sim.combustion.flame.tensor)
    return out


def _assert_same(a, b, label):
    fa, fb = _fields(a), _fields(b)
    for k in fa:
        assert bool((fa[k] == fb[k]).all()), f"{label}: {k} differs in " \
            f"{int((fa[k] != fb[k]).sum())} cells"


def _pair(device, grid, settings, **kw):
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    return (DenseProjectedSmoke(grid, settings, device, **kw),
            DenseProjectedSmoke(grid, settings, device, active_bricks=True, **kw))


def _run(pair, steps, dt, before=None, sources=None, label=""):
    fractions = []
    for k in range(steps):
        for sim in pair:
            if before:
                before(sim, k)
            sim.step(dt, sources=sources)
        _assert_same(*pair, f"{label} step {k + 1}")
        fractions.append(pair[1].active.fractions())
    return fractions


def _smoke_settings(n, **kw):
    from fluxfx_amd.physics.pressure import PressureSettings
    dt = 0.5 / n
    base = dict(source_center=(.5, .5, .16), source_radius=.09, source_rate=1 / dt, heat_source_rate=1 / dt,
                dissipation=0.0, cooling=0.05, thermal_lift=1.0, density_weight=0.0, source_profile="SOLID",
                initial_temperature=0.0, velocity_advection="MACCORMACK", scalar_advection="MACCORMACK",
                pressure_tolerance=1e-5, mg_pre=2, mg_post=2, mg_coarsest=8)
    base.update(kw)
    return PressureSettings(**base), dt


@pytest.mark.parametrize("precision", ["FP32", "BF16"])
def test_plume_with_boxes_equals_whole_grid(device, precision):
    import bench
    from fluxfx_amd.physics.config import GridSpec
    n = 64
    print(f"active bricks: plume with boxes at {n}^3, {precision}", flush=True)
    s, dt = _smoke_settings(n, pressure_precision=precision, pressure_krylov="PCG" if precision != "FP32" else "NONE")
    _, _, boxes = bench.scene(n)
    mask = bench.box_mask(n, boxes).ravel()
    pair = _pair(device, GridSpec((n,) * 3), s, solid_mask=mask)
    fr = _run(pair, 10, dt, label="plume")
    # the plume covers a small part of the grid: most bricks are skipped
    assert fr[-1]["transport"] < 0.5 and fr[-1]["predict"] < 0.7, fr[-1]


def test_fire_with_moving_sphere_equals_whole_grid(device):
    from fluxfx_amd.physics.config import GridSpec
    from tests.test_gpu_moving_tolerance import _fire_scene
    n = 48
    print(f"active bricks: fire with a moving sphere at {n}^3", flush=True)
    s, _, sphere, dt = _fire_scene(n, {"pressure_precision": "BF16", "pressure_krylov": "PCG"})
    pair = _pair(device, GridSpec((n,) * 3), s, colliders=(sphere(0),))

    def move(sim, k):
        sim.move_colliders((sphere(k + 1),), dt)

    fr = _run(pair, 10, dt, before=move, label="fire")
    assert fr[-1]["combustion"] < 0.5 and fr[-1]["transport"] < 0.7, fr[-1]


def test_semi_lagrangian_sources_on_a_ragged_grid(device):
    """40 x 36 x 44 cells (partial bricks at the high ends), semi-Lagrangian
    scalars, two moving emitters of the multi-source table, Jacobi projection."""
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.physics.emission import Emission, Source
    from fluxfx_amd.physics.pressure import PressureSettings
    shape = (40, 36, 44)
    print(f"active bricks: semi-Lagrangian emitters on {shape}", flush=True)
    dt = 0.5 / 44
    s = PressureSettings(dissipation=0.1, cooling=0.0, thermal_lift=1.0, density_weight=0.2, initial_temperature=0.0,
                         scalar_advection="SEMI_LAGRANGIAN", velocity_advection="MACCORMACK",
                         pressure_solver="JACOBI", pressure_iterations=40)
    pair = _pair(device, GridSpec(shape), s)
    srcs = [Source((.3, .4, .2), .08, 20.0, 10.0, Emission((.3, .4, .2), (0.0, 0.0, 0.5), 4.0), "RATE", "SOFT"),
            Source((.7, .6, .5), .06, 10.0, 5.0, Emission((.65, .6, .45)), "TARGET", "SOLID")]
    _run(pair, 8, dt, sources=srcs, label="ragged")


def test_fast_velocity_plans_every_brick_and_replaced_fields(device):
    """A velocity moving more than 2.5 cells per step: the plans compute every
    brick (still equal); then a density uploaded between steps is rescanned."""
    from fluxfx_amd.physics.config import GridSpec
    n = 32
    print(f"active bricks: fast velocity at {n}^3", flush=True)
    s, dt = _smoke_settings(n, pressure_tolerance=0.0, pressure_solver="JACOBI", pressure_iterations=30)
    pair = _pair(device, GridSpec((n,) * 3), s)
    rng = np.random.default_rng(11)
    vel = [rng.uniform(-20.0, 20.0, int(np.prod(sh))) for sh in pair[0].grid.face_shapes]
    for sim in pair:
        sim.upload_velocity(vel)
    fr = _run(pair, 2, dt, label="fast")
    assert fr[0]["transport"] == 1.0 and fr[0]["predict"] == 1.0, fr[0]
    blob = np.zeros((n, n, n))
    blob[20:26, 3:9, 10:14] = 1.0
    for sim in pair:
        sim.upload(blob.ravel())
        sim.upload_velocity([np.zeros(int(np.prod(sh))) for sh in sim.grid.face_shapes])
    _run(pair, 3, dt, label="replaced")
    # a write into the tensor in place, announced with invalidate_active_bricks()
    for sim in pair:
        sim.temperature.tensor.view(n, n, n)[2:5, 25:29, 4:8] = 0.75
    pair[1].invalidate_active_bricks()
    _run(pair, 3, dt, label="written in place")


@pytest.mark.parametrize("shape", [(48, 32, 40), (30, 20, 18)])
@pytest.mark.parametrize("walls", [False, True])
def test_divergence_bound_equals_divergence_and_max_abs(device, shape, walls):
    """fx_divergence_bound: the divergence of fx_divergence(_moving) bit for bit
    and max |U|, |V|, |W| over the whole arrays (vector path; a grid with
    nx % 4 != 0 takes the one-cell path)."""
    import ctypes as C
    from fluxfx_amd import _lib
    from fluxfx_amd.backend.device import grid_struct
    from fluxfx_amd.physics.config import GridSpec
    torch = device.torch
    print(f"divergence bound: {shape}, walls={walls}", flush=True)
    grid = GridSpec(shape, (1.0, 0.8, 1.2))
    g = grid_struct(grid)
    nx, ny, nz = shape
    rng = np.random.default_rng(3)
    solid = (rng.random((nz, ny, nx)) < 0.1).astype(np.uint8)
    solid_t = torch.from_numpy(solid.ravel()).to(device.device)
    flags = torch.zeros(nx * ny * nz, dtype=torch.uint8, device=device.device)
    _lib.call("fx_build_flags", g, solid_t.data_ptr(), flags.data_ptr(), device.stream)
    vel = [torch.from_numpy(rng.uniform(-3, 3, int(np.prod(s))).astype(np.float32)).to(device.device)
           for s in grid.face_shapes]
    vel[1][17] = -7.5  # a known maximum on V
    wall = [torch.from_numpy(rng.uniform(-1, 1, int(np.prod(s))).astype(np.float32)).to(device.device)
            for s in grid.face_shapes] if walls else [None, None, None]
    wp = [w.data_ptr() if w is not None else None for w in wall]
    n = nx * ny * nz
    div_ref = torch.empty(n, device=device.device)
    div = torch.empty(n, device=device.device)
    part = torch.zeros(3 * -(-n // 1024), device=device.device)
    vmax = torch.zeros(3, device=device.device)
    if walls:
        _lib.call("fx_divergence_moving", g, flags.data_ptr(), *(t.data_ptr() for t in vel), *wp,
                  div_ref.data_ptr(), device.stream)
    else:
        _lib.call("fx_divergence", g, flags.data_ptr(), *(t.data_ptr() for t in vel), div_ref.data_ptr(),
                  device.stream)
    _lib.call("fx_divergence_bound", g, flags.data_ptr(), *(t.data_ptr() for t in vel), *wp, div.data_ptr(),
              part.data_ptr(), vmax.data_ptr(), device.stream)
    torch.cuda.synchronize()
    assert torch.equal(div, div_ref)
    want = torch.stack([t.abs().max() for t in vel])
    assert torch.equal(vmax, want), (vmax.tolist(), want.tolist())
    assert float(vmax[1]) == 7.5
