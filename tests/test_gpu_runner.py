"""The caller side of the step on the GPU: the CFL timestep from the device
maxima (backend/timestep.py:35-53), the completion fence
(backend/completion.py:13-17) and the live loop that chains them
(blender/runtime.py:221-303 with physics/playback.py:27-40)."""
import numpy as np
import pytest

from oracle import mac, step as ostep, transport as otr
from tests.helpers import cells_np, faces_np

pytestmark = pytest.mark.gpu


def _sim(device, n=24, **kw):
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.physics.pressure import PressureSettings
    base = dict(source_center=(.5, .5, .2), source_radius=.15, source_rate=1.0, heat_source_rate=2.0,
                thermal_lift=1.0, density_weight=.5, dissipation=0.0, cooling=0.0, initial_temperature=0.0,
                source_profile="SOLID")
    base.update(kw)
    return DenseProjectedSmoke(GridSpec((n, n, n), (1.0, .8, 1.2)), PressureSettings(**base), device)


def test_select_closed_form_on_uniform_fields(device):
    """Uniform u = (0.5, -1.5, 2.0) m/s, density 0.4, temperature 3: the bound
    is rate = sum |u_a|/h_a and acc = (lift*T + weight*rho)/h_z, and dt is the
    oracle's choice under each policy (legacy Jacobi -> quantized, impulse
    multigrid -> continuous)."""
    from fluxfx_amd.backend.timestep import AdaptiveTimestep
    for solver, policy in (("JACOBI", "QUANTIZED"), ("MULTIGRID", "CONTINUOUS")):
        sim = _sim(device, pressure_solver=solver)
        n = sim.grid.shape[0]
        speeds = (.5, -1.5, 2.0)
        sim.upload_velocity([np.full(int(np.prod(sh)), v) for sh, v in zip(sim.grid.face_shapes, speeds)])
        sim.upload(np.full(n ** 3, .4))
        sim.upload_temperature(np.full(n ** 3, 3.0))
        ts = AdaptiveTimestep(sim.grid, sim.device)
        rep = ts.select(sim, .05, .75)
        h = sim.grid.cell_size
        rate = sum(abs(v) / hh for v, hh in zip(speeds, h))
        acc = (1.0 * 3.0 + .5 * .4) / h[2]
        assert rep["policy"] == policy
        assert rep["rate_bound"] == pytest.approx(rate, rel=1e-6)
        assert rep["acceleration_rate_bound"] == pytest.approx(acc, rel=1e-6)
        assert rep["dt"] == pytest.approx(ostep.choose_timestep(.05, .75, rate, acc, quantized=policy == "QUANTIZED"),
                                          rel=1e-6)
        assert rep["estimated_courant"] <= .75 * (1 + 1e-12)
        # a coupled jet faster than the field raises that axis's speed
        from fluxfx_amd.physics.emission import Emission
        jet = Emission((.5, .5, .2), (0., 0., 6.), 5.0)
        assert ts.select(sim, .05, .75, emission=jet)["max_abs_velocity"][2] == pytest.approx(6.0)
        sim.close()


def test_select_matches_host_maxima_of_a_random_state(device):
    """The fused max-abs reduction equals the FP64 max |.| of the downloaded
    fields exactly (fp32 values), and the dt is the oracle's for them."""
    from fluxfx_amd.backend.timestep import AdaptiveTimestep
    sim = _sim(device, pressure_solver="MULTIGRID", vorticity_strength=2.0, vorticity_limit=.3)
    og = mac.Grid(sim.grid.shape, sim.grid.extent)
    rng = np.random.default_rng(4)
    sim.upload_velocity([rng.uniform(-2, 2, int(np.prod(sh))) for sh in sim.grid.face_shapes])
    sim.upload(rng.uniform(0, 1, og.array_shape).ravel())
    sim.upload_temperature(rng.uniform(-3, 2, og.array_shape).ravel())
    ts = AdaptiveTimestep(sim.grid, sim.device)
    maxima = ts.bounds(sim)
    host = [otr.max_abs(f) for f in faces_np(og, sim._velocity)] + \
        [otr.max_abs(cells_np(og, sim.density)), otr.max_abs(cells_np(og, sim.temperature))]
    assert maxima == host
    rep = ts.select(sim, .1, .5)
    h = sim.grid.cell_size
    rate = sum(v / hh for v, hh in zip(host[:3], h))
    acc = (host[4] + .5 * host[3]) / h[2] + .3 * sum(1 / hh for hh in h)
    assert rep["dt"] == pytest.approx(ostep.choose_timestep(.1, .5, rate, acc, quantized=False), rel=1e-12)


def test_completion_fence_flags_nonfinite_state(device):
    from fluxfx_amd.backend.completion import StepCompletion
    sim = _sim(device, pressure_solver="JACOBI")
    sim.step(1 / 60)
    fence = StepCompletion(sim.device)
    fence.wait(sim)
    sim.density.tensor[0] = float("nan")
    with pytest.raises(RuntimeError):
        fence.wait(sim)


def test_live_loop_pays_the_owed_time_in_adaptive_substeps(device):
    """Three ticks of a fake wall clock: each pays the time owed in adaptive
    substeps (every one within the CFL bound of the state it started from),
    the simulated time equals the paid debt, and the emitter sweeps its path
    once per tick without gaps."""
    from fluxfx_amd.backend.runner import LiveRunner
    from fluxfx_amd.physics.emission import Emission
    sim = _sim(device, pressure_solver="MULTIGRID", source_rate=5.0, heat_source_rate=40.0)
    clock = {"now": 0.0}
    runner = LiveRunner(sim, time_step=1 / 60, cfl=.5, clock=lambda: clock["now"])
    spans = []
    orig = sim.step

    def recording_step(dt, emission=None, sources=None):
        spans.append((emission.start, sim.settings.source_center, dt))
        return orig(dt, emission, sources)
    sim.step = recording_step
    total = 0.0
    for k in range(3):
        clock["now"] += .05
        start = (.5, .5, .2 + .01 * k)
        sim.update_settings(__import__("dataclasses").replace(sim.settings, source_center=(.5, .5, .21 + .01 * k)))
        runner.tick(Emission(start, (0, 0, 1.0), 2.0), budget=10.0)
        rep = runner.playback.report
        assert rep["substeps"] >= 3 and not rep["limited"]
        assert runner.timestep_report["estimated_courant"] <= .5 * (1 + 1e-9)
        total += .05
        assert sim.time == pytest.approx(total) and runner.playback.simulated == pytest.approx(total)
        tick_spans = spans[-rep["substeps"]:]
        assert tick_spans[0][0] == pytest.approx(start)
        assert tick_spans[-1][1] == pytest.approx((.5, .5, .21 + .01 * k))
        for a, b in zip(tick_spans, tick_spans[1:]):
            assert a[1] == pytest.approx(b[0])
    assert sim.steps == len(spans)
    assert cells_np(mac.Grid(sim.grid.shape, sim.grid.extent), sim.density).max() > 0
