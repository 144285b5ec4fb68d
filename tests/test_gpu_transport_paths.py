"""The interior fast path of the transport kernels against the general path
(``fx_transport_fast_paths``): whole smoke steps of the benchmark scene at
96^3 (box colliders, MacCormack and semi-Lagrangian transport) and of a
solid-free fire give the same bits with the fast path on and off, for every
field.  At 96^3 most waves are interior (the fast path runs) and the walls,
the boxes and the emitter give waves of the general path in the same
launches."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(sim):
    out = {"density": sim.density, "temperature": sim.temperature, "pressure": sim.projector.pressure}
    for a, f in enumerate(sim._velocity):
        out["uvw"[a]] = f
    if sim.combustion is not None:
        out["fuel"] = sim.combustion.fuel
    import torch
    return {k: v.tensor.detach().view(torch.int32).cpu().numpy().copy() for k, v in out.items()}


def _run(device, make, steps, fast):
    from fluxfx_amd import _lib
    lib = _lib.load()
    prev = lib.fx_transport_fast_paths(1 if fast else 0)
    try:
        sim, dt = make()
        for _ in range(steps):
            sim.step(dt)
        device.torch.cuda.synchronize()
        out = _fields(sim)
        sim.close()
        return out
    finally:
        lib.fx_transport_fast_paths(prev)


def _compare(a, b):
    bad = {k: int(np.count_nonzero(a[k] != b[k])) for k in a if np.any(a[k] != b[k])}
    assert not bad, f"fields differ between the fast and general transport paths (cells): {bad}"


@pytest.mark.parametrize("advection", ["MACCORMACK", "SEMI_LAGRANGIAN"])
def test_fast_path_matches_general_path_on_the_bench_scene(device, advection):
    import bench
    print(f"transport paths: bench scene 96^3 advection={advection}", flush=True)

    def make():
        sim, dt = bench.make_sim(device, 96, "mixed")
        if advection != "MACCORMACK":
            sim.update_settings(dataclasses.replace(sim.settings, velocity_advection=advection,
                                                    scalar_advection=advection))
        return sim, dt

    fast = _run(device, make, 6, True)
    general = _run(device, make, 6, False)
    _compare(fast, general)
    assert np.count_nonzero(fast["density"]) > 0


def test_fast_path_matches_general_path_without_solids(device):
    """The fire workload (fuel, combustion, vorticity confinement) with its
    sphere removed: the solid-free kernel variants."""
    import bench
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    from fluxfx_amd.physics.config import GridSpec
    print("transport paths: fire 64^3 without solids", flush=True)

    def make():
        settings, dt = bench.settings_for(64, "plain", "fire")
        grid = GridSpec((64, 64, 64), (1.0, 1.0, 1.0), max_dim=1024)
        sim = DenseProjectedSmoke(grid, settings, device)
        center, radius, _ = bench.scene(64)
        sim.upload_velocity([v.ravel() for v in bench.initial_velocity(64, center, radius)])
        return sim, dt

    fast = _run(device, make, 6, True)
    general = _run(device, make, 6, False)
    _compare(fast, general)
