"""Training data and training loop of the learned pressure preconditioner.

Pairs come from the solver itself: seeded random smoke scenes (1-4 spherical
emitters, 0-8 sphere/box colliders, temperature rate U[0.5, 1.5]) are stepped
with plain FP32 tolerance multigrid, and after every step the pre-projection
divergence b (fluid-mean removed), the converged impulse q (L q = b) and the
solid mask are kept on the device.  The network learns
    (b / rms_fluid(b), solid)  ->  q / (rms_fluid(b) * h^2)
with a fluid-masked mean-squared error, one pair per Adam step.
"""
import dataclasses
import time

import numpy as np

from . import _lib


def random_scene(n, seed):
    """Emitters, collider mask and settings of one harvesting scene at n^3."""
    from .physics.emission import Emission, Source
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    sources = []
    for _ in range(int(rng.integers(1, 5))):
        c = (rng.uniform(.2, .8), rng.uniform(.2, .8), rng.uniform(.08, .3))
        radius = float(rng.uniform(4, 8)) * max(n / 128, 0.5) * h
        dt = 0.5 * h
        # density +1 and temperature +U[0.5, 1.5] per step inside the emitter
        sources.append(Source(c, radius, 1.0 / dt, float(rng.uniform(.5, 1.5)) / dt, Emission(c), profile="SOLID"))
    z, y, x = np.meshgrid(*(np.arange(n) + .5,) * 3, indexing="ij")
    solid = np.zeros((n, n, n), bool)
    for _ in range(int(rng.integers(0, 9))):
        c = rng.uniform(.15, .85, 3) * n
        if rng.uniform() < .5:
            r = rng.uniform(.04, .12) * n
            m = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r
        else:
            e = rng.uniform(.05, .2, 3) * n
            m = (np.abs(x - c[0]) <= e[0] / 2) & (np.abs(y - c[1]) <= e[1] / 2) & (np.abs(z - c[2]) <= e[2] / 2)
        solid |= m
    # keep every emitter free
    for s in sources:
        cx, cy, cz = (v * n for v in s.center)
        rr = s.radius * n + 2
        solid &= (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 > rr * rr
    return sources, (solid if solid.any() else None)


def harvest(device, n, scenes, steps_per_scene, seed0=10000, warmup=2, log=None):
    """Run `scenes` seeded scenes for warmup + steps_per_scene steps each and keep
    (b, target, solid, inv_count) per harvested step (device tensors)."""
    from .backend.projected import DenseProjectedSmoke
    from .physics.config import GridSpec
    from .physics.pressure import PressureSettings
    torch = device.torch
    grid = GridSpec((n, n, n), (1.0, 1.0, 1.0), max_dim=1024)
    dt = 0.5 / n
    h2 = (1.0 / n) ** 2
    samples = []
    t0 = time.time()
    for k in range(scenes):
        sources, solid = random_scene(n, seed0 + k)
        settings = PressureSettings(thermal_lift=1.0, density_weight=0.0, dissipation=0.0, cooling=0.0,
                                    velocity_advection="MACCORMACK", scalar_advection="MACCORMACK",
                                    pressure_tolerance=1e-5, pressure_max_cycles=60, mg_pre=2, mg_post=2,
                                    mg_coarsest=8)
        sim = DenseProjectedSmoke(grid, settings, device,
                                  solid_mask=None if solid is None else solid.ravel().astype(np.uint8))
        ts = None if solid is None else torch.from_numpy(solid.ravel().astype(np.uint8)).to(device.device)
        fluid = n ** 3 - (0 if solid is None else int(solid.sum()))
        for step in range(warmup + steps_per_scene):
            sim.step(dt, sources=sources)
            if step < warmup:
                continue
            pp = sim.projector
            b = pp.before.tensor.clone()
            q = pp.pressure.tensor.clone()
            mask = torch.ones_like(b, dtype=torch.bool) if ts is None else ts == 0
            rms = torch.sqrt((b[mask].double() ** 2).mean()).item()
            if not rms > 0:
                continue
            target = (q / (rms * h2)).float()
            samples.append((b, target, ts, 1.0 / fluid))
        sim.close()
        if log and (k + 1) % max(1, scenes // 8) == 0:
            log(f"harvest: {k + 1}/{scenes} scenes, {len(samples)} pairs, {time.time() - t0:.1f}s")
    return samples


def _restrict_mean(t, n):
    """8-child mean of an n^3 x-fast field (the multigrid restriction)."""
    return t.view(n // 2, 2, n // 2, 2, n // 2, 2).mean(dim=(1, 3, 5)).reshape(-1)


def harvest_increments(device, n, scenes, steps_per_scene, seed0=10000, warmup=4, log=None):
    """Pairs for the network run one multigrid level down inside warm-started
    solves: the same scenes stepped at n^3, and after every step the solve's
    increment problem L d = r0 (r0 = b - L x0 from the previous impulse) is
    restricted by the 8-child mean to (n/2)^3.  Returns (r0_c, target, solid_c,
    inv_count) with target = d_c / (rms_fluid(r0_c) * h_c^2), h_c = 2/n; a
    coarse cell is solid when all eight children are."""
    from .backend.projected import DenseProjectedSmoke
    from .physics.config import GridSpec
    from .physics.pressure import PressureSettings
    torch = device.torch
    grid = GridSpec((n, n, n), (1.0, 1.0, 1.0), max_dim=1024)
    dt = 0.5 / n
    nc = n // 2
    hc2 = (1.0 / nc) ** 2
    samples = []
    t0 = time.time()
    skipped = 0
    for k in range(scenes):
        sources, solid = random_scene(n, seed0 + k)
        # FP32 V-cycle preconditioned CG: the robust solver for cold starts at large n
        settings = PressureSettings(thermal_lift=1.0, density_weight=0.0, dissipation=0.0, cooling=0.0,
                                    velocity_advection="MACCORMACK", scalar_advection="MACCORMACK",
                                    pressure_tolerance=1e-5, pressure_max_cycles=60, mg_pre=2, mg_post=2,
                                    mg_coarsest=8, pressure_krylov="PCG")
        sim = DenseProjectedSmoke(grid, settings, device,
                                  solid_mask=None if solid is None else solid.ravel().astype(np.uint8))
        if solid is None:
            sc = None
            fluid = nc ** 3
        else:
            coarse = solid.reshape(nc, 2, nc, 2, nc, 2).all(axis=(1, 3, 5))
            sc = torch.from_numpy(coarse.ravel().astype(np.uint8)).to(device.device)
            fluid = int((~coarse).sum())
        # the cold start stalls on FP32 rounding near 2e-5 at 256^3: warm up at 1e-4,
        # then harvest the warm-started solves at the real tolerance
        sim.projector.settings = dataclasses.replace(settings, pressure_tolerance=1e-4)
        for step in range(warmup + steps_per_scene):
            if step == warmup:
                sim.projector.settings = settings
            try:
                sim.step(dt, sources=sources)
            except RuntimeError as exc:  # a solve that misses the tolerance ends the scene
                skipped += 1
                if log:
                    log(f"harvest (increments): scene {seed0 + k} stopped at step {step}: {exc}")
                break
            if step < warmup:
                continue
            inc = sim.projector._inc
            r0 = _restrict_mean(inc["r0"].tensor, n)
            d = _restrict_mean(inc["d"].tensor, n)
            mask = torch.ones_like(r0, dtype=torch.bool) if sc is None else sc == 0
            rms = torch.sqrt((r0[mask].double() ** 2).mean()).item()
            if not rms > 0:
                continue
            samples.append((r0.contiguous(), (d / (rms * hc2)).float().contiguous(), sc, 1.0 / fluid))
        sim.close()
        if log and (k + 1) % max(1, scenes // 8) == 0:
            log(f"harvest (increments): {k + 1}/{scenes} scenes, {len(samples)} pairs, {time.time() - t0:.1f}s")
    return samples


def train(net, samples, shape, steps, lr=1e-3, seed=0, log=None, log_every=200):
    """Adam on single pairs drawn in a seeded shuffled order; returns the loss history."""
    torch = net.device.torch
    ws = torch.zeros(_lib.FX_WS_DOUBLES, dtype=torch.float64, device=net.device.device)
    rng = np.random.default_rng(seed)
    order = []
    hist = []
    t0 = time.time()
    for it in range(steps):
        if not order:
            order = list(rng.permutation(len(samples)))
        b, target, solid, inv_count = samples[order.pop()]
        net.train_step(shape, b, target, solid, ws, inv_count, lr)
        if (it + 1) % log_every == 0 or it == steps - 1:
            loss = (ws[10] * inv_count).item()
            hist.append((it + 1, loss))
            if log:
                log(f"train: step {it + 1}/{steps} loss {loss:.4e} ({time.time() - t0:.1f}s)")
    return hist


def evaluate(net, samples, shape):
    """Mean masked MSE of the network over samples (no update)."""
    torch = net.device.torch
    ws = torch.zeros(_lib.FX_WS_DOUBLES, dtype=torch.float64, device=net.device.device)
    total = 0.0
    for b, target, solid, inv_count in samples:
        net.forward(shape, b, solid, ws, post_scale=1.0, output_rms=False)
        net.loss_grad(shape, target, solid, ws, inv_count)
        total += (ws[10] * inv_count).item()
    return total / max(len(samples), 1)
