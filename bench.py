"""Benchmark: 256^3 smoke step projected to 1e-5 relative residual on one MI355X.

Workload (the 256^3 mixed-precision projection scene): the 64^3 closed-box
plume scaled to the 256^3 domain (domain-relative x4: emitter radius 24 cells
centred at cells (128, 128, 40) with z vertical), density and temperature each
+1.0 per step inside the emitter (solid profile, rate 1/dt), buoyancy
f_z = 1.0*T, zero velocity except seeded U(-0.05, 0.05) inside the emitter,
MacCormack transport of density, temperature and U/V/W, plus 8 seeded
axis-aligned box colliders with edges of 8-32 cells.  Fixed dt = 0.5*h/(1 m/s)
(CFL 0.5).  Pressure: V(2,2) damped-Jacobi multigrid, 8-child mean restriction,
trilinear prolongation, FP64 direct solve on 8^3, warm-started impulse, run to
||b - A q|| <= 1e-5 ||b||.

Modes: plain (FP32 stationary V-cycles, the drop-in), mixed (BF16 V-cycles
preconditioning FP32 flexible CG) and learned (one FPCG step preconditioned by
a U-Net applied to the 8-child mean of the residual on the 128^3 level, its
output prolonged as a V-cycle correction, then BF16 V-cycle FPCG; the network was trained
on the warm-started solves' increment problems of seeded random 256^3 scenes,
restricted to 128^3),
whichever are built.  With the CPU leg the line also carries every field's
relative error against the FP64 oracle on the same scene at 64^3 after 7 steps
(``fp64_parity_64``).  ``value`` is the steps/s of the learned+mixed mode (of
the mixed mode when no trained weights are present): 1000 / the median of
the K steps' times, each step bracketed by HIP events on the stream its
kernels run on; every mode's rate (and its mean, minimum and wall-clock
figures) and its ratio to plain are in the JSON line.  ``roofline`` is the
level-0 smoother of the BF16 V-cycle (the kernel with the largest share of
the learned+mixed step), timed with HIP events over 60 back-to-back sweeps.

The scalar predictor, scalar transport and combustion run over the 8^3
bricks that can hold nonzero scalars (exact zero tracking: the values of the
whole-grid passes; ``modes.<mode>.scalar_bricks`` gives the shares of bricks
the last step's passes computed); ``--dense-scalars`` runs them over the
whole grid instead.

``--workload bricks`` times the sparse brick pool at 256^3 instead (support
scan, gather and scatter of the reference's 112^3-box occupancy case).

``--workload fire`` runs a 384^3 dense fire instead (fuel and heat emitter,
combustion, vorticity confinement, a sphere collider moving 0.5 cells per step;
no CPU leg).

Prints ONE JSON line.  Single GPU only.
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "256³ smoke steps/s at 1e-5 rel. residual: learned+mixed vs plain FP32 MG"
WEIGHTS = ROOT / "fluxfx_amd" / "weights" / "unet_128_inc.npz"
HBM_PEAK_GBS = 8000.0
LEARNED_LEVEL = 1
LEARNED_ITERATIONS = 1
# damped-Jacobi sweeps after the prolonged network correction: at 256^3 they
# buy no iterations (0/2/4/6 sweeps: 7.30/7.35/7.40/7.40 per step), so none run
LEARNED_SWEEPS = 0
# scalar passes over the bricks that can hold nonzero scalars (same values)
ACTIVE_BRICKS = True


def scene(n, seed=3):
    """Emitter, box colliders and initial velocity of the benchmark scene at n^3."""
    rng = np.random.default_rng(seed)
    s = n / 256.0
    center_cells = (128 * s, 128 * s, 40 * s)
    radius_cells = 24 * s
    boxes = []
    while len(boxes) < 8:
        edge = rng.integers(8, 33, 3) * s
        c = rng.uniform(edge / 2 + 4 * s, n - edge / 2 - 4 * s)
        # keep clear of the emitter sphere
        gap = np.maximum(np.abs(c - np.array(center_cells)) - edge / 2, 0)
        if np.sqrt((gap ** 2).sum()) < radius_cells + 6 * s:
            continue
        boxes.append((tuple(c), tuple(edge)))
    return center_cells, radius_cells, boxes


def box_mask(n, boxes):
    z, y, x = np.meshgrid(*(np.arange(n) + .5,) * 3, indexing="ij")
    m = np.zeros((n, n, n), bool)
    for c, e in boxes:
        m |= (np.abs(x - c[0]) <= e[0] / 2) & (np.abs(y - c[1]) <= e[1] / 2) & (np.abs(z - c[2]) <= e[2] / 2)
    return m


def initial_velocity(n, center_cells, radius_cells, seed=3):
    rng = np.random.default_rng(seed + 1000)
    out = []
    for a in range(3):
        shape = (n + (a == 2), n + (a == 1), n + (a == 0))
        z, y, x = np.meshgrid(*(np.arange(k) for k in shape), indexing="ij")
        off = [0.5, 0.5, 0.5]
        off[a] = 0.0
        d2 = (x + off[0] - center_cells[0]) ** 2 + (y + off[1] - center_cells[1]) ** 2 + \
             (z + off[2] - center_cells[2]) ** 2
        v = np.where(d2 < radius_cells ** 2, rng.uniform(-.05, .05, shape), 0.0)
        out.append(v.astype(np.float32))
    return out


def fire_sphere(n, seed=5):
    """Sphere collider of the fire workload at its start: radius 24 cells at
    384^3, seeded centre in the upper half of the domain (cells)."""
    rng = np.random.default_rng(seed)
    return tuple(rng.uniform(.3, .7, 3) * n * np.array([1, 1, 1.2])), 24 * n / 384


def fire_path(n, seed=5):
    """The fire workload's sphere at step k: from fire_sphere(n) along a seeded
    straight line at 0.5 cells per step (an analytic collider in metres)."""
    from fluxfx_amd.physics.collision import primitive
    center, radius = fire_sphere(n, seed)
    rng = np.random.default_rng(seed + 1)
    direction = rng.normal(size=3)
    direction /= np.linalg.norm(direction)
    c0 = np.array(center) / n
    return lambda k: primitive("SPHERE", tuple(float(v) for v in c0 + direction * 0.5 * k / n), (radius / n,) * 3)


def settings_for(n, mode, workload="smoke"):
    from fluxfx_amd.physics.pressure import PressureSettings
    center_cells, radius_cells, _ = scene(n)
    dt = 0.5 / n
    kw = dict(source_center=tuple(c / n for c in center_cells), source_radius=radius_cells / n,
              source_rate=1.0 / dt, heat_source_rate=1.0 / dt, dissipation=0.0, cooling=0.0,
              thermal_lift=1.0, density_weight=0.0, source_profile="SOLID", density_mode="RATE",
              initial_temperature=0.0, velocity_advection="MACCORMACK", scalar_advection="MACCORMACK",
              pressure_tolerance=1e-5, pressure_max_cycles=60, mg_pre=2, mg_post=2, mg_coarsest=8)
    if workload == "fire":
        # fuel and heat injected by the emitter; a quarter of the fuel above the
        # ignition temperature burns per step, heat release 1.0, smoke yield 0.5;
        # vorticity confinement 0.3
        kw.update(source_rate=0.0, fuel_source_rate=1.0 / dt, combustion_enabled=True, ignition_temperature=0.3,
                  burn_rate=-math.log(0.75) / dt, heat_yield=1.0, smoke_yield=0.5, vorticity_strength=0.3,
                  moving_colliders=True)
    if mode == "mixed":
        kw.update(pressure_precision="BF16", pressure_krylov="PCG")
    elif mode == "learned":
        if not WEIGHTS.exists():
            raise NotImplementedError(f"no trained U-Net weights at {WEIGHTS}")
        # the network runs one level down (on the 8-child mean of the 256^3
        # residual), its output prolonged like a V-cycle correction
        kw.update(pressure_precision="BF16", pressure_krylov="LEARNED", learned_weights=str(WEIGHTS),
                  learned_iterations=LEARNED_ITERATIONS, learned_sweeps=LEARNED_SWEEPS,
                  learned_level=LEARNED_LEVEL)
    return PressureSettings(**kw), dt


def make_sim(device, n, mode, workload="smoke"):
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.backend.projected import DenseProjectedSmoke
    center_cells, radius_cells, boxes = scene(n)
    settings, dt = settings_for(n, mode, workload)
    grid = GridSpec((n, n, n), (1.0, 1.0, 1.0), max_dim=1024)
    if workload == "smoke":
        sim = DenseProjectedSmoke(grid, settings, device, solid_mask=box_mask(n, boxes).ravel(),
                                  active_bricks=ACTIVE_BRICKS)
    else:
        # the sphere moves 0.5 cells per step: advance() moves it before each step
        path = fire_path(n)
        sim = DenseProjectedSmoke(grid, settings, device, colliders=(path(0),), active_bricks=ACTIVE_BRICKS)
        sim.collider_path, sim.collider_step = path, 0
    sim.upload(np.zeros(n ** 3, np.float32))
    sim.upload_temperature(np.zeros(n ** 3, np.float32))
    sim.upload_velocity([v.ravel() for v in initial_velocity(n, center_cells, radius_cells)])
    return sim, dt


def advance(sim, dt):
    """One step of the workload: the fire's sphere moves to its next pose
    (mask, wall velocities, remap) before the step."""
    path = getattr(sim, "collider_path", None)
    if path is not None:
        sim.collider_step += 1
        sim.move_colliders((path(sim.collider_step),), dt)
    sim.step(dt)


def active_bricks(sim, side=8, threshold=1e-4):
    """8^3 bricks of the fire grid that hold density or fuel above 1e-4 (the
    sparse engine's activity rule, here a diagnostic of the dense step): one
    support scan per field on the device, the flags OR-ed."""
    from fluxfx_amd.native import BrickPool
    nx, ny, nz = sim.grid.shape
    flags = BrickPool.support(sim.density.tensor.view(nz, ny, nx), side, threshold)
    if sim.combustion:
        flags |= BrickPool.support(sim.combustion.fuel.tensor.view(nz, ny, nx), side, threshold)
    active, total = int(flags.sum().item()), flags.numel()
    return {"side": side, "threshold": threshold, "active": active, "total": total, "fraction": active / total}


def timed_run(device, sim, dt, steps, warmup):
    """W untimed steps, then K steps each bracketed by HIP events on the stream
    the step's kernels run on; returns the per-step event times (ms), the wall
    time of the K steps (s), the solver iterations per step and the last
    solve's metrics."""
    torch = device.torch
    for _ in range(warmup):
        advance(sim, dt)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    marks = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    cycles = []
    t0 = time.perf_counter()
    for e0, e1 in marks:
        e0.record(stream)
        advance(sim, dt)
        e1.record(stream)
        cycles.append(sim.projector.last_cycles)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    m = sim.projector.solve_metrics()
    return [e0.elapsed_time(e1) for e0, e1 in marks], wall, cycles, m


def smoother_roofline(device, sim, reps=60):
    """Average duration of the level-0 smoothing kernel of the 16-bit V-cycle
    that preconditions FPCG (the kernel with the largest share of the
    learned+mixed step), timed with HIP events on the stream it runs on.
    Algorithmic bytes per sweep: p (2 B) + rhs (2 B) + face flags (1 B) read,
    the new p (2 B) written, 7 B per cell."""
    torch = device.torch
    eng = sim.projector._precond_engine()
    n = sim.grid.cells
    eng.smooth(0, 4)
    stream = torch.cuda.current_stream()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    eng.smooth(0, reps)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / reps
    bytes_per = 7 * n
    gbs = bytes_per / (ms * 1e-3) / 1e9
    return {"kernel": "k_smooth_q<bf16, FX_OP_FLAGS, 8> (level-0 smoother of the BF16 V-cycle preconditioner)",
            "avg_ms": ms, "bytes_per_launch": bytes_per, "achieved_gbs": gbs, "launches": reps}


def measured_traffic(kernel_key="k_smooth_q<fx::bf16, 1, 8>", summary="step_256_hbm_traffic.md"):
    """HBM bytes per launch of the roofline kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (None when absent or for another kernel)."""
    path = ROOT / "profiles" / summary
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("roofline-traffic:") and kernel_key in line:
            return float(line.rsplit("|", 1)[1])
    return None


def copy_calibration(device, sim, reps=5):
    """A plain copy kernel of known size (fx_scale with factor 1 over the cell
    field: reads and writes 4 B per cell) that the FETCH_SIZE / WRITE_SIZE
    counters of a profiled run are scaled by."""
    from fluxfx_amd import _lib
    pp = sim.projector
    n = sim.grid.cells
    for _ in range(reps):
        _lib.call("fx_scale", pp.pressure.ptr, pp.spare.ptr, n, 1.0, device.stream)
    device.torch.cuda.synchronize()
    return {"kernel": "k_scale", "read_bytes": 4 * n, "write_bytes": 4 * n}


def host_cpu():
    """The host's logical CPU count, the CPUs this process may run on, and the
    CPU model as lscpu names it (/proc/cpuinfo when lscpu is absent)."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        model = next((ln.split(":", 1)[1].strip() for ln in out.splitlines() if ln.startswith("Model name")), None)
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
        except OSError:
            pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"host_cpu_count": os.cpu_count(), "host_cpus_usable": usable, "host_cpu_model": model}


def cpu_baseline(n_sample=64, steps=6):
    """FP64 NumPy oracle on the same scene resampled to n_sample^3 (1 core)."""
    from oracle import mac, transport as otr, step as ostep
    center_cells, radius_cells, boxes = scene(n_sample)
    settings, dt = settings_for(n_sample, "plain")
    og = mac.Grid((n_sample,) * 3)
    solid = box_mask(n_sample, boxes)
    vel = tuple(v.astype(np.float64) for v in initial_velocity(n_sample, center_cells, radius_cells))
    state = ostep.State(og, mac.walls(vel, og, solid), np.zeros(og.array_shape), np.zeros(og.array_shape),
                        solid=solid)
    src = [otr.Src(settings.source_center, settings.source_radius, settings.source_rate,
                   settings.heat_source_rate, solid_profile=True)]
    cfg = ostep.PressureConfig("TOLERANCE", tol=1e-5, pre=2, post=2, coarsest=8, max_cycles=60)
    cache = {}
    ostep.step(state, settings, dt, src, cfg, mg_cache=cache)  # hierarchy set-up outside the timing
    t0 = time.perf_counter()
    for _ in range(steps):
        ostep.step(state, settings, dt, src, cfg, mg_cache=cache)
    wall = time.perf_counter() - t0
    cells_per_s = steps * n_sample ** 3 / wall
    return {"value": cells_per_s / 256 ** 3, "unit": "steps/s (256^3-equivalent)", "cores": 1, "kind": "port",
            "sample": f"same scene at {n_sample}^3, {steps} steps after 1 warm-up, FP64 NumPy oracle; "
                      f"{cells_per_s:.3e} cell-updates/s scaled to 256^3",
            "cell_updates_per_s": cells_per_s, "wall_s": wall, **host_cpu()}, state


def parity_vs_oracle(device, state, n_sample, steps, modes):
    """Relative L2 error of every field after `steps` GPU steps of each mode
    against the FP64 oracle state of the same scene (pressure over fluid
    cells with the mean removed; divergence as ||after|| / ||before||)."""
    og_shape = (n_sample,) * 3
    solid = state.solid
    fluid = ~solid if solid is not None else np.ones(og_shape, bool)

    def rel(a, b, mask=None):
        if mask is not None:
            a, b = a[mask], b[mask]
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    out = {}
    for mode in modes:
        try:
            sim, dt = make_sim(device, n_sample, mode)
        except NotImplementedError:
            continue
        for _ in range(steps):
            sim.step(dt)
        get = lambda f, shape: f.tensor.detach().cpu().double().numpy().reshape(shape)
        nz = ny = nx = n_sample
        faces = [(nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx)]
        q = get(sim.projector.pressure, og_shape)
        qo = state.pressure
        qa, qb = q[fluid] - q[fluid].mean(), qo[fluid] - qo[fluid].mean()
        before, after = get(sim.projector.before, og_shape), get(sim.projector.after, og_shape)
        out[mode] = {"density": rel(get(sim.density, og_shape), state.density),
                     "temperature": rel(get(sim.temperature, og_shape), state.temperature),
                     **{"uvw"[a]: rel(get(sim._velocity[a], faces[a]), state.velocity[a]) for a in range(3)},
                     "pressure": rel(qa, qb),
                     "divergence_after_over_before": float(np.linalg.norm(after) / max(np.linalg.norm(before),
                                                                                         1e-300))}
        sim.close()
        del sim
    return out


def bricks_bench(device, n, steps, warmup, cpu):
    """Sparse brick pool at n^3 with the reference's occupancy case (a 112^3
    box of 256^3, 8.374 %, 2744 8^3 bricks, docs/SPARSE_BRICKS.md): one cycle
    is the support scan of the dense field, the gather of the active bricks
    and their scatter back (topology update excluded: host work, timed apart)."""
    import torch
    from fluxfx_amd import _lib
    from fluxfx_amd.native import BrickPool
    lo, m = (72 * n) // 256 // 8 * 8, (112 * n) // 256 // 8 * 8
    g = torch.Generator(device="cuda").manual_seed(5)
    dense = torch.zeros((n, n, n), device="cuda")
    dense[lo:lo + m, lo:lo + m, lo:lo + m] = torch.rand((m, m, m), device="cuda", generator=g) + 0.5
    out = torch.zeros_like(dense)
    pool = BrickPool(8, 4096, 2 ** 31)
    flags = BrickPool.support(dense, 8, 0.0)
    t0 = time.perf_counter()
    pool.activate_flags(flags)
    pool.sync()
    topo_ms = 1e3 * (time.perf_counter() - t0)
    nb = len(pool.snapshot())

    def cycle():
        BrickPool.support(dense, 8, 0.0)
        pool.gather(dense)
        pool.scatter(out)

    def timed(fn, reps):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    # the copy kernel of known size that rocprof's FETCH/WRITE counters are scaled by
    for _ in range(5):
        _lib.call("fx_scale", dense.data_ptr(), out.data_ptr(), n ** 3, 1.0, None)
    torch.cuda.synchronize()
    cyc_ms = timed(cycle, steps)
    sup_ms = timed(lambda: BrickPool.support(dense, 8, 0.0), 4 * steps)
    assert torch.equal(out, dense)
    coords = [r[:3] for r in pool.snapshot()]
    # sparse RK2 transport of the brick field at the reference's displacement
    # (0.6, 0.2, 0.1) cells per step (native/__init__.py compare_transport)
    uvw = [torch.full(sh, d, device="cuda") for sh, d in (((n, n, n + 1), .6), ((n, n + 1, n), .2),
                                                           ((n + 1, n, n), .1))]
    need = BrickPool.required(dense, *uvw, 8, (1e6, 1e6, 1e6, 1.0, 0.0))
    pool.activate_flags(need.cpu().numpy())
    pool.sync()
    nb_t = len(pool.snapshot())
    pool.gather(dense)
    tr_ms = timed(lambda: pool.transport(n, *uvw), steps)
    scan_bytes = n ** 3 * 4
    line = {
        "metric": f"{n}³ brick pool cycles/s (support scan + gather + scatter, 8³ bricks, 8.4 % occupancy)",
        "value": 1e3 / cyc_ms, "unit": "cycles/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": cyc_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 field, {m}^3 occupied box at {lo}, {nb} active 8^3 bricks",
                   "parallelism": "single GPU"},
        "topology_update_ms": topo_ms,
        "transport": {"ms": tr_ms, "bricks": nb_t, "displacement_cells": [.6, .2, .1],
                      "voxel_updates_per_s": nb_t * 512 / (tr_ms * 1e-3)},
        "roofline": {"bound": "hbm", "achieved": scan_bytes / (sup_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": scan_bytes / (sup_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": measured_traffic("k_brick_support_rows", "bricks_256_hbm_traffic.md"),
                     "kernel": "k_brick_support_rows (support scan of the dense field)", "avg_ms": sup_ms,
                     "bytes_per_launch": scan_bytes},
    }
    if cpu:
        from oracle import bricks as ob
        host = dense.cpu().numpy()
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 10.0 or reps == 0:
            f = ob.support_flags(host, 8, 0.0)
            bricks = ob.gather(host, coords, 8)
            reps += 1
        cpu_ms = 1e3 * (time.perf_counter() - t0) / reps
        assert int(f.sum()) == nb and bricks.shape[0] == nb
        line["cpu_baseline"] = {"value": 1e3 / cpu_ms, "unit": "cycles/s", "cores": 1, "kind": "port",
                                "sample": f"oracle support scan + gather of the same {n}^3 field, {reps} cycles "
                                          "(NumPy, one core; the scatter is not timed on the CPU)", **host_cpu()}
    print(json.dumps(line), flush=True)


def hbm_bench(device, sizes=(256, 320, 384), copy_mib=(64, 128, 256, 512, 1024, 2048), reps=40):
    """The level-0 smoother per sweep (FP32 and BF16 storage, a 16^3 solid
    block at the centre) at grid sides whose working set lies below and above
    the 256 MB last-level cache, and the copy kernel fx_scale (4 B read and
    4 B written per element, one element per thread) beside the HIP runtime's
    device-to-device copy across array sizes: the HBM rate a stencil sustains
    once its arrays no longer fit that cache."""
    import torch
    from fluxfx_amd import _lib
    from fluxfx_amd.backend.collision import SolidMask
    from fluxfx_amd.backend.tolerance import ToleranceProjector
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.physics.pressure import PressureSettings

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    smoother = []
    for n in sizes:
        for prec, per_cell in (("FP32", 13), ("BF16", 7)):
            g = GridSpec((n,) * 3, (1.0,) * 3, max_dim=1024)
            mask = np.zeros((n, n, n), np.uint8)
            mask[n // 2 - 8:n // 2 + 8, n // 2 - 8:n // 2 + 8, n // 2 - 8:n // 2 + 8] = 1
            solids = SolidMask(g, device=device, mask=mask.ravel())
            s = PressureSettings(pressure_tolerance=1e-5, mg_coarsest=8, pressure_precision=prec,
                                 pressure_krylov="PCG")
            eng = ToleranceProjector(g, s, device, solids=solids)._precond_engine()
            ms = timed(lambda: eng.smooth(0, 1))
            smoother.append({"n": n, "storage": prec, "us_per_sweep": 1e3 * ms,
                             "working_set_mb": (3 * (4 if prec == "FP32" else 2) + 1) * n ** 3 / 1e6,
                             "gbs": per_cell * n ** 3 / (ms * 1e-3) / 1e9})
            print(f"# smoother {prec} {n}^3: {1e3 * ms:.1f} us per sweep, {smoother[-1]['gbs']:.0f} GB/s",
                  file=sys.stderr, flush=True)
            del eng, solids
            torch.cuda.empty_cache()
    copies = []
    for mib in copy_mib:
        count = mib * 2 ** 20 // 4
        a = torch.ones(count, device=device.device)
        b = torch.empty_like(a)
        ms = timed(lambda: _lib.call("fx_scale", a.data_ptr(), b.data_ptr(), count, 1.0, None))
        # the HIP runtime's device-to-device copy of the same arrays, for reference
        ms_rt = timed(lambda: b.copy_(a))
        copies.append({"mib_each_way": mib, "fx_scale_us": 1e3 * ms, "fx_scale_gbs": 8 * count / (ms * 1e-3) / 1e9,
                       "runtime_copy_gbs": 8 * count / (ms_rt * 1e-3) / 1e9})
        print(f"# copy {mib} MiB in + out: fx_scale {copies[-1]['fx_scale_gbs']:.0f} GB/s, runtime copy "
              f"{copies[-1]['runtime_copy_gbs']:.0f} GB/s", file=sys.stderr, flush=True)
        del a, b
        torch.cuda.empty_cache()
    big = [r for r in smoother if r["working_set_mb"] > 512]
    line = {"metric": "level-0 smoother GB/s with its working set beyond the last-level cache",
            "value": min(r["gbs"] for r in big) if big else None, "unit": "GB/s", "n_gpus": 1, "steps": reps,
            "warmup": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32/bf16",
            "data": "synthetic", "config": {"workload": "smoother sweeps and copies across sizes"},
            "smoother": smoother, "copy": copies, "peak_gbs": HBM_PEAK_GBS}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=None, help="grid side (default 256 smoke, 384 fire)")
    ap.add_argument("--modes", default="plain,mixed,learned")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("smoke", "fire", "bricks", "hbm"), default="smoke",
                    help="smoke: the 256^3 headline scene; fire: 384^3 dense fire (combustion, vorticity "
                         "confinement, a sphere collider moving 0.5 cells per step), no CPU leg; bricks: the "
                         "256^3 sparse brick pool; hbm: smoother and copy bandwidth across sizes")
    ap.add_argument("--weights", default=None, help="U-Net weights (.npz) of the learned mode")
    ap.add_argument("--dense-scalars", action="store_true",
                    help="scalar predictor, transport and combustion over the whole grid")
    args = ap.parse_args()
    global WEIGHTS, ACTIVE_BRICKS
    if args.weights:
        WEIGHTS = Path(args.weights).resolve()
    ACTIVE_BRICKS = not args.dense_scalars
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the step runs on a single MI355X; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    from fluxfx_amd.backend.device import HipDevice
    device = HipDevice(0)
    fire = args.workload == "fire"
    n = args.n or (384 if fire else 256)
    if args.workload == "bricks":
        bricks_bench(device, n, args.steps, args.warmup, not args.no_cpu)
        return
    if args.workload == "hbm":
        hbm_bench(device)
        return
    results = {}
    roof = None
    for mode in [m for m in args.modes.split(",") if m]:
        try:
            sim, dt = make_sim(device, n, mode, args.workload)
            step_ms, wall, cycles, m = timed_run(device, sim, dt, args.steps, args.warmup)
        except NotImplementedError as exc:
            results[mode] = {"status": f"not built: {exc}"}
            continue
        med = float(np.median(step_ms))
        results[mode] = {"steps_per_s": 1e3 / med, "ms_per_step": med, "ms_per_step_mean": float(np.mean(step_ms)),
                         "ms_per_step_min": float(np.min(step_ms)), "ms_per_step_max": float(np.max(step_ms)),
                         "wall_steps_per_s": args.steps / wall,
                         "mean_iterations": float(np.mean(cycles)), "relative_residual": m.get("relative_residual"),
                         "converged": m.get("converged"), "fallback": m.get("fallback")}
        results[mode]["scalar_bricks"] = sim.active.fractions() if sim.active else None
        if fire:
            results[mode]["active_bricks"] = active_bricks(sim)
        print(f"# {mode}: {1e3 / med:.2f} steps/s (median of {args.steps} event-timed steps {med:.3f} ms, wall "
              f"{1e3 * wall / args.steps:.3f} ms/step), iterations {np.mean(cycles):.2f}, residual "
              f"{m.get('relative_residual')}", file=sys.stderr, flush=True)
        if mode in ("mixed", "learned"):
            # the 16-bit level-0 smoother; the learned mode's figure replaces the mixed one
            roof = smoother_roofline(device, sim)
            copy_calibration(device, sim)
        del sim
        device.torch.cuda.empty_cache()
    ok = {k: v for k, v in results.items() if "steps_per_s" in v and v.get("converged")}
    # the headline is the learned+mixed step; without trained weights the mixed one
    best_mode = next((m for m in ("learned", "mixed", "plain") if m in ok), None)
    value = ok[best_mode]["steps_per_s"] if best_mode else 0.0
    plain = results.get("plain", {}).get("steps_per_s")
    if fire:
        metric = (f"{n}³ fire steps/s at 1e-5 rel. residual (dense grid: combustion, vorticity confinement, "
                  f"moving sphere)")
        workload = (f"{n}^3 fire step: fuel and heat emitter, combustion (25% burn per step above T=0.3, heat "
                    f"1.0, smoke 0.5), vorticity confinement 0.3, sphere collider moving 0.5 cells per step, "
                    f"MacCormack, projected to 1e-5 relative residual")
    else:
        metric = METRIC
        workload = f"{n}^3 smoke step with 8 box colliders, MacCormack, projected to 1e-5 relative residual"
    line = {
        "metric": metric, "value": value, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 / value if value else None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "grid": [n, n, n], "dt": 0.5 / n, "mode": best_mode,
                   "scalar_passes": "active 8^3 bricks (exact)" if ACTIVE_BRICKS else "whole grid",
                   "parallelism": "single GPU"},
        "modes": results,
        "speedup_vs_plain": (value / plain) if (plain and value) else None,
        "mixed_vs_plain": (results["mixed"]["steps_per_s"] / plain) if plain and "steps_per_s" in
        results.get("mixed", {}) else None,
        "cell_updates_per_s": value * n ** 3,
    }
    if roof:
        line["roofline"] = {"bound": "hbm", "achieved": roof["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": roof["achieved_gbs"] / HBM_PEAK_GBS, "traffic": measured_traffic(),
                            "kernel": roof["kernel"], "avg_ms": roof["avg_ms"],
                            "bytes_per_launch": roof["bytes_per_launch"]}
    if not args.no_cpu and not fire:
        base, state = cpu_baseline()
        line["cpu_baseline"] = base
        line["fp64_parity_64"] = parity_vs_oracle(device, state, 64, 7, [m for m in results if "steps_per_s" in
                                                                          results[m]])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
