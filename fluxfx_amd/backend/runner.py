"""Host driver of the live step loop: timestep choice, collider motion, the
step and the completion fence, paced by the wall clock.

This is the part of the reference's add-on runtime that calls the hot path,
without Blender: ``step`` follows ``runtime.step`` (blender/runtime.py:170-196)
and ``tick`` follows ``runtime.tick`` with its ``advance`` callback
(blender/runtime.py:221-303) and ``PlaybackClock.run``
(physics/playback.py:27-40).  Per substep:

1. dt from the device maxima (``AdaptiveTimestep.select``, the only host
   read besides the fence) or the fixed step, capped by the time owed;
2. with moving colliders, dt shortened so no wall moves more than half a
   cell, and the colliders moved along their path;
3. every source swept over its share [begin, end] of the frame's path;
4. the step, then ``StepCompletion.wait``.
"""
import time
from dataclasses import replace

from ..physics.collider_motion import MotionPath
from ..physics.emission import Emission, Source
from ..physics.playback import PlaybackClock, source_span
from .completion import StepCompletion
from .timestep import AdaptiveTimestep


class LiveRunner:
    def __init__(self, solver, *, time_step=1 / 30, cfl=0.75, adaptive=True, clock=time.perf_counter,
                 max_lag=0.25):
        self.solver = solver
        self.time_step = time_step
        self.cfl = cfl
        self.adaptive = adaptive
        self.clock = clock
        self.timestep = AdaptiveTimestep(solver.grid, solver.device)
        self.completion = StepCompletion(solver.device)
        self.playback = PlaybackClock(clock(), max_lag)
        self.timestep_report = None
        self.collider_report = None

    def _choose(self, maximum, emission, sources):
        if not self.adaptive:
            return {"mode": "FIXED", "dt": maximum}
        return self.timestep.select(self.solver, maximum, self.cfl, emission, sources)

    def _primary(self, emission):
        s = self.solver.settings
        return Source(s.source_center, s.source_radius, s.source_rate, s.heat_source_rate,
                      emission or Emission(s.source_center), s.density_mode, s.source_profile, s.fuel_source_rate)

    def _step_with(self, dt, spans, several):
        """One solver step with the swept sources: the table path for several
        sources, else the settings' single emitter moved to the span's end."""
        sim = self.solver
        if several:
            sim.step(dt, sources=spans)
            return
        settings = sim.settings
        sim.update_settings(replace(settings, source_center=spans[0].center))
        try:
            sim.step(dt, spans[0].motion)
        finally:
            sim.update_settings(settings)

    def step(self, emission=None, sources=None, colliders=None):
        """One interactive step (runtime.step): returns the timestep report."""
        sim = self.solver
        report = self._choose(self.time_step, emission, sources)
        if colliders is not None:
            path = MotionPath(sim.solids.colliders, colliders, self.time_step)
            report["dt"] = path.limit_dt(report["dt"], sim.grid.cell_size)
            sim.move_colliders(path.at(report["dt"]), report["dt"])
            self.collider_report = {"remaining_metres": max(0.0, path.distance - path.speed * report["dt"])}
        sim.step(report["dt"], emission, sources)
        self.completion.wait(sim)
        self.timestep_report = report
        return report

    def tick(self, emission=None, sources=None, colliders=None, budget=1 / 60, now=None):
        """Pay the wall time owed since the last tick in substeps (runtime.tick);
        ``emission``/``sources`` sweep from their start to their centre over the
        time owed, ``colliders`` is where the colliders are at its end.
        Returns the delay before the next tick."""
        started = self.clock() if now is None else now
        self.playback.begin(started)
        owed = max(self.playback.debt, 1e-6)
        extra = list(sources or ())
        sampled = [self._primary(emission)] + extra
        sim = self.solver
        path = None
        if colliders is not None and sim.settings.moving_colliders and sim.solids is not None:
            path = MotionPath(sim.solids.colliders, colliders, owed)
        consumed = 0.0

        def advance(remaining):
            nonlocal consumed
            report = self._choose(min(self.time_step, remaining), emission, sampled)
            dt = report["dt"]
            if path is not None:
                dt = report["dt"] = path.limit_dt(dt, sim.grid.cell_size)
                sim.move_colliders(path.at(consumed + dt), dt)
                self.collider_report = {"remaining_metres": max(0.0, path.distance - path.speed * (consumed + dt))}
            begin, end = min(1.0, consumed / owed), min(1.0, (consumed + dt) / owed)
            self._step_with(dt, [source_span(s, begin, end) for s in sampled], bool(extra))
            self.completion.wait(sim)
            consumed += dt
            self.timestep_report = report
            return dt

        return self.playback.run(advance, budget, self.clock, started=started)

    def close(self):
        self.timestep.close()
        self.completion.close()
