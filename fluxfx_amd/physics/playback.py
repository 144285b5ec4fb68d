"""Wall-clock accounting of the live loop and the splitting of swept sources
into substeps.

Interface of physics/playback.py:9-70: ``PlaybackClock(now, max_lag=0.25)``
with ``begin(now)``, ``run(advance, budget, clock, max_steps=32,
started=None)``, the ``debt``/``simulated``/``dropped``/``report`` fields,
and ``source_span(source, begin, end)``, with the reference's errors.

The clock owes the simulation the wall time that passed since the last call
(its debt), but never more than ``max_lag`` seconds: a stall (a UI freeze, a
laptop lid) is recorded as dropped time, not replayed.  ``run`` pays the
debt in solver steps until it is below a microsecond, the step budget is
spent or the callback has used its time budget; a step that has started is
never cut short, so at least one step runs whenever there is debt.
"""
from dataclasses import replace
from math import isfinite

from .emission import Emission

MIN_REMAINDER = 1e-6
FRAME = 1.0 / 30.0
IDLE = 0.001  # the shortest timer interval the host is asked for


class PlaybackClock:
    def __init__(self, now, max_lag=0.25):
        self.epoch = self.last = now
        self.max_lag = max_lag
        self.debt = 0.0
        self.simulated = 0.0
        self.dropped = 0.0
        self.report = None

    def begin(self, now):
        """Add the wall time since the last call to the debt; return the part
        of it beyond ``max_lag`` that is dropped."""
        self.debt += max(0.0, now - self.last)
        self.last = now
        excess = max(0.0, self.debt - self.max_lag)
        self.debt -= excess
        self.dropped += excess
        return excess

    def run(self, advance, budget, clock, max_steps=32, started=None):
        """Call ``advance(debt) -> dt`` until the debt is paid, ``max_steps``
        steps ran or ``budget`` seconds of wall time (counted from ``started``)
        are used after the first step; return the delay before the next call."""
        start = clock() if started is None else started
        steps = 0
        while self.debt >= MIN_REMAINDER and steps < max_steps:
            if steps and clock() - start >= budget:
                break
            dt = advance(self.debt)
            if not (isfinite(dt) and 0 < dt <= self.debt + 1e-10):
                raise ValueError("Playback step must consume positive available time")
            self.debt = max(0.0, self.debt - dt)
            self.simulated += dt
            steps += 1
        end = clock()
        spent = end - start
        elapsed = max(0.0, end - self.epoch)
        behind = self.debt >= MIN_REMAINDER
        self.report = {"callback_ms": 1000.0 * spent, "substeps": steps,
                       "speed": self.simulated / elapsed if elapsed else 0.0,
                       "lag_ms": 1000.0 * (self.debt + max(0.0, end - self.last)),
                       "dropped_seconds": self.dropped, "limited": behind}
        # behind: yield briefly and come back; otherwise sleep out the frame
        return IDLE if behind else max(IDLE, FRAME - spent)


def source_span(source, begin, end):
    """The part of a source swept during the fractions [begin, end] of its
    path from ``motion.start`` to ``center``: consecutive spans join end to
    start, and each keeps the rates (per second) and the jet."""
    if not 0 <= begin <= end <= 1:
        raise ValueError("Source fractions must be ordered within [0, 1]")
    a, b = source.motion.start, source.center

    def at(f):
        return tuple(x + (y - x) * f for x, y in zip(a, b))

    return replace(source, center=at(end), motion=Emission(at(begin), source.motion.velocity, source.motion.coupling))
