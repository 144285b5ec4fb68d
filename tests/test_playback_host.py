"""Playback clock and swept-source splitting of the live loop
(fluxfx_amd.physics.playback) against the reference's own cases
(FluxFX_Source/tests/test_playback.py:11-65)."""
import pytest

from fluxfx_amd.physics.emission import Emission, Source
from fluxfx_amd.physics.playback import PlaybackClock, source_span


class FakeClock:
    def __init__(self):
        self.now = 0.0

    def __call__(self):
        return self.now


def test_catches_up_multiple_steps_without_advancing_future():
    clock, playback = FakeClock(), PlaybackClock(0)
    clock.now = .033
    playback.begin(clock())

    def advance(remaining):
        clock.now += .002
        return min(.01, remaining)
    delay = playback.run(advance, .024, clock)
    assert playback.report["substeps"] == 4
    assert playback.simulated == pytest.approx(.033)
    assert playback.report["lag_ms"] == pytest.approx(8)
    assert delay == pytest.approx(1 / 30 - .008)


def test_completed_work_budget_and_resume_debt():
    clock, playback = FakeClock(), PlaybackClock(0)
    clock.now = .1
    playback.begin(clock())

    def advance(remaining):
        clock.now += .015
        return min(.01, remaining)
    assert playback.run(advance, .024, clock) == .001
    assert playback.report["substeps"] == 2
    assert playback.debt == pytest.approx(.08)
    assert playback.report["callback_ms"] == pytest.approx(30)
    clock.now += .001
    playback.begin(clock())
    assert playback.debt == pytest.approx(.111)


def test_one_expensive_step_yields_and_step_count_is_bounded():
    clock, playback = FakeClock(), PlaybackClock(0)
    playback.begin(.1)

    def advance(remaining):
        clock.now += .2
        return .001
    playback.run(advance, .001, clock)
    assert playback.report["substeps"] == 1
    playback.run(lambda remaining: .001, .1, clock, max_steps=3)
    assert playback.report["substeps"] == 3


def test_sleep_discard_is_visible_not_simulated():
    playback = PlaybackClock(10)
    assert playback.begin(20) == pytest.approx(9.75)
    assert playback.simulated == 0 and playback.debt == .25 and playback.dropped == 9.75
    resumed = PlaybackClock(30)
    resumed.begin(30.01)
    assert resumed.debt == pytest.approx(.01)


@pytest.mark.parametrize("dt", [0, -1, float("nan"), .2])
def test_invalid_step_rejected_without_accounting(dt):
    playback = PlaybackClock(0)
    playback.begin(.1)
    with pytest.raises(ValueError):
        playback.run(lambda _: dt, .02, lambda: .1)
    assert playback.simulated == 0


def test_source_path_subdivision_and_rate_preservation():
    source = Source((1, .5, .5), .1, 3, 2, Emission((0, .5, .5), (2, 0, 0), 4))
    spans = [source_span(source, i / 4, (i + 1) / 4) for i in range(4)]
    assert spans[0].motion.start == source.motion.start and spans[-1].center == source.center
    for previous, following in zip(spans, spans[1:]):
        assert previous.center == following.motion.start
    assert sum(span.density_rate * .025 for span in spans) == pytest.approx(.3)
    assert all(span.motion.velocity == (2, 0, 0) for span in spans)
    with pytest.raises(ValueError):
        source_span(source, .8, .2)
