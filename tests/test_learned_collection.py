"""Trajectory collection for the learned warm start: what happens at a failed step (host logic only,
with a stand-in model and time loop; runs without a GPU).

A failed step leaves the model at its old level with dt cut to a quarter, as adhk_time_loop does
(tc_scale.c:3-19).  Collection takes that step again as four quarter steps, collects neither them nor
the step after them (whose features would hold a quarter step's change as the previous change), and
goes on at the nominal dt.
"""
import pytest

from adh_kraken_amd import _lib as L
from adh_kraken_amd import learned as LR


class FakeModel:
    def __init__(self, dt, fail_at=(), quarter_fails=False, forcing_ends_at=None):
        self.dt, self.fail_at, self.quarter_fails = dt, set(fail_at), quarter_fails
        self.forcing_ends_at = forcing_ends_at
        self.calls = []  # (nsteps asked, dt) per time_loop call
        self.nominal_steps = 0

    def get_dt(self):
        return self.dt

    def set_dt(self, dt):
        self.dt = dt


def fake_time_loop(sm, nsteps, precision=None):
    sm.calls.append((nsteps, sm.dt))
    if nsteps == 1:
        sm.nominal_steps += 1
        if sm.forcing_ends_at is not None and sm.nominal_steps >= sm.forcing_ends_at:
            return 0, {"newton_its": 0}  # the series does not cover the step: not taken, dt unchanged
        if sm.nominal_steps in sm.fail_at:
            sm.dt *= 0.25  # the time loop's cut; the old level stays
            return 0, {"newton_its": 7}
        return 1, {"newton_its": 3}
    return (0 if sm.quarter_fails else nsteps), {"newton_its": 2 * nsteps}


class FakeMLP:
    def __init__(self):
        self.collected = []

    def collect(self, sm):
        self.collected.append(sm.nominal_steps)


@pytest.fixture(autouse=True)
def fake_loop(monkeypatch):
    monkeypatch.setattr(LR, "time_loop", fake_time_loop)


def test_no_failure_collects_every_step():
    sm, mlp = FakeModel(2.0), FakeMLP()
    its, k = LR.collect_trajectory(sm, mlp, 5, strict=False)
    assert (its, k) == (15, 5) and mlp.collected == [1, 2, 3, 4, 5]
    assert all(c == (1, 2.0) for c in sm.calls)


def test_failed_step_is_taken_in_quarter_steps_and_collection_goes_on():
    sm, mlp = FakeModel(2.0, fail_at=[3]), FakeMLP()
    its, k = LR.collect_trajectory(sm, mlp, 6, strict=False)
    # nominal steps 1, 2 collected; 3 fails -> four quarter steps; 4 is the step after them, not
    # collected; 5 and 6 collected
    assert sm.calls == [(1, 2.0), (1, 2.0), (1, 2.0), (4, 0.5), (1, 2.0), (1, 2.0), (1, 2.0)]
    assert mlp.collected == [1, 2, 5, 6]
    assert k == 4 and its == 3 * 5 + 7 + 8
    assert sm.dt == 2.0


def test_failed_quarter_steps_end_the_trajectory():
    sm, mlp = FakeModel(2.0, fail_at=[2], quarter_fails=True), FakeMLP()
    its, k = LR.collect_trajectory(sm, mlp, 6, strict=False)
    assert sm.calls == [(1, 2.0), (1, 2.0), (4, 0.5)]
    assert mlp.collected == [1] and k == 1


def test_strict_collection_raises_at_the_failed_step():
    sm, mlp = FakeModel(2.0, fail_at=[2]), FakeMLP()
    with pytest.raises(L.AdhkError, match="time step failed while collecting"):
        LR.collect_trajectory(sm, mlp, 6)
    assert mlp.collected == [1]


def test_forcing_stop_ends_the_trajectory_without_quarter_steps():
    """a step that the boundary series do not cover is not taken and dt is not cut: collection ends there
    instead of stepping past the series' end in quarter steps"""
    sm, mlp = FakeModel(2.0, forcing_ends_at=4), FakeMLP()
    its, k = LR.collect_trajectory(sm, mlp, 10, strict=False)
    assert (its, k) == (9, 3) and mlp.collected == [1, 2, 3]
    assert all(c == (1, 2.0) for c in sm.calls) and len(sm.calls) == 4
    sm2, mlp2 = FakeModel(2.0, forcing_ends_at=2), FakeMLP()
    assert LR.collect_trajectory(sm2, mlp2, 5, strict=True) == (3, 1)
