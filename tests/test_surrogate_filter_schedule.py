"""Host logic of engine.surrogate_filter on the CPU: which group forecasts with
the network or the fallback, when the live judge runs (waited for, or
enqueued beside the cycles), and what the history records.  The particles and
the surrogate are stand-ins that log the calls the filter makes (the device
work behind them is covered by the GPU tests)."""
from __future__ import annotations

import numpy as np
import pytest


class Particles:
    def __init__(self, log):
        self.log = log

    def pf_begin(self, seed, max_groups):
        self.log.append(("pf_begin", seed, max_groups))

    def pf_assimilate(self, dt, steps, x, y, v, sd, is_last, mean=True):
        self.log.append(("assimilate", len(x), bool(is_last)))
        return np.array([-1.0, 10.0, 0.2, 1.0]), (np.zeros((4, 4)) if mean else None)

    def pf_result(self):
        return -5.0, 3


class Surrogate:
    """Records calls; an enqueued judge finishes after `lag` polls without
    waiting, and its estimate is the group index it started at."""

    def __init__(self, log, lag):
        self.log, self.lag = log, lag
        self.error_estimate = 0.5
        self.pending = None  # (estimate, polls left)

    def judge(self, particles, judge, dt, steps, precision):
        self.log.append(("judge", steps))
        self.error_estimate = 1.0
        return 1.0

    def judge_start(self, particles, judge, dt, steps, precision):
        assert self.pending is None, "one judge in flight at a time"
        self.log.append(("judge_start", steps))
        self.pending = (len([c for c in self.log if c[0] == "assimilate"]), self.lag)

    def judge_poll(self, wait=False):
        self.log.append(("judge_poll", bool(wait)))
        if self.pending is None:
            return False, self.error_estimate
        est, left = self.pending
        if wait or left <= 1:
            self.pending = None
            self.error_estimate = float(est)
            return True, self.error_estimate
        self.pending = (est, left - 1)
        return False, self.error_estimate

    def forecast(self, particles, dt, steps, precision):
        self.log.append(("forecast", steps))

    def forecast_or_solve(self, particles, dt, steps, threshold, precision):
        self.log.append(("forecast_or_solve", steps))
        return self.error_estimate > threshold, 0.25


def rows(groups, every=50):
    return np.array([[(g + 1) * every, g % 8, 1, 0.1, 0.2] for g in range(groups)
                     for _ in range(2)])


def run(pkg, groups, judge_every, judge_async, lag=2, threshold=0.75):
    log = []
    sur = Surrogate(log, lag)
    res = pkg.surrogate_filter(Particles(log), sur, rows(groups), 1e-3, 7,
                               fallback_threshold=threshold, judge=object(),
                               judge_every=judge_every, judge_async=judge_async)
    return log, res


def test_waited_judges_run_every_k_groups(pkg):
    log, res = run(pkg, groups=7, judge_every=3, judge_async=False)
    assert [c for c in log if c[0] == "judge"] == [("judge", 50)] * 3  # groups 0, 3, 6
    assert not any(c[0] in ("judge_start", "judge_poll") for c in log)
    assert [h["judged"] for h in res["history"]] == [g % 3 == 0 for g in range(7)]
    # every group forecasts through the fallback comparison, then assimilates
    kinds = [c[0] for c in log if c[0] in ("judge", "forecast_or_solve", "assimilate")]
    assert kinds[:3] == ["judge", "forecast_or_solve", "assimilate"]
    assert kinds.count("forecast_or_solve") == kinds.count("assimilate") == 7
    assert res["log_likelihood"] == -5.0 and res["resampling_events"] == 3


@pytest.mark.parametrize("lag", [1, 2, 5])
def test_enqueued_judges_after_the_first(pkg, lag):
    """Group 0 waits for its judge; from then on a judge is started every k
    groups (after waiting for the previous one), every other group only polls,
    a finished judge's estimate is used from the group that folds it in, and
    the last one is waited for after the final group."""
    groups, k = 10, 3
    log, res = run(pkg, groups, k, judge_async=True, lag=lag)
    assert [c[0] for c in log[:3]] == ["pf_begin", "judge", "forecast_or_solve"]
    assert [c for c in log if c[0] == "judge_start"] == [("judge_start", 50)] * 3  # 3, 6, 9
    polls = [c for c in log if c[0] == "judge_poll"]
    assert len(polls) == groups - 1 + 1  # groups 1..9, then the final wait
    assert [w for _, w in polls] == [g % k == 0 for g in range(1, groups)] + [True]
    judged = [h["judged"] for h in res["history"]]
    assert judged[0]
    # the judge started at group 3 lands at min(3 + lag, 6), the one at 6 by 9
    assert judged.index(True, 1) == min(3 + lag, 6)
    assert sum(judged) == 3
    # each group's estimate: 1.0 (group 0's judge) until the next judge lands
    est = [h["error_estimate"] for h in res["history"]]
    first = min(3 + lag, 6)
    assert est[:first] == [1.0] * first
    assert est[first] == 3.0  # started after the assimilation of groups 0..2
    assert log[-1] == ("judge_poll", True)


def test_fallback_follows_the_estimate(pkg):
    """forecast_or_solve is asked every group; with the threshold below the
    first judge's estimate every group falls back to the solver."""
    log, res = run(pkg, groups=4, judge_every=2, judge_async=False, threshold=0.5)
    assert all(h["solver"] for h in res["history"])
    log, res = run(pkg, groups=4, judge_every=2, judge_async=False, threshold=2.0)
    assert not any(h["solver"] for h in res["history"])
