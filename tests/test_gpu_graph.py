"""Captured HIP graphs replay the launch-by-launch results exactly.

Each ensemble runs twice from the same seed: graphs on (steps and filter
cycles are recorded on the second call with a key and replayed afterwards,
reading the step index, filter group and last-group flag from device memory)
and graphs off.  Every replay is compared bit for bit with the eager run,
with host work and synchronisations between the replays.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def pair(pkg, n, members, precision, seed=11, **kw):
    args = dict(nu=1e-3, sigma=0.05, amp=0.1, kf=4, seed=seed)
    args.update(kw)
    a = pkg.Ensemble(n, members, precision, **args)
    b = pkg.Ensemble(n, members, precision, **args)
    b.set_graph(False)
    a.init_grf()
    b.init_grf()
    return a, b


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_step_replays_match_eager(pkg, precision):
    a, b = pair(pkg, 64, 6, precision)
    with a, b:
        # 5-step calls: eager, capture, replays; a second key (7) in between,
        # and 450 steps = two replays of a 200-step graph + 50 eager steps
        for steps in (5, 5, 5, 7, 5, 7, 7, 450, 5, 450):
            a.step(1e-3, steps)
            b.step(1e-3, steps)
            ga, gb = a.get_state(), b.get_state()
            assert np.array_equal(ga, gb), f"state differs after a {steps}-step call"
            a.synchronize()
            b.synchronize()
            assert a.step_index == b.step_index
        da, db = a.diagnostics(), b.diagnostics()
        assert np.array_equal(da, db)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_filter_cycle_replays_match_eager(pkg, precision):
    n, P, dt, every, groups = 32, 32, 2e-3, 6, 9
    rng = np.random.default_rng(5)
    rows = []
    for g in range(groups):
        xs, ys = rng.integers(0, n, 16), rng.integers(0, n, 16)
        # alternate sharp and loose observations: some groups resample
        sd = 0.05 if g % 2 == 0 else 5.0
        vals = rng.normal(0.0, 0.3, 16)
        rows += [((g + 1) * every, x, y, v, sd) for x, y, v in zip(xs, ys, vals)]
    rows = np.array(rows, dtype=np.float64)
    a, b = pair(pkg, n, P, precision, seed=21, nu=2e-3, sigma=0.2)
    with a, b:
        ra = pkg.bootstrap_filter(a, rows, dt, seed=4)
        rb = pkg.bootstrap_filter(b, rows, dt, seed=4)
        sa, sb = a.get_state(), b.get_state()
    flags = [h["resampled"] for h in rb["history"]]
    print(precision, flags, ra["log_likelihood"], rb["log_likelihood"])
    assert any(flags) and not all(flags[:-1])
    assert not flags[-1]  # never after the last group (src/filter.cpp:102)
    assert ra["log_likelihood"] == rb["log_likelihood"]
    assert ra["resampling_events"] == rb["resampling_events"]
    for ha, hb in zip(ra["history"], rb["history"]):
        for k in ("increment", "ess", "max_weight", "resampled"):
            assert ha[k] == hb[k], k
        assert np.array_equal(ha["mean_field"], hb["mean_field"])
    assert np.array_equal(sa, sb)
