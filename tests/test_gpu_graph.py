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


@pytest.mark.parametrize("case", ["serial-ou-localized", "serial-white", "batch"])
def test_enkf_cycle_replays_match_eager(pkg, case):
    """A whole EnKF cycle (forecast steps, inflation, forecast RMSE, device
    perturbations, analysis, mean and spread) replays one captured graph from
    the third group on; records, mean fields and states equal the launch-by-
    launch run bit for bit, group by group, with host reads in between."""
    n, M, dt, every, groups = 32, 24, 2e-3, 5, 6
    rng = np.random.default_rng(9)
    rows = []
    for g in range(groups):
        xs, ys = rng.integers(0, n, 12), rng.integers(0, n, 12)
        rows += [((g + 1) * every, x, y, v, 0.1) for x, y, v in
                 zip(xs, ys, rng.normal(0.0, 0.3, 12))]
    rows = np.array(rows, dtype=np.float64)
    precision = "fp32" if case == "batch" else "fp64"
    a, b = pair(pkg, n, M, precision, seed=31, nu=2e-3, sigma=0.2, tau=0.4)
    mode = "batch" if case == "batch" else "serial"
    inflation, radius = (1.05, 4.0) if case == "serial-ou-localized" else (1.0, 0.0)
    with a, b:
        if case == "serial-ou-localized":
            a.set_forcing_process("ou")
            b.set_forcing_process("ou")
        a.enkf_begin(17, mode)
        b.enkf_begin(17, mode)
        steps = rows[:, 0].astype(int)
        current = 0
        for target in np.unique(steps):
            sel = rows[steps == target]
            ra, ma = a.enkf_assimilate(dt, int(target - current), sel[:, 1], sel[:, 2],
                                       sel[:, 3], sel[:, 4], inflation, radius, mean=True)
            rb, mb = b.enkf_assimilate(dt, int(target - current), sel[:, 1], sel[:, 2],
                                       sel[:, 3], sel[:, 4], inflation, radius, mean=True)
            assert np.array_equal(ra, rb), (target, ra, rb)
            assert np.array_equal(ma, mb)
            assert np.array_equal(a.get_state(), b.get_state()), target
            current = int(target)
        assert a.step_index == b.step_index == current
        # a forecast after the filter continues the same noise stream
        a.step(dt, 3)
        b.step(dt, 3)
        assert np.array_equal(a.get_state(), b.get_state())


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_joint_filter_cycle_replays_match_eager(pkg, precision):
    """The mixed-grid filter's cycle (fork to both grids' streams for steps and
    log-weights, join for normalisation and the cross-grid gather) replays one
    captured graph; every group's record and the final states equal the run
    with graph replay off, bit for bit."""
    rng = np.random.default_rng(7)
    rows = []
    for g in range(6):
        xs, ys = rng.integers(0, 32, 12), rng.integers(0, 32, 12)
        sd = 0.05 if g % 2 == 0 else 5.0
        rows += [((g + 1) * 4, x, y, v, sd) for x, y, v in zip(xs, ys, rng.normal(0, 0.3, 12))]
    rows = np.array(rows, dtype=np.float64)
    runs = []
    for graphs in (True, False):
        with pkg.Ensemble(32, 10, precision, nu=2e-3, sigma=0.2, amp=0.1, seed=1) as a, \
                pkg.Ensemble(64, 6, precision, nu=2e-3, sigma=0.2, amp=0.1, seed=2) as b:
            a.init_grf()
            b.init_grf()
            a.set_graph(graphs)
            b.set_graph(graphs)
            res = pkg.joint_bootstrap_filter([a, b], rows, 2e-3, 13)
            runs.append((res, a.get_state(), b.get_state(), a.step_index, b.step_index))
    (ra, a1, b1, ia, ib), (rb, a2, b2, ja, jb) = runs
    flags = [h["resampled"] for h in rb["history"]]
    assert any(flags) and (ia, ib) == (ja, jb) == (24, 24)
    assert ra["log_likelihood"] == rb["log_likelihood"]
    for ha, hb in zip(ra["history"], rb["history"]):
        for k in ("increment", "ess", "max_weight", "resampled"):
            assert ha[k] == hb[k], k
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
