"""GPU particle filter parity (members are particles), through the C-ABI.

Checks:
  * the reference's own bootstrap_filter in the deterministic limit
    (tests/verification.cpp:184-197) from the committed golden vectors:
    log-likelihood, per-group increments / ESS / max weight, mean fields;
  * the FP64 oracle filter (tests/oracle_filters.py) with stochastic forcing:
    identical resampling decisions and ancestors (through the states), log-
    likelihood to 1e-10 relative in FP64 and 1e-4 absolute in FP32.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle_filters import change_grid, oracle_bootstrap_filter, oracle_joint_filter

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def full_from_half(h, n):
    K = (n - 1) // 3
    out = np.zeros((n, n), dtype=np.complex128)
    for c in range(K + 1):
        for j, ky in enumerate(range(-K, K + 1)):
            out[ky % n, c] = h[c, j]
            if c > 0:
                out[(-ky) % n, n - c] = np.conj(h[c, j])
    return out


def test_pf_deterministic_limit_matches_reference_golden(pkg):
    g = np.load(os.path.join(GOLDEN, "pf_deterministic_16.npz"))
    meta = json.load(open(os.path.join(GOLDEN, "pf_deterministic_16.json")))
    n, P = 16, meta["particles"]
    w0 = full_from_half(g["initial"], n)
    with pkg.Ensemble(n, P, "fp64", nu=meta["nu"], alpha=meta["alpha"], sigma=0.0, amp=0.0,
                      kmin=2, kmax=3, seed=meta["seed"]) as ens:
        ens.set_state(np.stack([w0] * P))
        res = pkg.bootstrap_filter(ens, g["rows"], meta["dt"], meta["seed"])
    assert abs(res["log_likelihood"] - meta["log_likelihood"]) < 1e-10
    assert res["resampling_events"] == meta["resampling_events"] == 0
    for h, rec, mean in zip(res["history"], g["records"], g["means"]):
        assert abs(h["increment"] - rec[0]) < 1e-10
        assert abs(h["ess"] - rec[1]) < 1e-9 and abs(h["max_weight"] - rec[2]) < 1e-12
        phys = (np.fft.ifft2(h["mean_field"]) * n * n).real
        np.testing.assert_allclose(phys, mean, rtol=0, atol=1e-10)


def synthetic(orc, n, nu, dt, groups, every, nobs, truth_key, seed, sd=0.05):
    rng = np.random.default_rng(seed)
    w = orc.grf(n, truth_key)
    rows, step = [], 0
    for g in range(groups):
        w = orc.trajectory(n, nu, 0.0, dt, every, truth_key, step0=step, initial=w)
        step += every
        ph = (orc.fft2d(w, inverse=True)).real
        xs, ys = rng.integers(0, n, nobs), rng.integers(0, n, nobs)
        vals = ph[ys, xs] + sd * rng.standard_normal(nobs)
        rows += [(step, x, y, v, sd) for x, y, v in zip(xs, ys, vals)]
    return np.array(rows, dtype=np.float64), w


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_pf_matches_oracle_with_stochastic_forcing(pkg, orc, precision):
    n, P, nu, dt, seed = 32, 16, 2e-3, 2e-3, 77
    # truth starts from the prior mean; observation noise sd 1 keeps ESS in the
    # middle range (groups resample / do not resample / resample / last)
    rows, _ = synthetic(orc, n, nu, dt, groups=4, every=10, nobs=24, truth_key=999, seed=3,
                        sd=1.0)
    base = orc.grf(n, 999)
    init = np.stack([base + 0.05 * orc.grf(n, orc.stream_seed(5, i)) for i in range(P)])
    want = oracle_bootstrap_filter(orc, init, rows, dt, seed, nu, sigma=0.2)
    with pkg.Ensemble(n, P, precision, nu=nu, sigma=0.2, amp=0.1, kf=4, seed=seed) as ens:
        ens.set_state(init)
        got = pkg.bootstrap_filter(ens, rows, dt, seed)
    print(precision, got["log_likelihood"], want["log_likelihood"],
          got["resampling_events"], want["resampling_events"])
    assert got["resampling_events"] == want["resampling_events"]
    assert [h["resampled"] for h in want["history"]] == [True, False, True, False]
    tol = 1e-10 * abs(want["log_likelihood"]) if precision == "fp64" else 1e-4
    assert abs(got["log_likelihood"] - want["log_likelihood"]) < max(tol, 1e-12)
    for a, b in zip(got["history"], want["history"]):
        assert a["resampled"] == b["resampled"]
        rel = np.linalg.norm(a["mean_field"] - b["mean_field"]) / np.linalg.norm(b["mean_field"])
        assert rel < (1e-12 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("P", [4096, 4100], ids=["sums_in_lds", "sums_in_memory"])
def test_pf_matches_oracle_at_thousands_of_particles(pkg, orc, P):
    """Up to 4096 particles per filter the resampling sums and the ancestor
    search stay in the normalising workgroup's LDS, above that in device
    memory: both give the oracle's resampling (one event, in the second of
    three groups) and log-likelihood to 1e-10 relative (FP64, 16^2)."""
    n, nu, dt, seed = 16, 2e-3, 2e-3, 78
    rows, _ = synthetic(orc, n, nu, dt, groups=3, every=4, nobs=24, truth_key=998, seed=4,
                        sd=1.0)
    base = orc.grf(n, 998)
    init = np.stack([base + 0.05 * orc.grf(n, orc.stream_seed(6, i)) for i in range(P)])
    want = oracle_bootstrap_filter(orc, init, rows, dt, seed, nu, sigma=0.2)
    with pkg.Ensemble(n, P, "fp64", nu=nu, sigma=0.2, amp=0.1, kf=4, seed=seed) as ens:
        ens.set_state(init)
        got = pkg.bootstrap_filter(ens, rows, dt, seed)
    print(P, got["log_likelihood"], want["log_likelihood"])
    assert [h["resampled"] for h in want["history"]] == [False, True, False]
    assert [h["resampled"] for h in got["history"]] == [False, True, False]
    assert abs(got["log_likelihood"] - want["log_likelihood"]) < 1e-10 * abs(
        want["log_likelihood"])
    for a, b in zip(got["history"], want["history"]):
        assert abs(a["ess"] - b["ess"]) < 1e-8 * b["ess"]
        rel = np.linalg.norm(a["mean_field"] - b["mean_field"]) / np.linalg.norm(b["mean_field"])
        assert rel < 1e-12


def test_pf_all_minus_infinity_gives_minus_infinity(pkg):
    """All -inf weights end the filter with log-likelihood -inf (src/filter.cpp:76-80)."""
    n, P = 16, 4
    with pkg.Ensemble(n, P, "fp64", sigma=0.0, amp=0.0, seed=1) as ens:
        ens.pf_begin(1, 4)
        rec, _ = ens.pf_assimilate(0.01, 1, [1], [2], [1e300], [1e-300], False)
        ll, _ = ens.pf_result()
    assert rec[0] == -np.inf and ll == -np.inf


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_joint_filter_over_two_grids_matches_oracle(pkg, orc, precision):
    """One filter over 8 particles at 32^2 and 4 at 64^2 (the uneven-size
    configuration in miniature): joint weights, cross-grid resampling."""
    nu, dt, seed = 2e-3, 2e-3, 91
    rows, _ = synthetic(orc, 32, nu, dt, groups=4, every=10, nobs=24, truth_key=999, seed=3,
                        sd=1.0)
    base = orc.grf(32, 999)
    coarse = np.stack([base + 0.05 * orc.grf(32, orc.stream_seed(5, i)) for i in range(8)])
    fine = np.stack([change_grid(base + 0.05 * orc.grf(32, orc.stream_seed(6, i)), 64)
                     for i in range(4)])
    want = oracle_joint_filter(orc, [coarse, fine], rows, dt, seed, nu, sigma=0.2)
    with pkg.Ensemble(32, 8, precision, nu=nu, sigma=0.2, amp=0.1, seed=1) as a, \
            pkg.Ensemble(64, 4, precision, nu=nu, sigma=0.2, amp=0.1, seed=2) as b:
        a.set_state(coarse)
        b.set_state(fine)
        got = pkg.joint_bootstrap_filter([a, b], rows, dt, seed)
        states = list(a.get_state()) + list(b.get_state())
    print(precision, got["log_likelihood"], want["log_likelihood"],
          [h["resampled"] for h in want["history"]])
    assert [h["resampled"] for h in got["history"]] == [h["resampled"] for h in want["history"]]
    assert any(h["resampled"] for h in want["history"])
    tol = 1e-10 * abs(want["log_likelihood"]) if precision == "fp64" else 1e-4
    assert abs(got["log_likelihood"] - want["log_likelihood"]) < max(tol, 1e-12)
    for g, w in zip(states, want["states"]):
        assert np.linalg.norm(g - w) / np.linalg.norm(w) < (1e-11 if precision == "fp64" else 1e-5)
