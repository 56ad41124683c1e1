"""The reference's OU shell forcing on the GPU (HZL_FORCING_OU), through the C-ABI.

Checks:
  * the reference's own Solver with sigma > 0 (golden ou_16, written by
    oracle/_ref): final state and forcing after 25 and 40 steps, FP64 to 1e-12,
    FP32 to 1e-5;
  * the reference's own bootstrap_filter over a stochastic dataset (golden
    ou_16): log-likelihood, resampling events, per-group records and means;
  * the FP64 oracle filter with OU forcing (tests/oracle_filters.py) on a
    32^2 case that resamples, with Kolmogorov forcing on top;
  * graph replay of OU steps and filter cycles bit-identical to eager launches;
  * the reference's own ensemble_kalman_filter with sigma > 0 (golden ou_16):
    the OU forcing analysed as a latent state, plain and with inflation and
    localisation;
  * errors: forcing of a white-noise ensemble, batch EnKF with OU forcing,
    tau <= 0.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle_filters import oracle_bootstrap_filter, oracle_enkf

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


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def golden():
    return (np.load(os.path.join(GOLDEN, "ou_16.npz")),
            json.load(open(os.path.join(GOLDEN, "ou_16.json"))))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_ou_trajectory_matches_reference_golden(pkg, orc, precision, tol):
    g, meta = golden()
    n = meta["n"]
    w0 = full_from_half(g["initial"], n)
    for run in meta["runs"]:
        key = f"run_{run['seed']}"
        print(precision, key, flush=True)
        # member 1 carries the reference solver's seed; members 0 and 2 others
        seeds = [orc.stream_seed(5, 0), run["seed"], orc.stream_seed(5, 2)]
        with pkg.Ensemble(n, 3, precision, nu=meta["nu"], alpha=meta["alpha"],
                          sigma=meta["sigma"], tau=meta["tau"], kmin=run["kmin"],
                          kmax=run["kmax"], amp=0.0, seed=0) as ens:
            ens.set_state(np.stack([w0] * 3))
            ens.set_member_seeds(seeds)
            ens.set_forcing_process("ou")
            ens.step(meta["dt"], run["steps"])
            w = ens.get_state(1)
            F = ens.get_forcing(1)
            others = ens.get_state()
        want_w = full_from_half(g[key + "_state"], n)
        want_F = full_from_half(g[key + "_forcing"], n)
        assert rel(w, want_w) < tol
        assert rel(F, want_F) < (1e-13 if precision == "fp64" else 1e-12)
        # the other members follow their own streams (oracle)
        for m in (0, 2):
            ow, _ = orc.run_ou(w0, meta["nu"], meta["alpha"], meta["sigma"], meta["tau"],
                               run["kmin"], run["kmax"], seeds[m], meta["dt"], run["steps"])
            assert rel(others[m], ow) < tol


def test_ou_filter_matches_reference_golden(pkg):
    g, meta = golden()
    n, P = meta["n"], meta["particles"]
    w0 = full_from_half(g["initial"], n)
    with pkg.Ensemble(n, P, "fp64", nu=meta["nu"], alpha=meta["alpha"], sigma=meta["sigma"],
                      tau=meta["tau"], kmin=2, kmax=3, amp=0.0, seed=meta["seed"]) as ens:
        ens.set_state(np.stack([w0] * P))
        ens.set_forcing_process("ou")
        res = pkg.bootstrap_filter(ens, g["rows"], meta["dt"], meta["seed"])
    print(res["log_likelihood"], meta["log_likelihood"], res["resampling_events"])
    assert res["resampling_events"] == meta["resampling_events"]
    assert abs(res["log_likelihood"] - meta["log_likelihood"]) < 1e-9 * abs(meta["log_likelihood"])
    for h, rec, mean in zip(res["history"], g["records"], g["means"]):
        assert abs(h["increment"] - rec[0]) < 1e-9
        assert abs(h["ess"] - rec[1]) < 1e-8 and abs(h["max_weight"] - rec[2]) < 1e-10
        phys = (np.fft.ifft2(h["mean_field"]) * n * n).real
        np.testing.assert_allclose(phys, mean, rtol=0, atol=1e-10)


@pytest.mark.parametrize("key,inflation,radius", [("enkf", 1.0, 0.0),
                                                  ("enkf_loc", 1.05, 3.0)])
def test_ou_enkf_matches_reference_golden(pkg, key, inflation, radius):
    """Serial EnKF with the OU forcing as a latent state against the reference's
    ensemble_kalman_filter with sigma 0.2 (golden ou_16)."""
    g, meta = golden()
    n, P = meta["n"], meta["particles"]
    w0 = full_from_half(g["initial"], n)
    with pkg.Ensemble(n, P, "fp64", nu=meta["nu"], alpha=meta["alpha"], sigma=meta["sigma"],
                      tau=meta["tau"], kmin=2, kmax=3, amp=0.0, seed=meta["seed"]) as ens:
        ens.set_state(np.stack([w0] * P))
        ens.set_forcing_process("ou")
        hist = pkg.ensemble_kalman_filter(ens, g["rows"], meta["dt"], meta["seed"],
                                          inflation=inflation, localization_radius=radius)
    assert len(hist) == len(g[key + "_records"])
    for h, rec, mean in zip(hist, g[key + "_records"], g[key + "_means"]):
        assert h["step"] == rec[0]
        assert abs(h["forecast_sensor_rmse"] - rec[1]) < 1e-10
        assert abs(h["posterior_spread"] - rec[2]) < 1e-10
        phys = (np.fft.ifft2(h["mean_field"]) * n * n).real
        np.testing.assert_allclose(phys, mean, rtol=0, atol=1e-10)


def test_ou_enkf_fp32_matches_oracle(pkg, orc):
    """FP32 states with the FP64 OU forcing state: serial EnKF with inflation and
    localisation against the FP64 oracle (OU latent forcing)."""
    n, P, nu, dt, seed, tau = 32, 16, 2e-3, 2e-3, 21, 0.3
    rows = synthetic(orc, n, nu, dt, groups=3, every=10, nobs=12, truth_key=41, seed=5, sd=0.3)
    base = orc.grf(n, 41)
    init = np.stack([base + 0.05 * orc.grf(n, orc.stream_seed(6, i)) for i in range(P)])
    want = oracle_enkf(orc, init, rows, dt, seed, seed, nu, sigma=1.0, forcing="ou", tau=tau,
                       inflation=1.05, radius=4.0)
    with pkg.Ensemble(n, P, "fp32", nu=nu, sigma=1.0, tau=tau, amp=0.1, kf=4, seed=seed) as ens:
        ens.set_state(init)
        ens.set_forcing_process("ou")
        got = pkg.ensemble_kalman_filter(ens, rows, dt, seed, inflation=1.05,
                                         localization_radius=4.0)
    for a, b in zip(got, want):
        assert abs(a["forecast_sensor_rmse"] - b["forecast_sensor_rmse"]) < 1e-5 * max(
            1.0, b["forecast_sensor_rmse"])
        assert abs(a["posterior_spread"] - b["posterior_spread"]) < 1e-5 * b["posterior_spread"]
        assert rel(a["mean_field"], b["mean_field"]) < 1e-5


def synthetic(orc, n, nu, dt, groups, every, nobs, truth_key, seed, sd):
    rng = np.random.default_rng(seed)
    w = orc.grf(n, truth_key)
    rows, step = [], 0
    for _ in range(groups):
        w = orc.trajectory(n, nu, 0.0, dt, every, truth_key, step0=step, initial=w)
        step += every
        ph = (orc.fft2d(w, inverse=True)).real
        xs, ys = rng.integers(0, n, nobs), rng.integers(0, n, nobs)
        vals = ph[ys, xs] + sd * rng.standard_normal(nobs)
        rows += [(step, x, y, v, sd) for x, y, v in zip(xs, ys, vals)]
    return np.array(rows, dtype=np.float64)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ou_filter_matches_oracle_with_resampling(pkg, orc, precision):
    n, P, nu, dt, seed, tau = 32, 16, 2e-3, 2e-3, 77, 0.3
    # sigma 3 spreads the particles enough that groups 1 and 2 resample
    rows = synthetic(orc, n, nu, dt, groups=4, every=10, nobs=24, truth_key=999, seed=3, sd=0.5)
    base = orc.grf(n, 999)
    init = np.stack([base + 0.05 * orc.grf(n, orc.stream_seed(5, i)) for i in range(P)])
    want = oracle_bootstrap_filter(orc, init, rows, dt, seed, nu, sigma=3.0, forcing="ou",
                                   tau=tau)
    with pkg.Ensemble(n, P, precision, nu=nu, sigma=3.0, tau=tau, amp=0.1, kf=4,
                      seed=seed) as ens:
        ens.set_state(init)
        ens.set_forcing_process("ou")
        got = pkg.bootstrap_filter(ens, rows, dt, seed)
    print(precision, got["log_likelihood"], want["log_likelihood"],
          got["resampling_events"], [h["resampled"] for h in want["history"]])
    assert [h["resampled"] for h in want["history"]] == [True, True, False, False]
    assert got["resampling_events"] == want["resampling_events"]
    tol = 1e-10 * abs(want["log_likelihood"]) if precision == "fp64" else 1e-4
    assert abs(got["log_likelihood"] - want["log_likelihood"]) < max(tol, 1e-12)
    for a, b in zip(got["history"], want["history"]):
        assert a["resampled"] == b["resampled"]
        assert rel(a["mean_field"], b["mean_field"]) < (1e-12 if precision == "fp64" else 1e-5)


def test_ou_graph_replay_is_bit_identical(pkg, orc):
    n, P, dt = 32, 8, 2e-3
    rows = synthetic(orc, n, 2e-3, dt, groups=3, every=6, nobs=16, truth_key=5, seed=9, sd=0.5)
    init = np.stack([orc.grf(n, orc.stream_seed(8, i)) for i in range(P)])
    out = []
    for graph in (False, True):
        with pkg.Ensemble(n, P, "fp32", nu=2e-3, sigma=0.3, tau=0.2, seed=11) as ens:
            ens.set_graph(graph)
            ens.set_state(init)
            ens.set_forcing_process("ou")
            ens.step(dt, 7)
            ens.step(dt, 7)  # second call replays the captured chunk
            res = pkg.bootstrap_filter(ens, rows, dt, seed=4)
            out.append((ens.get_state(), ens.get_forcing(), res["log_likelihood"]))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_ou_errors(pkg):
    with pkg.Ensemble(16, 2, "fp64", sigma=0.1, tau=0.5) as ens:
        with pytest.raises(pkg.HzlInvalidArgument):
            ens.get_forcing(0)  # white noise selected
        with pytest.raises(pkg.HzlInvalidArgument):
            ens.set_forcing_process("pink")
        ens.set_forcing_process("ou")
        with pytest.raises(pkg.HzlInvalidArgument):  # batch needs white forcing (and FP32)
            pkg.ensemble_kalman_filter(ens, np.array([[2, 1, 1, 0.0, 0.1]]), 0.01, 3,
                                       mode="batch")
        ens.set_forcing_process("white")
        with pytest.raises(pkg.HzlInvalidArgument):
            ens.get_forcing(0)
    with pkg.Ensemble(16, 4, "fp64", sigma=0.1) as a, pkg.Ensemble(32, 2, "fp64", sigma=0.1) as b:
        a.set_forcing_process("ou")
        with pytest.raises(pkg.HzlInvalidArgument):  # joint filters run white forcing
            pkg.joint_bootstrap_filter([a, b], np.array([[2, 1, 1, 0.0, 0.1]]), 0.01, 3)
    with pytest.raises(pkg.HzlInvalidArgument):  # tau must be positive at creation
        pkg.Ensemble(16, 2, "fp64", sigma=0.1, tau=0.0)
