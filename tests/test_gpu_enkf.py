"""GPU ensemble Kalman filter parity (csrc/enkf.hip), through the C-ABI.

  * the reference's own ensemble_kalman_filter in the deterministic limit
    (tests/verification.cpp:204-210) from the committed golden vectors:
    forecast sensor RMSE, zero spread and mean fields per group;
  * the FP64 oracle replay (tests/oracle_filters.py) with stochastic forcing,
    spread, inflation and the reference's perturbation stream: serial mode to
    1e-9 in FP64 and 1e-4 in FP32; batch mode (MFMA products, FP64 Cholesky)
    against the dense closed-form analysis with the same perturbations.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle_filters import oracle_enkf
from test_gpu_filter import full_from_half, synthetic

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_enkf_deterministic_limit_matches_reference_golden(pkg):
    g = np.load(os.path.join(GOLDEN, "pf_deterministic_16.npz"))
    meta = json.load(open(os.path.join(GOLDEN, "pf_deterministic_16.json")))
    n, P = 16, meta["particles"]
    w0 = full_from_half(g["initial"], n)
    with pkg.Ensemble(n, P, "fp64", nu=meta["nu"], alpha=meta["alpha"], sigma=0.0, amp=0.0,
                      kmin=2, kmax=3, seed=meta["seed"]) as ens:
        ens.set_state(np.stack([w0] * P))
        hist = pkg.ensemble_kalman_filter(ens, g["rows"], meta["dt"], meta["seed"])
    assert len(hist) == len(g["enkf_records"])
    for h, rec, mean in zip(hist, g["enkf_records"], g["enkf_means"]):
        assert h["step"] == int(rec[0])
        assert abs(h["forecast_sensor_rmse"] - rec[1]) < 1e-10
        assert abs(h["posterior_spread"] - rec[2]) < 1e-12
        phys = (np.fft.ifft2(h["mean_field"]) * n * n).real
        np.testing.assert_allclose(phys, mean, rtol=0, atol=1e-10)


def setup(orc, n=32, P=16, nu=2e-3, dt=2e-3):
    rows, _ = synthetic(orc, n, nu, dt, groups=3, every=10, nobs=8, truth_key=4242, seed=8,
                        sd=0.3)
    base = orc.grf(n, 4242)
    init = np.stack([base + 0.1 * orc.grf(n, orc.stream_seed(9, i)) for i in range(P)])
    return rows, init


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-4)])
def test_enkf_serial_matches_oracle(pkg, orc, precision, tol):
    n, P, nu, dt, ens_seed, seed = 32, 16, 2e-3, 2e-3, 31, 77
    rows, init = setup(orc, n, P, nu, dt)
    want = oracle_enkf(orc, init, rows, dt, ens_seed, seed, nu, sigma=0.2, amp=0.1,
                       inflation=1.05)
    with pkg.Ensemble(n, P, precision, nu=nu, sigma=0.2, amp=0.1, kf=4, seed=ens_seed) as ens:
        ens.set_state(init)
        got = pkg.ensemble_kalman_filter(ens, rows, dt, seed, inflation=1.05)
        final = ens.get_state()
    for a, b in zip(got, want):
        print(precision, a["forecast_sensor_rmse"], b["forecast_sensor_rmse"],
              a["posterior_spread"], b["posterior_spread"])
        assert a["step"] == b["step"]
        assert abs(a["forecast_sensor_rmse"] - b["forecast_sensor_rmse"]) <= \
            tol * b["forecast_sensor_rmse"]
        assert abs(a["posterior_spread"] - b["posterior_spread"]) <= tol * b["posterior_spread"]
        assert rel(a["mean_field"], b["mean_field"]) < tol
    assert rel(final, want[-1]["states"]) < tol
    # the analysis moved the members (a non-trivial test)
    assert want[-1]["posterior_spread"] > 0


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-4)])
def test_enkf_localized_matches_oracle(pkg, orc, precision, tol):
    """Gaspari-Cohn tapered covariances (src/filter.cpp:290-308), radius 5 cells."""
    n, P, nu, dt, ens_seed, seed = 32, 16, 2e-3, 2e-3, 31, 77
    rows, init = setup(orc, n, P, nu, dt)
    want = oracle_enkf(orc, init, rows, dt, ens_seed, seed, nu, sigma=0.2, amp=0.1, radius=5.0)
    plain = oracle_enkf(orc, init, rows, dt, ens_seed, seed, nu, sigma=0.2, amp=0.1)
    with pkg.Ensemble(n, P, precision, nu=nu, sigma=0.2, amp=0.1, kf=4, seed=ens_seed) as ens:
        ens.set_state(init)
        got = pkg.ensemble_kalman_filter(ens, rows, dt, seed, localization_radius=5.0)
        final = ens.get_state()
    for a, b in zip(got, want):
        assert abs(a["posterior_spread"] - b["posterior_spread"]) <= tol * b["posterior_spread"]
        assert rel(a["mean_field"], b["mean_field"]) < tol
    assert rel(final, want[-1]["states"]) < tol
    # the taper changed the analysis
    assert rel(want[-1]["mean_field"], plain[-1]["mean_field"]) > 1e-3


def test_enkf_batch_matches_dense_closed_form(pkg, orc):
    n, P, nu, dt, ens_seed, seed = 32, 16, 2e-3, 2e-3, 31, 77
    rows, init = setup(orc, n, P, nu, dt)
    want = oracle_enkf(orc, init, rows, dt, ens_seed, seed, nu, sigma=0.2, amp=0.1,
                       mode="batch")
    with pkg.Ensemble(n, P, "fp32", nu=nu, sigma=0.2, amp=0.1, kf=4, seed=ens_seed) as ens:
        ens.set_state(init)
        got = pkg.ensemble_kalman_filter(ens, rows, dt, seed, mode="batch")
        final = ens.get_state()
    for a, b in zip(got, want):
        print(a["posterior_spread"], b["posterior_spread"], rel(a["mean_field"], b["mean_field"]))
        assert abs(a["forecast_sensor_rmse"] - b["forecast_sensor_rmse"]) <= \
            1e-4 * b["forecast_sensor_rmse"]
        assert abs(a["posterior_spread"] - b["posterior_spread"]) <= 1e-3 * b["posterior_spread"]
        assert rel(a["mean_field"], b["mean_field"]) < 1e-4
    assert rel(final, want[-1]["states"]) < 1e-3


def test_enkf_free_forecast_records_and_errors(pkg, orc):
    n, P = 32, 8
    rows, init = setup(orc, n, P)
    with pkg.Ensemble(n, P, "fp64", nu=2e-3, sigma=0.2, seed=3) as ens:
        ens.set_state(init[:P])
        hist = pkg.ensemble_kalman_filter(ens, rows[:8], 2e-3, 5, forecast_steps=5,
                                          forecast_every=2)
        assert [h["step"] for h in hist] == [10, 12, 14, 15]
        assert all(h["is_forecast"] for h in hist[1:]) and np.isnan(hist[1]["forecast_sensor_rmse"])
        sp, mean = ens.mean_spread()
        assert abs(sp - hist[-1]["posterior_spread"]) < 1e-12 * sp
        with pytest.raises(pkg.HzlInvalidArgument):
            ens.enkf_assimilate(2e-3, 0, [1], [1], [0.0], [0.1], inflation=0.5)
        with pytest.raises(pkg.HzlInvalidArgument):
            ens.enkf_begin(1, "batch")  # FP64 ensemble
