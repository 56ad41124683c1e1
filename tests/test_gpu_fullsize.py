"""The benchmark workloads' shapes checked on the GPU, not only timed.

* the EnKF workload: one batch (MFMA) analysis of 1024 FP32 members at 128^2 with
  64 observations against a dense FP64 closed form computed on the host from
  the downloaded forecast states and the same perturbations (the split-K
  products over the 7310-long state axis at their real size);
* the two-grid workload: the joint particle filter over 256^2 and 512^2 particles (FP32,
  the three-pass transforms at 512^2, the cross-grid gather on resampling)
  against the FP64 oracle's joint filter;
* the pMCMC workload: batched chains at 256^2 (FP32) against the oracle's chains.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle_filters import change_grid, oracle_joint_filter, oracle_pmcmc
from test_gpu_filter import synthetic

pytestmark = pytest.mark.gpu


def half_plane(fields):
    """[m][c][ky + K] retained half plane of reference Fields (m, n, n)."""
    n = fields.shape[-1]
    K = (n - 1) // 3
    ky = np.arange(-K, K + 1) % n
    return np.ascontiguousarray(fields[:, ky, :K + 1].transpose(0, 2, 1))


def observation_operator(n, xs, ys):
    """E[o, c, j]: physical value at (x_o, y_o) = Re sum E * w over the half
    plane (kx > 0 modes stand for their conjugates too)."""
    K = (n - 1) // 3
    c = np.arange(K + 1)[None, :, None]
    ky = np.arange(-K, K + 1)[None, None, :]
    phase = np.exp(2j * np.pi * (c * xs[:, None, None] + ky * ys[:, None, None]) / n)
    return np.where(c > 0, 2.0, 1.0) * phase


def test_batch_enkf_analysis_at_1024_members_128(pkg, orc):
    n, M, nobs, sd, seed, ens_seed = 128, 1024, 64, 0.01, 2027, 99
    rng = np.random.default_rng(17)
    xs, ys = rng.integers(0, n, nobs), rng.integers(0, n, nobs)
    with pkg.Ensemble(n, M, "fp32", nu=1e-3, sigma=0.05, amp=0.1, kf=4, seed=ens_seed) as ens:
        ens.init_grf()
        ens.step(1e-3, 50)
        ens.synchronize()
        forecast = half_plane(ens.get_state())
        E = observation_operator(n, xs, ys)
        HX = np.real(np.einsum("mcj,ocj->mo", forecast, E))  # M x m
        # observations: one member's values plus noise, so the innovation is O(sd)
        values = HX[0] + sd * rng.standard_normal(nobs)
        ens.enkf_begin(seed, "batch")
        rec, _ = ens.enkf_assimilate(1e-3, 0, xs, ys, values, np.full(nobs, sd))
        analysis = half_plane(ens.get_state())
    # the engine's perturbations: NormalStream(stream_seed(seed, M)), observation-major
    stream = orc.mt(orc.stream_seed(seed, M))
    eps = np.array([[orc.normal(stream) for _ in range(M)] for _ in range(nobs)])
    X = forecast.reshape(M, -1)
    A = X - X.mean(0)
    HA = HX - HX.mean(0)
    S = HA.T @ HA / (M - 1) + np.diag(np.full(nobs, sd * sd))
    D = values[:, None] + sd * eps - HX.T  # m x M
    C = np.linalg.solve(S, D).T @ HA.T / (M - 1)  # M x M
    want = X + C @ A
    got = analysis.reshape(M, -1)
    incr_err = np.linalg.norm((got - X) - (want - X)) / np.linalg.norm(want - X)
    mean_err = np.linalg.norm(got.mean(0) - want.mean(0)) / np.linalg.norm(want.mean(0))
    print(f"batch EnKF 1024x128^2: increment rel err {incr_err:.2e}, mean rel err {mean_err:.2e}")
    assert incr_err < 1e-3 and mean_err < 1e-4
    # the analysis pulls the mean towards the observations
    fc_rmse = np.sqrt(np.mean((HX.mean(0) - values) ** 2))
    an_rmse = np.sqrt(np.mean((np.real(np.einsum("cj,ocj->o", want.reshape(M, *forecast.shape[1:])
                                                 .mean(0), E)) - values) ** 2))
    assert abs(rec[0] - fc_rmse) <= 1e-4 * fc_rmse and an_rmse < fc_rmse


def test_joint_filter_256_and_512(pkg, orc):
    """8 particles at 256^2 and 4 at 512^2, nu = dt = 5e-4, two groups of 3
    steps with 48 observations (sd 0.05); the first group resamples, so
    children cross grids."""
    nu, dt, seed = 5e-4, 5e-4, 4242
    rows, _ = synthetic(orc, 256, nu, dt, groups=2, every=3, nobs=48, truth_key=71, seed=12,
                        sd=0.05)
    base = orc.grf(256, 71)
    coarse = np.stack([base + 0.2 * orc.grf(256, orc.stream_seed(13, i)) for i in range(8)])
    fine = np.stack([change_grid(base + 0.2 * orc.grf(256, orc.stream_seed(14, i)), 512)
                     for i in range(4)])
    want = oracle_joint_filter(orc, [coarse, fine], rows, dt, seed, nu, sigma=0.05)
    with pkg.Ensemble(256, 8, "fp32", nu=nu, sigma=0.05, amp=0.1, seed=1) as a, \
            pkg.Ensemble(512, 4, "fp32", nu=nu, sigma=0.05, amp=0.1, seed=2) as b:
        a.set_state(coarse)
        b.set_state(fine)
        got = pkg.joint_bootstrap_filter([a, b], rows, dt, seed)
        states = list(a.get_state()) + list(b.get_state())
    flags = [h["resampled"] for h in want["history"]]
    print("joint 256+512:", got["log_likelihood"], want["log_likelihood"], flags)
    assert [h["resampled"] for h in got["history"]] == flags and flags[0]
    for h, w in zip(got["history"], want["history"]):
        assert abs(h["increment"] - w["increment"]) <= 1e-4 * abs(w["increment"]) + 1e-3
        assert abs(h["ess"] - w["ess"]) <= 2e-3 * w["ess"]
    for g, w in zip(states, want["states"]):
        assert np.linalg.norm(g - w) / np.linalg.norm(w) < 1e-5


def test_pmcmc_at_256(pkg, orc):
    """2 chains x 8 particles x 2 Metropolis steps at 256^2 (FP32) over a
    two-group window: the oracle's proposals, decisions and estimates."""
    n, P, C, S, dt, seed = 256, 8, 2, 2, 5e-4, 606
    rows, _ = synthetic(orc, n, 5e-4, dt, groups=2, every=4, nobs=32, truth_key=72, seed=15,
                        sd=0.05)
    init = orc.grf(n, 72)
    q0 = np.array([[np.log(6e-4), np.log(0.1)], [np.log(1.5e-3), np.log(0.08)]])
    lower, upper = (np.log(2e-4), np.log(0.02)), (np.log(5e-3), np.log(0.5))
    want = oracle_pmcmc(orc, init, rows, dt, q0, S, 0.3, seed, lower, upper, P, sigma=0.05)
    with pkg.Ensemble(n, P * C, "fp32", nu=5e-4, sigma=0.05, amp=0.1, kf=4, seed=1) as ens:
        got = pkg.particle_metropolis(ens, rows, dt, init, q0, S, 0.3, seed, friction=0.0,
                                      forcing_sigma=0.05)
    for c in range(C):
        for s in range(S):
            q, ll, dens, acc = want[c][s][:4]
            print("chain", c, "sample", s, got["log_likelihood"][c, s], ll, acc)
            assert got["accepted"][c, s] == acc, (c, s)
            np.testing.assert_array_equal(got["q"][c, s], q)
            assert abs(got["log_likelihood"][c, s] - ll) <= 1e-4 * abs(ll) + 1e-3
