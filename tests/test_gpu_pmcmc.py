"""GPU batched particle-marginal Metropolis-Hastings (csrc/pmcmc.hip) against
the FP64 oracle replay of the reference's particle_metropolis
(tests/oracle_filters.py), chain by chain: identical proposals, acceptance
decisions and retained states, likelihood estimates to 1e-9 relative; the
batched likelihood evaluation against the oracle's filters, and with a
surrogate forecast against the same filter composed by hand from the network
forecast and the filter steps.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle_filters import oracle_bootstrap_filter, oracle_pmcmc
from test_gpu_filter import synthetic

pytestmark = pytest.mark.gpu


def test_batched_pmcmc_matches_oracle_chains(pkg, orc):
    n, P, C, S, dt, seed = 32, 8, 3, 4, 2e-3, 2026
    rows, _ = synthetic(orc, n, 1e-3, dt, groups=3, every=10, nobs=16, truth_key=555, seed=4,
                        sd=0.5)
    init = orc.grf(n, 555)
    q0 = np.array([[np.log(1e-3), np.log(0.1)], [np.log(2e-3), np.log(0.05)],
                   [np.log(6e-4), np.log(0.2)]])
    lower, upper = (np.log(2e-4), np.log(0.02)), (np.log(5e-3), np.log(0.5))
    want = oracle_pmcmc(orc, init, rows, dt, q0, S, 0.4, seed, lower, upper, P, sigma=0.1)
    with pkg.Ensemble(n, P * C, "fp64", nu=1e-3, sigma=0.1, amp=0.1, kf=4, seed=1) as ens:
        got = pkg.particle_metropolis(ens, rows, dt, init, q0, S, 0.4, seed, friction=0.0,
                                      forcing_sigma=0.1)
    accepted = 0
    for c in range(C):
        for s in range(S):
            q, ll, dens, acc = want[c][s][:4]
            assert got["accepted"][c, s] == acc, (c, s)
            np.testing.assert_array_equal(got["q"][c, s], q)
            assert abs(got["log_likelihood"][c, s] - ll) <= 1e-9 * abs(ll)
            assert abs(got["log_density"][c, s] - dens) <= 1e-9 * abs(dens)
            accepted += acc
    # the window is informative: some proposals are accepted, some rejected
    assert 0 < accepted < C * S


def test_batched_pmcmc_with_ou_forcing_matches_oracle_chains(pkg, orc):
    """The same chains with the reference's OU forcing: every likelihood
    evaluation starts new particles (seeds stream_seed(lseed, i), stationary OU
    start), as the reference's particle_log_likelihood builds new Solvers."""
    n, P, C, S, dt, seed = 32, 8, 2, 3, 2e-3, 77
    rows, _ = synthetic(orc, n, 1e-3, dt, groups=3, every=10, nobs=16, truth_key=556, seed=5,
                        sd=0.5)
    init = orc.grf(n, 556)
    q0 = np.array([[np.log(1e-3), np.log(0.1)], [np.log(2e-3), np.log(0.05)]])
    lower, upper = (np.log(2e-4), np.log(0.02)), (np.log(5e-3), np.log(0.5))
    want = oracle_pmcmc(orc, init, rows, dt, q0, S, 0.4, seed, lower, upper, P, sigma=0.5,
                        forcing="ou", tau=0.3)
    with pkg.Ensemble(n, P * C, "fp64", nu=1e-3, sigma=0.5, tau=0.3, amp=0.1, kf=4,
                      seed=1) as ens:
        ens.set_forcing_process("ou")
        got = pkg.particle_metropolis(ens, rows, dt, init, q0, S, 0.4, seed, friction=0.0,
                                      forcing_sigma=0.5)
    for c in range(C):
        for s in range(S):
            q, ll, dens, acc = want[c][s][:4]
            assert got["accepted"][c, s] == acc, (c, s)
            np.testing.assert_array_equal(got["q"][c, s], q)
            assert abs(got["log_likelihood"][c, s] - ll) <= 1e-9 * abs(ll)


def test_pmcmc_rejects_bad_configuration(pkg, orc):
    n = 32
    init = orc.grf(n, 1)
    rows = np.array([[10, 1, 2, 0.1, 0.5]])
    with pkg.Ensemble(n, 8, "fp64", seed=1) as ens:
        with pytest.raises(pkg.HzlInvalidArgument):  # 8 members do not split into 3 chains
            pkg.particle_metropolis(ens, rows, 2e-3, init, np.zeros((3, 2)), 2, 0.1, 1)
        with pytest.raises(pkg.HzlInvalidArgument):  # outside the prior
            pkg.particle_metropolis(ens, rows, 2e-3, init, [[0.0, 0.0], [0.0, 0.0]], 2, 0.1, 1)


def test_batched_likelihoods_match_oracle_filters(pkg, orc):
    """hzl_pf_log_likelihoods: four filters of 8 FP64 particles at their own
    (nu, amplitude) and seeds, each the oracle's bootstrap filter from the
    same initial field; a point outside the prior box gives -inf."""
    n, P, dt = 32, 8, 2e-3
    rows, _ = synthetic(orc, n, 1e-3, dt, groups=3, every=10, nobs=16, truth_key=557, seed=6,
                        sd=0.5)
    init = orc.grf(n, 557)
    q = np.log([[1e-3, 0.1], [2e-3, 0.05], [6e-4, 0.3], [1e-2, 0.1]])  # last: nu outside
    seeds = [11, 2 ** 63 + 5, 77, 3]
    with pkg.Ensemble(n, 4 * P, "fp64", nu=1e-3, sigma=0.1, amp=0.1, kf=4, seed=1) as ens:
        got = pkg.pf_log_likelihoods(ens, rows, dt, init, q, seeds, forcing_sigma=0.1)
    assert got[3] == -np.inf
    for f in range(3):
        want = oracle_bootstrap_filter(orc, np.stack([init] * P), rows, dt, seeds[f],
                                       np.exp(q[f, 0]), amp=np.exp(q[f, 1]),
                                       sigma=0.1)["log_likelihood"]
        print(f, got[f], want)
        assert abs(got[f] - want) <= 1e-9 * abs(want)


def test_surrogate_likelihoods_are_the_network_forecast_filter(pkg, orc):
    """With a surrogate the batched likelihoods (and so the pMCMC chains)
    forecast every interval with the network: the same numbers as filters
    composed by hand from Surrogate.forecast and pf_assimilate with no solver
    steps, for the same members, parameters and seeds."""
    n, P, F, dt, H = 64, 16, 3, 2e-3, 25
    rows, _ = synthetic(orc, n, 1e-3, dt, groups=3, every=H, nobs=16, truth_key=558, seed=7,
                        sd=0.3)
    init = orc.grf(n, 558)
    q = np.log([[1e-3, 0.1], [2e-3, 0.05], [6e-4, 0.3]])
    seeds = [5, 6, 7]
    with pkg.Surrogate(n, 32, 16, 2, 4, seed=9) as sur, \
            pkg.Ensemble(n, F * P, "fp32", nu=1e-3, sigma=0.05, amp=0.1, seed=1) as ens:
        sur.set_scales(1.5, 4.0)
        got = pkg.pf_log_likelihoods(ens, rows, dt, init, q, seeds, surrogate=sur,
                                     surrogate_precision="fp32")
        ens.fill_state(init)
        ens.step_index = 0
        for f in range(F):
            for i in range(P):
                ens.set_member_parameters(f * P + i, np.exp(q[f, 0]), 0.0, 0.05,
                                          np.exp(q[f, 1]))
        groups = np.unique(rows[:, 0].astype(int))
        ens.pf_begin_filters(seeds, max_groups=len(groups) + 1)
        now = 0
        for g, t in enumerate(groups):
            sel = rows[rows[:, 0] == t]
            sur.forecast(ens, dt, int(t - now), "fp32")
            ens.pf_assimilate(dt, 0, sel[:, 1], sel[:, 2], sel[:, 3], sel[:, 4],
                              g == len(groups) - 1)
            now = int(t)
        want, _ = ens.pf_result()
        solver = pkg.pf_log_likelihoods(ens, rows, dt, init, q, seeds)
    print(got, want, solver)
    np.testing.assert_array_equal(got, want)
    assert np.all(np.isfinite(got)) and not np.array_equal(got, solver)
