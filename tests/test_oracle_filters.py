"""CPU pins of the filter oracle helpers (tests/oracle_filters.py)."""
from __future__ import annotations

import numpy as np

import json
import os

from oracle_filters import (change_grid, full_from_half, gaspari_cohn, localized_scalar_update,
                            oracle_bootstrap_filter, oracle_joint_filter, oracle_pmcmc,
                            stochastic_log_prior, taper_field)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_gaspari_cohn_published_values():
    """Gaspari & Cohn (1999) eq. 4.10: 1 at 0, compact support 2c, the two
    branches meet at c (value 5/24) and at 2c (value 0), monotone decay."""
    c = 3.0
    assert gaspari_cohn(0.0, c) == 1.0
    assert gaspari_cohn(2 * c, c) == 0.0 and gaspari_cohn(2.5 * c, c) == 0.0
    x = 1.0
    a = 1 - 5 / 3 * x**2 + 5 / 8 * x**3 + 0.5 * x**4 - 0.25 * x**5
    b = 4 - 5 * x + 5 / 3 * x**2 + 5 / 8 * x**3 - 0.5 * x**4 + x**5 / 12 - 2 / (3 * x)
    assert abs(a - 5 / 24) < 1e-15 and abs(b - 5 / 24) < 1e-15
    assert abs(gaspari_cohn(c, c) - 5 / 24) < 1e-15
    d = np.linspace(0, 2 * c, 200)
    v = gaspari_cohn(d, c)
    assert np.all(np.diff(v) <= 1e-15) and abs(v[-1]) < 1e-12


def test_taper_is_periodic_distance():
    t = taper_field(16, 1, 2, 4.0)
    assert t[2, 1] == 1.0
    assert abs(t[2, 15] - gaspari_cohn(2.0, 4.0)) < 1e-15  # x distance wraps: |15 - 1| -> 2
    assert t[10, 9] == 0.0


def test_localized_update_without_taper_is_the_plain_update(orc):
    """With an infinite radius the taper is 1 everywhere and the localized
    update equals the pinned scalar update."""
    rng = np.random.default_rng(3)
    n, P = 16, 6
    fields = np.stack([orc.fft2d(rng.standard_normal((n, n))) * orc.active_mask(n)
                       for _ in range(P)])
    h = rng.standard_normal(P)
    eps = rng.standard_normal(P)
    got = localized_scalar_update(orc, fields, h, 0.3, 0.2, eps, 3, 4, 1e9)
    want = orc.enkf_scalar_update(fields.reshape(P, -1), h, 0.3, 0.2, eps).reshape(P, n, n)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)


def test_likelihood_adjoint_matches_reference_and_differences(orc, ref):
    """The restated discrete adjoint equals the reference's likelihood_adjoint
    (value and gradient) and the central differences of its own value."""
    from oracle_filters import oracle_likelihood_adjoint
    h, rows = ref.synthetic(16, 0.01, 30, 10, 0.03, 0.1, 0.0, 0.5, 2, 3, 0.2, 987)
    init = ref.initialize(16, "multimode")
    for q in ((np.log(0.03), np.log(0.1)), (np.log(0.02), np.log(0.15)), (-13.0, -1.0)):
        want = ref.likelihood_adjoint(h, q[0], q[1], 7)
        v, g = oracle_likelihood_adjoint(orc, init, rows, 0.01, q, 7)
        if not np.isfinite(want["value"]):
            assert not np.isfinite(v)
            continue
        assert abs(v - want["value"]) < 1e-12 * abs(want["value"])
        np.testing.assert_allclose(g, want["gradient"], rtol=1e-10, atol=0)
        e = 1e-5
        fd = [(oracle_likelihood_adjoint(orc, init, rows, 0.01, (q[0] + e, q[1]), 7)[0] -
               oracle_likelihood_adjoint(orc, init, rows, 0.01, (q[0] - e, q[1]), 7)[0]) / (2 * e),
              (oracle_likelihood_adjoint(orc, init, rows, 0.01, (q[0], q[1] + e), 7)[0] -
               oracle_likelihood_adjoint(orc, init, rows, 0.01, (q[0], q[1] - e), 7)[0]) / (2 * e)]
        np.testing.assert_allclose(g, fd, rtol=1e-6, atol=0)
    ref.dataset_free(h)


def test_step_tangent_and_likelihood_gradient_match_reference(orc, ref):
    """The restated step_tangent equals the reference's; the forward-
    sensitivity gradient equals the reference's likelihood_gradient and the
    adjoint gradient of the same discrete objective."""
    from oracle_filters import (oracle_likelihood_adjoint, oracle_likelihood_gradient,
                                oracle_step_tangent)
    n, nu, alpha, dt = 32, 0.01, 0.1, 5e-3
    w = orc.grf(n, 5)
    T = [orc.grf(n, 6), orc.grf(n, 7)]
    ws, Ts = w, T
    for _ in range(3):
        ws, Ts = oracle_step_tangent(orc, ws, Ts, nu, alpha, dt)
    rw, rT = ref.step_tangent(w, np.stack(T), nu, alpha, dt, 3)
    assert np.linalg.norm(ws - rw) < 1e-14 * np.linalg.norm(rw)
    for j in range(2):
        assert np.linalg.norm(Ts[j] - rT[j]) < 1e-13 * np.linalg.norm(rT[j])
    h, rows = ref.synthetic(16, 0.01, 30, 10, 0.03, 0.1, 0.0, 0.5, 2, 3, 0.2, 987)
    init = ref.initialize(16, "multimode")
    for q in ((np.log(0.03), np.log(0.1)), (np.log(0.02), np.log(0.15))):
        rv, rg = ref.likelihood_gradient(h, *q)
        v, g = oracle_likelihood_gradient(orc, init, rows, 0.01, q)
        assert abs(v - rv) < 1e-12 * abs(rv)
        np.testing.assert_allclose(g, rg, rtol=1e-10, atol=0)
        _, ga = oracle_likelihood_adjoint(orc, init, rows, 0.01, q, 7)
        np.testing.assert_allclose(g, ga, rtol=1e-9, atol=0)
    ref.dataset_free(h)


# ---------------------------------------------------------------- pMCMC --
def test_oracle_pmcmc_reproduces_the_reference_chain(orc):
    """With the reference's prior, streams and OU forcing, the oracle chain is
    the reference's own particle_metropolis(data, log(0.03, 0.1, 0.2, 0.5),
    0.1, 6, 8, 123) (golden ou_16, written by the unmodified reference):
    proposals, likelihood seeds and decisions identical, values to 1e-10."""
    g = np.load(os.path.join(GOLDEN, "ou_16.npz"))
    meta = json.load(open(os.path.join(GOLDEN, "ou_16.json")))
    n = meta["n"]
    q0 = np.log([meta["nu"], meta["alpha"], meta["sigma"], meta["tau"]])
    init = full_from_half(g["initial"], n)
    chain = oracle_pmcmc(orc, init, g["rows"], meta["dt"], q0, 6, 0.1, 0, None, None, 8,
                         kmin=2, kmax=3, forcing="ou", prior="stochastic", chain_seeds=[123])[0]
    for s, (q, ll, dens, acc, lseed) in enumerate(chain):
        np.testing.assert_allclose(q, g["chain_q"][s], rtol=1e-14, atol=0)
        assert abs(ll - g["chain_values"][s][0]) <= 1e-10 * abs(ll)
        assert abs(dens - g["chain_values"][s][1]) <= 1e-10 * abs(dens)
        assert acc == bool(g["chain_accepted"][s])
        assert lseed == int(g["chain_seeds"][s])


def test_oracle_pmcmc_rejection_keeps_the_estimate(orc):
    """Pseudo-marginal invariant (reference tests/verification.cpp:225-236):
    a rejected step keeps the previous state, likelihood estimate and seed;
    the chain both accepts and rejects."""
    g = np.load(os.path.join(GOLDEN, "ou_16.npz"))
    init = full_from_half(g["initial"], 16)
    q0 = np.log([0.03, 0.1, 0.2, 0.5])
    chain = oracle_pmcmc(orc, init, g["rows"], 0.01, q0, 10, 0.1, 0, None, None, 8, kmin=2,
                         kmax=3, forcing="ou", prior="stochastic", chain_seeds=[123])[0]
    decisions = [r[3] for r in chain]
    assert any(decisions) and not all(decisions[1:])
    for prev, cur in zip(chain, chain[1:]):
        if not cur[3]:
            np.testing.assert_array_equal(cur[0], prev[0])
            assert cur[1] == prev[1] and cur[4] == prev[4]


def test_stochastic_prior_closed_form():
    """The reference prior: -0.5 sum ((q - c) / 1.5)^2 inside the bounds,
    -inf on any violated bound or non-finite component."""
    c = np.log([0.03, 0.1, 0.2, 0.5])
    assert stochastic_log_prior(c) == 0.0
    q = c + np.array([0.3, -0.6, 0.9, 1.5])
    assert abs(stochastic_log_prior(q) + 0.5 * np.sum((q - c) ** 2) / 2.25) < 1e-15
    for i, bad in ((0, 0.01), (1, 0.01), (2, 2.01), (3, 2.01), (0, -10.01), (3, np.nan)):
        v = c.copy()
        v[i] = bad
        assert stochastic_log_prior(v) == -np.inf
    assert stochastic_log_prior([0.0, 0.0, 2.0, -10.0]) > -np.inf  # bounds are inclusive


# ---------------------------------------------------------- joint filter --
def test_change_grid_pad_then_truncate_is_identity(orc):
    """Zero padding to a finer grid then truncation back restores every
    retained mode, and the padded field adds no mode the coarse grid lacks."""
    rng = np.random.default_rng(11)
    for n, fine in ((16, 32), (32, 64), (32, 128)):
        f = orc.fft2d(rng.standard_normal((n, n))) * orc.active_mask(n)
        up = change_grid(f, fine)
        np.testing.assert_array_equal(change_grid(up, n), f)
        K = (n - 1) // 3
        k = np.rint(np.fft.fftfreq(fine, 1.0 / fine)).astype(int)
        outside = (np.abs(k)[None, :] > K) | (np.abs(k)[:, None] > K)
        assert np.all(up[outside] == 0)


def test_change_grid_padded_field_takes_the_coarse_values(orc):
    """The padded field is the same trigonometric polynomial: at fine grid
    point (2x, 2y) it equals the coarse field at (x, y); truncating a fine field
    keeps exactly the modes with |kx|, |ky| <= K_coarse."""
    rng = np.random.default_rng(12)
    n = 32
    f = orc.fft2d(rng.standard_normal((n, n))) * orc.active_mask(n)
    coarse = orc.fft2d(f, inverse=True).real
    fine = orc.fft2d(change_grid(f, 2 * n), inverse=True).real
    np.testing.assert_allclose(fine[::2, ::2], coarse, rtol=0, atol=1e-13)
    g = orc.fft2d(rng.standard_normal((64, 64))) * orc.active_mask(64)
    t = change_grid(g, n)
    K = (n - 1) // 3
    for ky in range(-K, K + 1):
        for kx in range(-K, K + 1):
            assert t[ky % n, kx % n] == (g[ky % 64, kx % 64] if (kx or ky) else 0)
    assert np.count_nonzero(t) == np.count_nonzero(orc.active_mask(n) * t)


def test_joint_filter_on_one_grid_is_the_bootstrap_filter(orc):
    """Two ensembles on the same grid: the joint filter is the pinned
    bootstrap filter over their union, bit for bit (log-likelihood, records,
    final states), including resampling groups."""
    n, dt, nu = 32, 2e-3, 1e-3
    a = np.stack([orc.grf(n, 100 + i) for i in range(5)])
    b = np.stack([orc.grf(n, 200 + i) for i in range(3)])
    rng = np.random.default_rng(5)
    truth = orc.grf(n, 7)
    rows = []
    for t in (4, 8, 12):
        truth = orc.trajectory(n, nu, 0.0, dt, 4, orc.stream_seed(9, 0), initial=truth,
                               step0=t - 4)
        ph = orc.fft2d(truth, inverse=True).real
        for _ in range(6):
            x, y = rng.integers(0, n, 2)
            rows.append((t, x, y, ph[y, x] + 0.3 * rng.standard_normal(), 0.3))
    rows = np.array(rows, dtype=np.float64)
    joint = oracle_joint_filter(orc, [a, b], rows, dt, 31, nu)
    single = oracle_bootstrap_filter(orc, np.concatenate([a, b]), rows, dt, 31, nu)
    assert joint["log_likelihood"] == single["log_likelihood"]
    assert joint["resampling_events"] == single["resampling_events"] >= 1
    for hj, hs in zip(joint["history"], single["history"]):
        for k in ("step", "increment", "ess", "max_weight", "resampled"):
            assert hj[k] == hs[k], k
