"""Batched GPU likelihood_adjoint (src/adjoint.cpp:43-161): every member
evaluates the deterministic negative log-likelihood and its gradient with
respect to (log nu, log alpha) at its own parameter vector.

Checked against the reference's own likelihood_adjoint values (golden
fixture tests/golden/adjoint_16.json, written from the reference library) and
against the oracle restatement (tests/oracle_filters.py, pinned to the
reference and to central differences in tests/test_oracle_filters.py)."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle_filters import oracle_likelihood_adjoint

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


@pytest.mark.parametrize("precision,vtol,gtol", [("fp64", 1e-11, 1e-9), ("fp32", 1e-5, 2e-3)])
def test_matches_reference_golden(pkg, precision, vtol, gtol):
    meta = json.load(open(os.path.join(GOLDEN, "adjoint_16.json")))
    g = np.load(os.path.join(GOLDEN, "pf_deterministic_16.npz"))
    init = full_from_half(g["initial"], 16)
    rows = g["rows"]
    for stride in (7, 4):
        evals = [e for e in meta["evaluations"] if e["stride"] == stride]
        q = np.array([e["q"] for e in evals])
        print(precision, "stride", stride, flush=True)
        with pkg.Ensemble(16, len(evals), precision, seed=1) as ens:
            v, grad, counts = pkg.likelihood_adjoint(ens, init, rows, meta["dt"], q, stride)
        for k, e in enumerate(evals):
            if not np.isfinite(e["value"]):
                assert not np.isfinite(v[k]) and np.all(grad[k] == 0)
                continue
            assert abs(v[k] - e["value"]) < vtol * abs(e["value"])
            np.testing.assert_allclose(grad[k], e["gradient"], rtol=gtol, atol=0)
        assert counts[0] == evals[0]["checkpoint_fields"]
        assert counts[2] == evals[0]["replayed_steps"]


def test_matches_oracle_at_32(pkg, orc):
    n, dt, stride = 32, 5e-3, 6
    rng = np.random.default_rng(8)
    init = orc.grf(n, 77)
    rows = []
    for step in (0, 9, 17, 24):
        xs, ys = rng.integers(0, n, 12), rng.integers(0, n, 12)
        rows += [(step, x, y, rng.normal(0, 0.3), 0.1) for x, y in zip(xs, ys)]
    rows = np.array(rows, dtype=np.float64)
    q = np.array([[np.log(0.01), np.log(0.1)], [np.log(0.004), np.log(0.03)],
                  [np.log(0.02), np.log(0.2)]])
    with pkg.Ensemble(n, len(q), "fp64", seed=1) as ens:
        v, grad, _ = pkg.likelihood_adjoint(ens, init, rows, dt, q, stride)
    for k in range(len(q)):
        wv, wg = oracle_likelihood_adjoint(orc, init, rows, dt, q[k], stride)
        assert abs(v[k] - wv) < 1e-11 * abs(wv)
        np.testing.assert_allclose(grad[k], wg, rtol=1e-9, atol=0)


@pytest.mark.parametrize("precision,vtol,gtol", [("fp64", 1e-11, 1e-9), ("fp32", 1e-5, 2e-3)])
def test_likelihood_gradient_matches_reference_golden(pkg, precision, vtol, gtol):
    """Forward-sensitivity gradient (batched step_tangent) against the
    reference's likelihood_gradient values."""
    meta = json.load(open(os.path.join(GOLDEN, "adjoint_16.json")))
    g = np.load(os.path.join(GOLDEN, "pf_deterministic_16.npz"))
    init = full_from_half(g["initial"], 16)
    evals = meta["likelihood_gradient"]
    q = np.array([e["q"] for e in evals])
    with pkg.Ensemble(16, len(evals), precision, seed=1) as ens:
        v, grad = pkg.likelihood_gradient(ens, init, g["rows"], meta["dt"], q)
    for k, e in enumerate(evals):
        if not np.isfinite(e["value"]):
            assert not np.isfinite(v[k]) and np.all(grad[k] == 0)
            continue
        assert abs(v[k] - e["value"]) < vtol * abs(e["value"])
        np.testing.assert_allclose(grad[k], e["gradient"], rtol=gtol, atol=0)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_step_tangent_matches_oracle(pkg, orc, precision, tol):
    from oracle_filters import oracle_step_tangent
    n, dt, steps = 32, 5e-3, 3
    nus, alphas = [0.01, 0.004], [0.1, 0.02]
    S = np.stack([orc.grf(n, 500 + m) for m in range(2)])
    T = np.stack([[orc.grf(n, 600 + 2 * m), orc.grf(n, 601 + 2 * m)] for m in range(2)])
    with pkg.Ensemble(n, 2, precision, sigma=0.0, amp=0.0, seed=1) as ens:
        for m in range(2):
            ens.set_member_parameters(m, nus[m], alphas[m], 0.0, 0.0)
        ens.set_state(S)
        got_t = ens.step_tangent(dt, steps, T)
        got_w = ens.get_state()
    for m in range(2):
        w, t = S[m], [T[m, 0], T[m, 1]]
        for _ in range(steps):
            w, t = oracle_step_tangent(orc, w, t, nus[m], alphas[m], dt)
        assert np.linalg.norm(got_w[m] - w) / np.linalg.norm(w) < tol
        for j in range(2):
            assert np.linalg.norm(got_t[m, j] - t[j]) / np.linalg.norm(t[j]) < tol
