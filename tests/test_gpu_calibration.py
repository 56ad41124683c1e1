"""GPU SSPRK(3,3) stepping and the calibration drifts (drift, tangent,
adjoint) against the FP64 oracle, which is pinned to the reference's
Solver::step(dt, ssprk3), drift, linearized_drift and adjoint_drift
(src/solver.cpp:235-323) in tests/test_oracle.py."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [32, 64, 256])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_drifts_match_oracle(pkg, orc, n, precision, tol):
    members = 3
    nus, alphas = [0.01, 0.004, 0.02], [0.1, 0.0, 0.05]
    S = np.stack([orc.grf(n, 100 + m) for m in range(members)])
    D = np.stack([orc.grf(n, 200 + m) for m in range(members)])
    L = np.stack([orc.grf(n, 300 + m) for m in range(members)])
    print(f"n={n} {precision}", flush=True)
    with pkg.Ensemble(n, members, precision, sigma=0.0, amp=0.0, seed=1) as ens:
        for m in range(members):
            ens.set_member_parameters(m, nus[m], alphas[m], 0.0, 0.0)
        ens.set_state(S)
        drift = ens.drift()
        tangent = ens.linearized_drift(D)
        adjoint = ens.adjoint_drift(L)
        assert np.array_equal(ens.get_state(), ens.get_state())  # states untouched
    for m in range(members):
        nu, al = nus[m], alphas[m]
        assert rel(drift[m], orc.rhs(S[m], nu, al)) < tol
        assert rel(tangent[m], orc.linearized_drift(S[m], D[m], nu, al)) < tol
        assert rel(adjoint[m], orc.adjoint_drift(S[m], L[m], nu, al)) < tol
    if precision == "fp64":
        # adjoint identity on the GPU results: Re <lam, L d> = Re <L* lam, d>
        with pkg.Ensemble(n, 1, "fp64", sigma=0.0, amp=0.0, seed=1) as e1:
            e1.set_member_parameters(0, nus[0], alphas[0], 0.0, 0.0)
            e1.set_state(S[:1])
            Ld = e1.linearized_drift(D[:1])[0]
            Al = e1.adjoint_drift(L[:1])[0]
        lhs, rhs = np.vdot(L[0], Ld).real, np.vdot(Al, D[0]).real
        assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.parametrize("n", [32, 128])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_ssprk3_matches_oracle(pkg, orc, n, precision, tol):
    """SSPRK(3,3) with the Kolmogorov forcing held (white noise off)."""
    members, nu, alpha, dt, steps = 2, 2e-3, 0.05, 2e-3, 20
    S = np.stack([orc.grf(n, 400 + m) for m in range(members)])
    F = orc.step_forcing(n, 0.1, 4, 0.0, 1, 4, dt, 0, 0)
    with pkg.Ensemble(n, members, precision, nu=nu, alpha=alpha, sigma=0.0, amp=0.1, kf=4,
                      seed=1) as ens:
        ens.set_state(S)
        ens.step(dt, steps, integrator="ssprk3")
        ens.synchronize()
        got = ens.get_state()
        if precision == "fp32" and n == 128:
            # a tuned column shape (E = 32 split) must not change the stages
            ens.set_state(S)
            ens.tuning = (2, 1)
            ens.step(dt, steps, integrator="ssprk3")
            got_tuned = ens.get_state()
    for m in range(members):
        want = orc.ssprk3(S[m], nu, alpha, dt, steps, F)
        assert rel(got[m], want) < tol
        if precision == "fp32" and n == 128:
            assert rel(got_tuned[m], want) < tol


def test_bad_integrator_is_invalid_argument(pkg):
    lib = pkg.load_library()
    with pkg.Ensemble(16, 1, "fp64") as ens:
        with pytest.raises(pkg.HzlInvalidArgument):
            ens.step(1e-3, 1, integrator="euler")
        assert lib.hzl_ensemble_step_integrator(ens._h, 1e-3, 1, 7) == pkg.engine.HZL_EINVAL
