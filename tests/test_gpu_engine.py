"""GPU parity of the batched RK4 engine through the C-ABI (libhzl_amd.so).

Checker: the FP64 C oracle (oracle/), itself pinned to the reference build,
and the golden vectors of tests/golden/ written by the reference.
Tolerances (relative L2 per member):
  FP64 engine vs FP64 oracle        <= 1e-12 after a few steps
  FP32 engine vs FP64 oracle        <= 1e-5  (after 100 steps at 256^2, the accuracy bound of
                                              the FP32 ensemble path)
"""
from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20261203


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def half(field):
    n = field.shape[0]
    K = (n - 1) // 3
    ky = np.arange(-K, K + 1) % n
    return field[ky][:, : K + 1].T


def full_from_half(h, n):
    K = (n - 1) // 3
    out = np.zeros((n, n), dtype=np.complex128)
    for c in range(K + 1):
        for j, ky in enumerate(range(-K, K + 1)):
            out[ky % n, c] = h[c, j]
            if c > 0:
                out[(-ky) % n, n - c] = np.conj(h[c, j])
    return out


def test_native_library_is_the_one_loaded(pkg):
    lib = pkg.load_library()
    assert os.path.basename(lib._name) == "libhzl_amd.so"
    for name in pkg.EXPORTS:
        assert hasattr(lib, name)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("n", [16, 32, 64, 128, 256])
def test_steps_match_oracle(pkg, orc, n, precision, tol):
    members, steps, nu, dt = 3, 4, 1e-3, 1e-3
    with pkg.Ensemble(n, members, precision, nu=nu, sigma=0.05, amp=0.1, kf=4,
                      kmax=min(4, (n - 1) // 3), seed=SEED) as e:
        e.init_grf()
        e.step(dt, steps)
        e.synchronize()
        got = e.get_state()
    want = orc.trajectories(n, nu, 0.0, dt, steps, SEED, 0, members,
                            kmax=min(4, (n - 1) // 3))
    errs = [rel(got[m], want[m]) for m in range(members)]
    print(n, precision, errs)
    assert max(errs) < tol


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("fp64", 1e-13)])
def test_512_matches_oracle(pkg, orc, precision, tol):
    """512^2: the default split is the three-pass transform (8 values over 64
    lanes), in FP32 and FP64."""
    n, members, steps, nu, dt = 512, 2, 2, 5e-4, 5e-4
    with pkg.Ensemble(n, members, precision, nu=nu, seed=SEED) as e:
        assert e.tuning == (0, 0) or precision == "fp32"
        e.tuning = (0, 0)
        e.init_grf()
        e.step(dt, steps)
        e.synchronize()
        got = e.get_state()
    want = orc.trajectories(n, nu, 0.0, dt, steps, SEED, 0, members)
    errs = [rel(got[m], want[m]) for m in range(members)]
    print(precision, errs)
    assert max(errs) < tol


def test_grf_initial_state_matches_oracle(pkg, orc):
    n, members = 64, 4
    for precision, tol in (("fp64", 1e-13), ("fp32", 1e-6)):
        with pkg.Ensemble(n, members, precision, seed=SEED) as e:
            e.init_grf()
            got = e.get_state()
        for m in range(members):
            assert rel(got[m], orc.grf(n, orc.stream_seed(SEED, m))) < tol


def test_forced_step_matches_oracle(pkg, orc):
    """CudaSolver::step(dt, F) semantics: host forcing held through the stages."""
    n, nu, alpha, dt = 32, 0.01, 0.1, 0.01
    rng = np.random.default_rng(7)
    F = np.fft.fft2(rng.standard_normal((n, n))) / n**2 * 0.5
    F *= orc.active_mask(n)
    w0 = orc.grf(n, 99)
    for precision, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        with pkg.Ensemble(n, 2, precision, nu=nu, alpha=alpha, sigma=0.0, amp=0.0) as e:
            e.set_state(np.stack([w0, w0]))
            for _ in range(3):
                e.step_forced(dt, F, shared=True)
            e.synchronize()
            got = e.get_state()
        want = orc.rk4(w0, nu, alpha, dt, 3, F)
        assert rel(got[0], want) < tol and rel(got[1], want) < tol


def test_state_roundtrip_masks_like_cudasolver(pkg, orc):
    n = 32
    rng = np.random.default_rng(3)
    w = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    with pkg.Ensemble(n, 1, "fp64") as e:
        e.set_state(w, member=0)
        got = e.get_state(0)
    # retained half plane is kept exactly; kx < 0 entries become conjugate mirrors
    np.testing.assert_array_equal(half(got), np.where(
        (np.arange(half(w).shape[0])[:, None] == 0) & (np.arange(half(w).shape[1])[None, :]
                                                       == (n - 1) // 3), 0, half(w)))


def test_diagnostics_match_oracle(pkg, orc):
    n, nu, dt, steps = 64, 1e-3, 1e-3, 3
    with pkg.Ensemble(n, 2, "fp64", nu=nu, seed=SEED) as e:
        e.init_grf()
        e.step(dt, steps)
        e.synchronize()
        d = e.diagnostics()
    for m in range(2):
        key = orc.stream_seed(SEED, m)
        w = orc.trajectory(n, nu, 0.0, dt, steps, key)
        F = orc.step_forcing(n, 0.1, 4, 0.05, 1, 4, dt, key, steps - 1)
        o = orc.diagnostics(w, nu, F)
        for i, k in enumerate(("energy", "enstrophy", "dissipation", "injection", "cfl_rate")):
            assert abs(d[m, i] - o[k]) <= 1e-10 * abs(o[k]), (k, d[m, i], o[k])


def test_invalid_dt_is_invalid_argument(pkg):
    with pkg.Ensemble(32, 1, "fp32", nu=0.5) as e:
        with pytest.raises(pkg.HzlInvalidArgument):
            e.step(float("nan"))
        with pytest.raises(pkg.HzlInvalidArgument):
            e.step(1.0)  # dt * 2 nu K^2 > 2.5


def test_cfl_violation_reported_at_synchronize(pkg):
    with pkg.Ensemble(32, 1, "fp64", nu=0.0, sigma=0.0, amp=0.0, seed=1) as e:
        e.init_grf()
        e.set_state(e.get_state() * 1e3)
        e.step(0.05, 1)
        with pytest.raises(pkg.HzlError, match="CFL"):
            e.synchronize()


def test_nonsymmetric_forcing_rejected(pkg):
    n = 16
    F = np.zeros((n, n), dtype=np.complex128)
    F[0, 1] = 1.0  # partner (0, n-1) left at zero
    with pkg.Ensemble(n, 1, "fp64", sigma=0.0, amp=0.0) as e:
        with pytest.raises(pkg.HzlInvalidArgument, match="conjugate"):
            e.step_forced(0.01, F)


def test_golden_reference_trajectory_64(pkg):
    """The 64^2 trajectory (nu = dt = 1e-3, Kolmogorov 0.1) against the reference
    solver's own."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "trajectory_64.npz"))
    n, nu, dt = 64, 1e-3, 1e-3
    errs = {}
    for precision in ("fp64", "fp32"):
        with pkg.Ensemble(n, 1, precision, nu=nu, sigma=0.05, amp=0.1, kf=4, seed=SEED) as e:
            e.init_grf()
            e.step(dt, 100)
            e.synchronize()
            s100 = half(e.get_state(0))
            e.step(dt, 900)
            e.synchronize()
            s1000 = half(e.get_state(0))
        errs[precision] = (rel(s100, g["w100"]), rel(s1000, g["w1000"]))
    print(errs)
    assert errs["fp64"][0] < 1e-12 and errs["fp64"][1] < 1e-10
    assert errs["fp32"][0] < 1e-5 and errs["fp32"][1] < 1e-4


def test_fp32_accuracy_256_after_100_steps(pkg, orc):
    """Batched RK4 at 256^2 in FP32 within 1e-5 relative of the FP64 oracle after 100 steps."""
    n, members, steps, nu, dt = 256, 8, 100, 5e-4, 5e-4
    with pkg.Ensemble(n, members, "fp32", nu=nu, seed=SEED) as e:
        e.init_grf()
        e.step(dt, steps)
        e.synchronize()
        got = e.get_state()
    want = orc.trajectories(n, nu, 0.0, dt, steps, SEED, 0, members)
    errs = [rel(got[m], want[m]) for m in range(members)]
    print("256^2 FP32 100-step relative errors:", errs)
    assert max(errs) <= 1e-5


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_energy_spectrum_matches_oracle(pkg, orc, precision, tol):
    """Device shell spectrum (Solver::energy_spectrum) of every member after a
    few steps against the FP64 oracle's spectrum of the same downloaded state;
    the bins sum to the member's energy."""
    n, members = 64, 3
    with pkg.Ensemble(n, members, precision, nu=1e-3, sigma=0.05, seed=5) as ens:
        ens.init_grf()
        ens.step(1e-3, 3)
        ens.synchronize()
        bins = ens.energy_spectrum()
        states = ens.get_state()
        diag = ens.diagnostics()
    for m in range(members):
        want = orc.energy_spectrum(states[m])
        assert np.abs(bins[m] - want).max() <= tol * want.max()
        assert abs(bins[m].sum() - diag[m, 0]) <= tol * diag[m, 0]
