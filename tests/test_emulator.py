"""CPU checks of the engine's kernel bodies through the lane emulator.

tests/emulator/emu.cpp runs the unmodified kernel bodies of
hzl-stochasticflowinference_amd/csrc/kernels.hpp on the host (one thread per
lane, std::barrier for the workgroup barrier) with the engine's launch
sequence, so layouts, FFT index maps, the fused RK4 stages and the forcing are
checked against the FP64 oracle without a GPU.
"""
from __future__ import annotations

import numpy as np
import pytest

SEED = 20261203


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision,tol", [("fp32", 3e-7), ("fp64", 2e-15)])
@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512])
def test_group_fft_lines(emu, n, precision, tol):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    assert rel(emu.fft_lines(x, False, precision), np.fft.fft(x, axis=1)) < tol
    assert rel(emu.fft_lines(x, True, precision), np.fft.ifft(x, axis=1) * n) < tol


@pytest.mark.parametrize("precision,tol", [("fp32", 3e-7), ("fp64", 2e-15)])
@pytest.mark.parametrize("n,elems", [(64, 8), (128, 8), (128, 32), (256, 8), (256, 32),
                                     (512, 8)])
def test_group_fft_splits(emu, n, elems, precision, tol):
    """Other T x E splits of the launch-shape variants: two passes where each
    lane holds at least T values, three passes (E = 8 over 8 F lanes) where
    it holds fewer; forward and inverse against numpy."""
    rng = np.random.default_rng(n + elems)
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    assert rel(emu.fft_lines(x, False, precision, elems), np.fft.fft(x, axis=1)) < tol
    assert rel(emu.fft_lines(x, True, precision, elems), np.fft.ifft(x, axis=1) * n) < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 1e-6)])
@pytest.mark.parametrize("n", [16, 32, 64])
def test_rk4_pipeline_matches_oracle(emu, orc, n, precision, tol):
    kmax = min(4, (n - 1) // 3)
    got, diag = emu.run(n, 3, precision, 1e-3, 3, kmax=kmax, seed=SEED)
    want = orc.trajectories(n, 1e-3, 0.0, 1e-3, 3, SEED, 0, 3, kmax=kmax)
    for m in range(3):
        assert rel(got[m], want[m]) < tol
        d = orc.diagnostics(want[m], 1e-3)
        assert abs(diag[m, 0] - d["energy"]) < 1e-6 * d["energy"]
        assert abs(diag[m, 1] - d["enstrophy"]) < 1e-6 * d["enstrophy"]


def test_forced_pipeline_matches_oracle(emu, orc):
    n, nu, alpha, dt = 32, 0.01, 0.1, 0.01
    rng = np.random.default_rng(7)
    F = np.fft.fft2(rng.standard_normal((n, n))) / n**2 * orc.active_mask(n)
    w0 = orc.grf(n, 99)
    got, _ = emu.run(n, 1, "fp64", dt, 3, nu=nu, alpha=alpha, sigma=0.0, amp=0.0, init=w0,
                     forcing=F)
    assert rel(got[0], orc.rk4(w0, nu, alpha, dt, 3, F)) < 1e-13


@pytest.mark.parametrize("n,cv,rv", [(128, 1, 0), (128, 2, 0), (128, 3, 0), (128, 0, 1),
                                     (128, 0, 2), (128, 0, 3), (256, 2, 0), (256, 0, 3),
                                     (256, 0, 4), (128, 0, 4), (256, 0, 5),
                                     (256, 7, 0), (256, 0, 9), (256, 0, 10),
                                     (256, 0, 11), (256, 9, 0), (256, 10, 0), (256, 0, 12),
                                     (256, 0, 13), (256, 0, 14), (512, 4, 0), (512, 0, 4),
                                     (512, 0, 5), (128, 5, 0), (128, 6, 0), (128, 7, 0),
                                     (128, 0, 6), (128, 0, 7), (128, 0, 8)])
def test_launch_shape_variants_match_oracle(emu, orc, n, cv, rv):
    """Autotuner variants (other T x E splits, workgroup sizes, row tiles)
    compute the same RK4 step as the FP64 oracle."""
    got, _ = emu.run_variant(n, 1, cv, rv, 1e-3, 2, seed=SEED)
    want = orc.trajectories(n, 1e-3, 0.0, 1e-3, 2, SEED, 0, 1, kmax=4)
    assert rel(got[0], want[0]) < 2e-7


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-14), ("fp32", 2e-6)])
def test_calibration_drifts_and_ssprk3_match_oracle(emu, orc, precision, tol):
    """Drift, its tangent and adjoint (two-state column pass, row pass,
    adjoint products) and SSPRK(3,3) stages, from the kernel bodies."""
    n, nu, alpha = 32, 0.01, 0.1
    S = np.stack([orc.grf(n, 60 + i) for i in range(2)])
    D = np.stack([orc.grf(n, 70 + i) for i in range(2)])
    got = {k: emu.calibration(k, S, D if k in ("tangent", "adjoint") else None, nu=nu,
                              alpha=alpha, precision=precision, dt=2e-3, steps=3)
           for k in ("drift", "tangent", "adjoint", "ssprk3")}
    for m in range(2):
        assert rel(got["drift"][m], orc.rhs(S[m], nu, alpha)) < tol
        assert rel(got["tangent"][m], orc.linearized_drift(S[m], D[m], nu, alpha)) < tol
        assert rel(got["adjoint"][m], orc.adjoint_drift(S[m], D[m], nu, alpha)) < tol
        assert rel(got["ssprk3"][m], orc.ssprk3(S[m], nu, alpha, 2e-3, 3)) < tol


@pytest.mark.parametrize("n,kmin,kmax,members,calls", [(32, 1, 4, 3, 20), (64, 1, 8, 2, 5)])
def test_ou_body_matches_oracle(emu, orc, n, kmin, kmax, members, calls):
    """The device OU kernel body (parallel draws, two-phase regeneration of the
    mt19937_64 block, resampling reseed) on the CPU lanes equals the oracle's
    sequential NormalStream process; 98 pairs exercise more pairs than lanes and
    a regeneration inside one call.  The midpoint forcing lands at the engine
    pair index, conjugated where the canonical mode is the partner."""
    first = orc.ou_pairs(n, kmin, kmax)
    P = len(first)
    rng = np.random.default_rng(1)
    pmap = np.arange(P)
    rng.shuffle(pmap)
    conj = rng.random(P) < 0.4
    emap = np.where(conj, ~pmap, pmap)
    keys0 = np.array([orc.stream_seed(99, m) for m in range(members)], dtype=np.uint64)
    modes = [0] + [1] * (calls - 1)
    modes[calls // 2] = 2
    keys = np.tile(keys0, (calls, 1))
    keys[calls // 2] = [orc.stream_seed(7, m) for m in range(members)]
    hs = [0.0] + [0.004] * (calls - 1)
    sigma = np.array([0.05 * (m + 1) for m in range(members)])
    tau = np.array([0.37 * (1 + 0.5 * m) for m in range(members)])  # per-member time constants
    F, noise = emu.ou(P, emap, keys, sigma, tau, modes, hs)
    for m in range(members):
        proc = orc.ou_process(n, sigma[m], tau[m], kmin, kmax, int(keys0[m]))
        for c in range(calls):
            if c > 0:
                if modes[c] == 2:
                    proc.reseed(int(keys[c][m]))
                else:
                    proc.transition(hs[c])
            want = proc.F.ravel()[first]
            assert np.abs(F[c, m] - want).max() <= 1e-15 * np.abs(want).max()
            if modes[c] == 1:
                got = np.where(conj, np.conj(noise[c, m][pmap]), noise[c, m][pmap])
                assert np.abs(got - want).max() <= 1e-15 * np.abs(want).max()


@pytest.mark.parametrize("lanes", [64, 256])
def test_device_normal_stream_is_the_reference_stream(emu, orc, lanes):
    """normal_stream_body (the EnKF's perturbations on the device) continues
    one NormalStream across calls of any size: a kept r sin first, odd
    remainders, several 312-word blocks in one call, empty calls; every draw
    equals the oracle's NormalStream (itself the reference's, src/solver.cpp:41-54)."""
    seed = orc.stream_seed(77, 1024)
    counts = [1, 7, 0, 311, 312, 313, 2, 1000, 5, 156, 157]
    got = emu.normal_stream(seed, counts, lanes)
    st = orc.mt(seed)
    want = np.array([orc.normal(st) for _ in range(sum(counts))])
    # same draws; libm sin / cos may differ by an ulp (a misordered draw is O(1))
    np.testing.assert_allclose(got, want, rtol=1e-15, atol=1e-15)
