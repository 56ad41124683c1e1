"""The C++ drop-in layer (csrc/host/hzl_gpu.hpp: the reference's class and
function signatures over the C-ABI) and the observation-file formats.

CPU: the reference-written HZL_OBSERVATIONS_V2 fixture round-trips byte for
byte through Dataset::load/save and through the Python reader/writer; the host
FFT matches numpy; initial_condition matches the reference's own
Solver::initialize fixtures.  GPU: CudaSolver against the FP64 oracle, and the
reference-signature bootstrap_filter / ensemble_kalman_filter against the
reference golden records.
"""
from __future__ import annotations

import json
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
CHECK = os.path.join(HERE, "dropin", "dropin_check")
CSV = os.path.join(GOLDEN, "observations_v2_16.csv")


def run(*args, timeout=120):
    if not os.path.exists(CHECK):
        pytest.fail(f"{CHECK} not built: run __graft_entry__.build()")
    p = subprocess.run([CHECK, *map(str, args)], capture_output=True, text=True,
                       timeout=timeout)
    assert p.returncode == 0, p.stderr + p.stdout
    return p.stdout


def field(raw, n, count=1):
    c = raw[0::2] + 1j * raw[1::2]
    return c.reshape(count, n, n) if count > 1 else c.reshape(n, n)


def full_from_half(h, n):
    K = (n - 1) // 3
    out = np.zeros((n, n), dtype=np.complex128)
    for c in range(K + 1):
        for j, ky in enumerate(range(-K, K + 1)):
            out[ky % n, c] = h[c, j]
            if c > 0:
                out[(-ky) % n, n - c] = np.conj(h[c, j])
    return out


def test_dataset_csv_round_trip_is_byte_identical(tmp_path, pkg):
    out = tmp_path / "out.csv"
    run("dataset", CSV, out)
    assert out.read_bytes() == open(CSV, "rb").read()
    meta, rows = pkg.load_observations(CSV)
    g = np.load(os.path.join(GOLDEN, "pf_deterministic_16.npz"))
    np.testing.assert_array_equal(rows, g["rows"])
    assert meta == dict(n=16, dt=0.01, stochastic=False, forcing_min=2, forcing_max=3)
    py = tmp_path / "py.csv"
    pkg.save_observations(py, meta, rows)
    assert py.read_bytes() == open(CSV, "rb").read()


def test_dataset_validation_errors(tmp_path, pkg):
    meta, rows = pkg.load_observations(CSV)
    bad = rows.copy()
    bad[0, 4] = 0.0
    with pytest.raises(pkg.HzlInvalidArgument):
        pkg.save_observations(tmp_path / "x.csv", meta, bad)
    bad = rows[::-1].copy()
    with pytest.raises(pkg.HzlInvalidArgument):
        pkg.validate_observations(meta, bad)
    (tmp_path / "h.csv").write_text("HZL_OBSERVATIONS_V3,16,0.01\n")
    with pytest.raises(pkg.HzlError):
        pkg.load_observations(tmp_path / "h.csv")
    p = subprocess.run([CHECK, "dataset", tmp_path / "h.csv", tmp_path / "o.csv"],
                       capture_output=True, text=True)
    assert p.returncode == 1 and "invalid dataset header" in p.stderr


@pytest.mark.parametrize("n", [16, 64])
def test_host_fft_matches_numpy(tmp_path, n):
    run("fft", n, tmp_path / "f.bin")
    raw = np.fromfile(tmp_path / "f.bin")
    fwd, back = field(raw[:2 * n * n], n), field(raw[2 * n * n:], n)
    i = np.arange(n * n)
    x = (np.sin(0.37 * i) + 0.25 + 1j * np.cos(1.3 * i)).reshape(n, n)
    np.testing.assert_allclose(fwd, np.fft.fft2(x) / (n * n), rtol=0, atol=1e-15)
    np.testing.assert_allclose(back, x, rtol=0, atol=1e-13)


def test_initial_condition_matches_reference_fixtures(tmp_path):
    g = np.load(os.path.join(GOLDEN, "pf_deterministic_16.npz"))
    run("initial", 16, "multimode", tmp_path / "a.bin")
    np.testing.assert_allclose(field(np.fromfile(tmp_path / "a.bin"), 16),
                               full_from_half(g["initial"], 16), rtol=0, atol=1e-15)
    d = np.load(os.path.join(GOLDEN, "drift_multimode_32.npz"))
    run("initial", 32, "multimode", tmp_path / "b.bin")
    np.testing.assert_allclose(field(np.fromfile(tmp_path / "b.bin"), 32),
                               full_from_half(d["state"], 32), rtol=0, atol=1e-15)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["solver", "forced"])
def test_cudasolver_matches_oracle(tmp_path, orc, mode):
    n, steps = 64, 100
    run(mode, n, steps, tmp_path / "s.bin")
    got = field(np.fromfile(tmp_path / "s.bin"), n)
    run("initial", n, "multimode", tmp_path / "i.bin")
    w0 = field(np.fromfile(tmp_path / "i.bin"), n)
    F = None
    if mode == "forced":
        F = np.zeros((n, n), dtype=np.complex128)
        F[0, 2], F[0, n - 2], F[3, 0], F[n - 3, 0] = -0.15j, 0.15j, 0.1, 0.1
    want = orc.rk4(w0, 1e-2, 0.1, 1e-3, steps, forcing=F)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12


@pytest.mark.gpu
def test_reference_signature_filters_match_golden(tmp_path):
    g = np.load(os.path.join(GOLDEN, "pf_deterministic_16.npz"))
    meta = json.load(open(os.path.join(GOLDEN, "pf_deterministic_16.json")))
    n = 16
    run("pf", CSV, tmp_path / "pf.bin")
    raw = np.fromfile(tmp_path / "pf.bin")
    assert abs(raw[0] - meta["log_likelihood"]) < 1e-10 and raw[1] == 0
    rec = 4 + 2 * n * n
    for k, (want, mean) in enumerate(zip(g["records"], g["means"])):
        r = raw[2 + k * rec: 2 + (k + 1) * rec]
        assert abs(r[1] - want[0]) < 1e-10 and abs(r[2] - want[1]) < 1e-9
        assert abs(r[3] - want[2]) < 1e-12
        np.testing.assert_allclose(field(r[4:], n).real, mean, rtol=0, atol=1e-10)
    run("enkf", CSV, tmp_path / "enkf.bin")
    raw = np.fromfile(tmp_path / "enkf.bin")
    rec = 3 + 2 * n * n
    for k, (want, mean) in enumerate(zip(g["enkf_records"], g["enkf_means"])):
        r = raw[k * rec:(k + 1) * rec]
        assert r[0] == want[0] and abs(r[1] - want[1]) < 1e-10 and abs(r[2] - want[2]) < 1e-12
        np.testing.assert_allclose(field(r[3:], n).real, mean, rtol=0, atol=1e-10)


OU_CSV = os.path.join(GOLDEN, "ou_observations_16.csv")


def ou_golden():
    return (np.load(os.path.join(GOLDEN, "ou_16.npz")),
            json.load(open(os.path.join(GOLDEN, "ou_16.json"))))


@pytest.mark.gpu
def test_reference_signature_filter_with_ou_forcing_matches_golden(tmp_path):
    """bootstrap_filter(data, p, 8, 555) with sigma > 0 runs the reference's OU
    forcing: the reference's own result over its stochastic dataset (golden
    ou_16), resampling events and per-group records included."""
    g, meta = ou_golden()
    n = meta["n"]
    run("pf_ou", OU_CSV, tmp_path / "pf.bin")
    raw = np.fromfile(tmp_path / "pf.bin")
    assert raw[1] == meta["resampling_events"] >= 1
    assert abs(raw[0] - meta["log_likelihood"]) < 1e-9 * abs(meta["log_likelihood"])
    rec = 4 + 2 * n * n
    for k, (want, mean) in enumerate(zip(g["records"], g["means"])):
        r = raw[2 + k * rec: 2 + (k + 1) * rec]
        assert abs(r[1] - want[0]) < 1e-9 and abs(r[2] - want[1]) < 1e-8
        assert abs(r[3] - want[2]) < 1e-10
        np.testing.assert_allclose(field(r[4:], n).real, mean, rtol=0, atol=1e-10)


@pytest.mark.gpu
def test_reference_signature_particle_metropolis_matches_golden(tmp_path):
    """particle_metropolis with the reference's streams and prior: the same
    proposals, likelihood seeds and accept/reject decisions as the reference's
    chain, likelihoods to 1e-9 (each one a GPU filter with OU forcing)."""
    g, _ = ou_golden()
    count = len(g["chain_accepted"])
    run("pmcmc", OU_CSV, count, tmp_path / "chain.bin", timeout=300)
    raw = np.fromfile(tmp_path / "chain.bin").reshape(count, 8)
    seeds = raw[:, 6].copy().view(np.uint64)
    assert np.array_equal(seeds, g["chain_seeds"])
    assert np.array_equal(raw[:, 7].astype(bool), g["chain_accepted"])
    np.testing.assert_allclose(raw[:, :4], g["chain_q"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(raw[:, 4:6], g["chain_values"], rtol=1e-9, atol=0)


@pytest.mark.gpu
def test_reference_signature_ensemble_with_ou_forcing_matches_golden(tmp_path):
    """ensemble(n, p, dt, steps, count, workers, seed) from zero vorticity with
    OU forcing: every member's energy equals the reference's."""
    g, _ = ou_golden()
    run("ensemble_ou", tmp_path / "e.bin")
    raw = np.fromfile(tmp_path / "e.bin").reshape(-1, 3)
    np.testing.assert_allclose(raw[:, 0], g["ensemble_energy"], rtol=1e-10, atol=0)
    assert np.all(raw[:, 0] > 0)


@pytest.mark.gpu
def test_reference_error_contract():
    assert run("errors").strip() == "ok"


CKPT = os.path.join(GOLDEN, "checkpoint_v1_16.txt")


def test_checkpoint_round_trip_is_byte_identical(tmp_path, pkg):
    """A reference-written HZL_CHECKPOINT_V1 file (Solver::save) loads into
    exactly the fields the reference's own Solver::load produced, and writes
    back byte for byte, from C++ and from Python."""
    g = np.load(os.path.join(GOLDEN, "checkpoint_v1_16.npz"))
    meta = json.load(open(os.path.join(GOLDEN, "checkpoint_v1_16.json")))
    out = tmp_path / "c.txt"
    run("checkpoint", CKPT, out)
    assert out.read_bytes() == open(CKPT, "rb").read()
    c = pkg.load_checkpoint(CKPT)
    assert c["n"] == 16 and c["time"] == meta["reference_load_diagnostics"]["time"]
    assert c["parameters"] == dict(viscosity=0.02, friction=0.1, forcing_sigma=0.3,
                                   forcing_tau=0.5, forcing_min=2, forcing_max=3)
    assert len(c["random_engine"]) == 313 and c["random_engine"][-1] <= 312
    np.testing.assert_array_equal(c["vorticity"], g["vorticity"])
    np.testing.assert_array_equal(c["forcing"], g["forcing"])
    assert np.any(c["forcing"] != 0)  # the OU forcing of the reference run is carried
    py = tmp_path / "p.txt"
    pkg.save_checkpoint(py, c)
    assert py.read_bytes() == open(CKPT, "rb").read()


def test_checkpoint_diagnostics_match_reference_load(orc):
    """The loaded fields give the reference's own diagnostics of the loaded
    solver (energy, enstrophy, dissipation, injection)."""
    meta = json.load(open(os.path.join(GOLDEN, "checkpoint_v1_16.json")))
    import __graft_entry__
    c = __graft_entry__.load_package().load_checkpoint(CKPT)
    want = meta["reference_load_diagnostics"]
    d = orc.diagnostics(c["vorticity"], meta["nu"], forcing=c["forcing"])
    for k in ("energy", "enstrophy", "dissipation", "injection"):
        assert abs(d[k] - want[k]) < 1e-12 * max(1.0, abs(want[k])), k


@pytest.mark.parametrize("edit,message", [
    (lambda t: t.replace("HZL_CHECKPOINT_V1", "HZL_CHECKPOINT_V2", 1), "invalid checkpoint header"),
    (lambda t: t + "1\n", "unexpected checkpoint data"),
    (lambda t: "\n".join(t.split("\n")[:-40]) + "\n", "invalid checkpoint field"),
    # a vorticity value on the highest (masked) wave number, y = 0, x = n/2
    (lambda t: "\n".join(line if i != 5 + 8 else "1 0 0 0"
                         for i, line in enumerate(t.split("\n"))), "spectral constraints"),
])
def test_checkpoint_errors_match_reference(tmp_path, pkg, edit, message):
    bad = tmp_path / "bad.txt"
    bad.write_text(edit(open(CKPT).read()))
    with pytest.raises(pkg.HzlError, match=message):
        pkg.load_checkpoint(bad)
    p = subprocess.run([CHECK, "checkpoint", bad, tmp_path / "o.txt"], capture_output=True,
                       text=True)
    assert p.returncode == 1 and message in p.stderr


@pytest.mark.gpu
def test_cudasolver_resumes_from_checkpoint(tmp_path, orc):
    """CudaSolver built from a reference checkpoint, advanced holding its
    forcing, matches the FP64 oracle's held-forcing RK4 from the loaded state."""
    run("resume", CKPT, 50, tmp_path / "r.bin")
    raw = np.fromfile(tmp_path / "r.bin")
    n = 16
    w0, F, got = (field(raw[k * 2 * n * n:(k + 1) * 2 * n * n], n) for k in range(3))
    want = orc.rk4(w0, 0.02, 0.1, 1e-3, 50, forcing=F)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12


@pytest.mark.gpu
def test_reference_signature_calibration_matches_oracle(tmp_path, orc):
    """hzl::Ensemble::drift / linearized_drift / adjoint_drift /
    step(dt, steps, TimeIntegrator::ssprk3) through the C++ drop-in layer."""
    n = 32
    run("calibration", n, tmp_path / "c.bin")
    raw = np.fromfile(tmp_path / "c.bin")
    f = [field(raw[k * 2 * n * n:(k + 1) * 2 * n * n], n) for k in range(8)]
    run("initial", n, "multimode", tmp_path / "a.bin")
    run("initial", n, "taylor-green", tmp_path / "b.bin")
    a, b = field(np.fromfile(tmp_path / "a.bin"), n), field(np.fromfile(tmp_path / "b.bin"), n)
    nu, al = 1e-2, 0.1
    want = [orc.rhs(a, nu, al), orc.rhs(b, nu, al),
            orc.linearized_drift(a, b, nu, al), orc.linearized_drift(b, a, nu, al),
            orc.adjoint_drift(a, b, nu, al), orc.adjoint_drift(b, a, nu, al),
            orc.ssprk3(a, nu, al, 1e-3, 5), orc.ssprk3(b, nu, al, 1e-3, 5)]
    for g, w in zip(f, want):
        scale = max(np.linalg.norm(w), 1e-300)
        assert np.linalg.norm(g - w) / scale < 1e-12


@pytest.mark.gpu
def test_reference_signature_likelihood_adjoint_matches_golden(tmp_path):
    """hzl::likelihood_adjoint(Dataset, q, stride) through the C++ drop-in on
    the reference-written observation file: value, gradient and the
    reference's bookkeeping counts."""
    meta = json.load(open(os.path.join(GOLDEN, "adjoint_16.json")))
    for stride in (7, 4):
        run("adjoint", CSV, stride, tmp_path / "a.bin")
        raw = np.fromfile(tmp_path / "a.bin").reshape(3, 6)
        want = [e for e in meta["evaluations"] if e["stride"] == stride]
        for got, w in zip(raw, [want[0], want[1], want[3]]):
            if not np.isfinite(w["value"]):
                assert not np.isfinite(got[0]) and got[1] == got[2] == 0
                continue
            assert abs(got[0] - w["value"]) < 1e-11 * abs(w["value"])
            np.testing.assert_allclose(got[1:3], w["gradient"], rtol=1e-9, atol=0)
            assert got[3] == w["checkpoint_fields"] and got[5] == w["replayed_steps"]


@pytest.mark.gpu
def test_stochastic_resume_from_reference_checkpoint(tmp_path, pkg):
    """A GPU member loads a reference checkpoint of a stochastic (OU) run --
    vorticity, mt19937_64 stream and forcing -- and continues it: after 15
    steps its state, forcing and stream equal the reference Solver's own
    checkpoint 15 steps later (C++ drop-in and Python binding)."""
    before = os.path.join(GOLDEN, "checkpoint_ou_before_16.txt")
    after = pkg.load_checkpoint(os.path.join(GOLDEN, "checkpoint_ou_after_16.txt"))
    run("resume_ou", before, 15, tmp_path / "resumed.txt")
    got = pkg.load_checkpoint(str(tmp_path / "resumed.txt"))
    assert got["random_engine"] == after["random_engine"]
    assert got["time"] == after["time"]
    for key, tol in (("vorticity", 1e-12), ("forcing", 1e-13)):
        err = np.linalg.norm(got[key] - after[key]) / np.linalg.norm(after[key])
        assert err < tol, (key, err)
    # the same through the Python binding
    c = pkg.load_checkpoint(before)
    p = c["parameters"]
    with pkg.Ensemble(16, 1, "fp64", nu=p["viscosity"], alpha=p["friction"],
                      sigma=p["forcing_sigma"], tau=p["forcing_tau"], kmin=p["forcing_min"],
                      kmax=p["forcing_max"], amp=0.0) as ens:
        ens.set_state(c["vorticity"][None])
        ens.set_forcing_process("ou")
        ens.set_forcing_state(0, c["random_engine"], c["forcing"])
        ens.step(0.01, 15)
        words, F = ens.get_forcing_state(0)
        w = ens.get_state(0)
    assert [int(x) for x in words] == after["random_engine"]
    assert np.linalg.norm(w - after["vorticity"]) < 1e-12 * np.linalg.norm(after["vorticity"])
    assert np.linalg.norm(F - after["forcing"]) < 1e-13 * np.linalg.norm(after["forcing"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["stochastic", "deterministic"])
def test_synthetic_observations_match_reference(tmp_path, case):
    """synthetic_observations on the GPU (truth member with the reference's OU
    forcing, noise from NormalStream(stream_seed(seed, 1))) against the rows the
    reference's own synthetic_observations wrote: sensor positions and steps
    exact, values to 1e-10."""
    if case == "stochastic":  # golden ou_16: (16, 0.01, 60, 10, sigma 0.2, noise 0.05, 987)
        g, _ = ou_golden()
        args, rows = (0.2, 10, 60, 0.05), g["rows"]
    else:  # golden pf_deterministic_16: (16, 0.01, 30, 10, sigma 0, noise 0.2, 987)
        args = (0.0, 10, 30, 0.2)
        rows = np.load(os.path.join(GOLDEN, "pf_deterministic_16.npz"))["rows"]
    run("synthetic", *args, tmp_path / "s.bin")
    raw = np.fromfile(tmp_path / "s.bin")
    got = raw[:rows.size].reshape(rows.shape)
    np.testing.assert_array_equal(got[:, [0, 1, 2, 4]], rows[:, [0, 1, 2, 4]])
    np.testing.assert_allclose(got[:, 3], rows[:, 3], rtol=0, atol=1e-10)
    truth = field(raw[rows.size:], 16, count=len(rows) // 8)
    assert np.abs(truth.imag).max() < 1e-12  # physical fields of the truth run


@pytest.mark.gpu
def test_energy_spectra_match_oracle(tmp_path, orc):
    n = 32
    run("spectrum", n, tmp_path / "b.bin")
    bins = np.fromfile(tmp_path / "b.bin").reshape(2, n)
    for m, name in enumerate(("multimode", "taylor-green")):
        run("initial", n, name, tmp_path / "w.bin")
        want = orc.energy_spectrum(field(np.fromfile(tmp_path / "w.bin"), n))
        np.testing.assert_allclose(bins[m], want, rtol=1e-13, atol=1e-17)
    assert abs(bins[1][1] - np.pi ** 2) < 1e-12  # Taylor-Green: pi^2 in shell 1


@pytest.mark.gpu
def test_particle_likelihood_family_batches(tmp_path):
    """particle_log_likelihood runs one GPU filter; the averaged likelihood runs
    its replicates as one batch of filters and equals the log-mean-exp of the
    single evaluations with the replicate seeds stream_seed(seed, i); one
    replicate is the single evaluation itself (reference
    tests/verification.cpp:198-202); a q outside the prior is -inf."""
    run("likelihoods", OU_CSV, tmp_path / "l.bin")
    v = np.fromfile(tmp_path / "l.bin")
    single, avg3, avg1, prior, outside = v[:3], v[3], v[4], v[5], v[6]
    assert np.all(np.isfinite(single))
    top = single.max()
    assert abs(avg3 - (top + np.log(np.mean(np.exp(single - top))))) < 1e-12 * abs(avg3)
    assert avg1 == single[0]
    assert prior == 0.0 and outside == -np.inf
