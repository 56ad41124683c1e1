"""The reference CLI's `ensemble`, `filter` and `particle-sample` flows on the GPU
engine (hzl_amd_cli), against the files the reference's own executable wrote
for the same command lines (tests/golden/cli, from tests/golden/make_golden.py
cli16): the same file set, the same CSV headers and integer / text columns,
floating-point columns to 1e-9 relative (FP64 GPU filters vs the reference's
CPU solver), the same summary keys and values, the same resolved options, and
byte-identical dataset copies.  The CPU tests check the error contract, which
fails before any GPU call."""
from __future__ import annotations

import csv
import json
import math
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
CLI = os.path.join(ROOT, "hzl-stochasticflowinference_amd", "hzl_amd_cli")
CASES = json.load(open(os.path.join(GOLDEN, "cli", "cases.json")))


def cli(*args, cwd=None, timeout=300):
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} not built: run __graft_entry__.build()")
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, cwd=cwd,
                          timeout=timeout)


def test_error_contract(tmp_path):
    """Messages and exit status of the reference's option handling
    (src/main.cpp:23-111,1517-1521); none of these reaches the GPU."""
    cases = [
        (["frob"], "error: unknown command: frob\n"),
        (["filter", "--particles"], "error: options require unique --name value pairs\n"),
        (["filter", "--seed", "1", "--seed", "2"],
         "error: options require unique --name value pairs\n"),
        (["filter", "particles", "8"], "error: options require unique --name value pairs\n"),
        (["filter", "--bogus", "1", "--output", str(tmp_path / "a")],
         "error: unknown option --bogus\n"),
        (["filter", "--particles", "8x"], "error: invalid value for --particles\n"),
        (["ensemble", "--dt", "inf"], "error: nonfinite value for --dt\n"),
        (["particle-sample", "--initial-nu", "-1"],
         "error: initial stochastic parameters must be positive\n"),
    ]
    for args, message in cases:
        p = cli(*args, cwd=tmp_path)
        assert p.returncode == 1 and p.stderr == message, (args, p.stderr)
    full = tmp_path / "full"
    full.mkdir()
    (full / "x").write_text("x")
    p = cli("ensemble", "--output", full, cwd=tmp_path)
    assert p.returncode == 1
    assert p.stderr == "error: output directory is not empty; choose a new --output directory\n"
    p = cli("--help")
    assert p.returncode == 0 and "particle-sample" in p.stdout


def read_csv(path):
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    return rows[0], rows[1:]


def close(a: str, b: str, rtol: float) -> bool:
    if a == b:
        return True
    try:
        x, y = float(a), float(b)
    except ValueError:
        return False
    if math.isnan(x) or math.isnan(y):
        return math.isnan(x) and math.isnan(y)
    return abs(x - y) <= rtol * max(abs(x), abs(y), 1e-300) or abs(x - y) < 1e-13


def compare_csv(got, want, exact_columns, rtol=1e-9):
    gh, gr = read_csv(got)
    wh, wr = read_csv(want)
    assert gh == wh and len(gr) == len(wr), (got, len(gr), len(wr))
    for r_got, r_want in zip(gr, wr):
        for name, a, b in zip(wh, r_got, r_want):
            if name in exact_columns:
                assert a == b, (got, name, a, b)
            else:
                assert close(a, b, rtol), (got, name, a, b)


def compare_json(got, want, skip=("wall_seconds",), rtol=1e-9):
    g, w = json.load(open(got)), json.load(open(want))
    assert set(g) == set(w)
    for k in w:
        if k in skip:
            continue
        if isinstance(w[k], float):
            assert close(repr(g[k]), repr(w[k]), rtol), (got, k, g[k], w[k])
        else:
            assert g[k] == w[k], (got, k, g[k], w[k])


def run_case(case, tmp_path):
    args = [a.replace("{golden}", GOLDEN) for a in CASES[case]]
    p = cli(*args, "--output", "out", cwd=tmp_path)
    assert p.returncode == 0, p.stderr
    out, want = tmp_path / "out", os.path.join(GOLDEN, "cli", case)
    produced = sorted(os.listdir(out))
    assert produced == sorted(f for f in os.listdir(want) if f != "stdout.txt")
    gi, wi = json.load(open(out / "invocation.json")), json.load(open(os.path.join(want, "invocation.json")))
    assert gi["command"] == wi["command"] and gi["schema"] == wi["schema"]
    go, wo = dict(gi["options"]), dict(wi["options"])
    go.pop("data", None), wo.pop("data", None)
    assert go == wo
    first = open(os.path.join(want, "stdout.txt")).read().split(":")[0]
    assert p.stdout.split(":")[0] == first
    return out, want


@pytest.mark.gpu
def test_filter_particle_matches_reference_cli(tmp_path):
    out, want = run_case("filter_particle", tmp_path)
    compare_csv(out / "filter.csv", os.path.join(want, "filter.csv"), {"step"})
    compare_csv(out / "reconstruction.csv", os.path.join(want, "reconstruction.csv"),
                {"step", "ix", "iy"}, rtol=1e-8)
    compare_json(out / "summary.json", os.path.join(want, "summary.json"))
    assert open(out / "observations.csv", "rb").read() == \
        open(os.path.join(want, "observations.csv"), "rb").read()


@pytest.mark.gpu
def test_filter_enkf_matches_reference_cli(tmp_path):
    out, want = run_case("filter_enkf", tmp_path)
    compare_csv(out / "filter.csv", os.path.join(want, "filter.csv"), {"step", "phase"})
    compare_csv(out / "reconstruction.csv", os.path.join(want, "reconstruction.csv"),
                {"step", "phase", "ix", "iy"}, rtol=1e-8)
    compare_json(out / "summary.json", os.path.join(want, "summary.json"))


@pytest.mark.gpu
def test_filter_enkf_with_ou_forcing_matches_reference_cli(tmp_path):
    """filter --method enkf with sigma > 0: the OU forcing analysed as a latent
    state, with inflation, localisation and free forecasts."""
    out, want = run_case("filter_enkf_ou", tmp_path)
    compare_csv(out / "filter.csv", os.path.join(want, "filter.csv"), {"step", "phase"})
    compare_csv(out / "reconstruction.csv", os.path.join(want, "reconstruction.csv"),
                {"step", "phase", "ix", "iy"}, rtol=1e-8)
    compare_json(out / "summary.json", os.path.join(want, "summary.json"))


@pytest.mark.gpu
def test_particle_sample_matches_reference_cli(tmp_path):
    out, want = run_case("particle_sample", tmp_path)
    compare_csv(out / "chain.csv", os.path.join(want, "chain.csv"),
                {"iteration", "likelihood_seed", "accepted"})
    compare_json(out / "summary.json", os.path.join(want, "summary.json"))
    assert open(out / "training.csv", "rb").read() == \
        open(os.path.join(want, "training.csv"), "rb").read()


@pytest.mark.gpu
def test_ensemble_matches_reference_cli(tmp_path):
    out, want = run_case("ensemble", tmp_path)
    compare_csv(out / "members.csv", os.path.join(want, "members.csv"),
                {"member", "seed", "time"}, rtol=1e-10)
    compare_json(out / "summary.json", os.path.join(want, "summary.json"),
                 skip=("wall_seconds", "energy_standard_error"))
    # both sides print the standard error with 6 decimals (std::to_string):
    # equal, or one unit of the sixth decimal apart where the FP64 values
    # straddle a rounding boundary
    g_se = json.load(open(out / "summary.json"))["energy_standard_error"]
    w_se = json.load(open(os.path.join(want, "summary.json")))["energy_standard_error"]
    assert g_se > 0 and abs(g_se - w_se) <= 1e-6 + 1e-12, (g_se, w_se)
    g = json.load(open(out / "metadata.json"))
    w = json.load(open(os.path.join(want, "metadata.json")))
    for k in ("schema", "n", "dt", "integrator", "steps_this_run", "seed", "initial", "nu",
              "alpha", "sigma", "tau", "kmin", "kmax", "final_time"):
        assert g[k] == w[k], k
