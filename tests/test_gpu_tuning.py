"""Launch-shape autotuner on the GPU.

Every compiled variant of the FP32 column and row kernels at 128^2, 256^2 and
512^2 reproduces the FP64 oracle within the FP32 tolerance; the tuner times
them, keeps a valid choice, remembers it in its cache file, and a new
ensemble of the same size applies the remembered choice.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SEED = 20261203


@pytest.mark.parametrize("n", [128, 256, 512])
def test_every_variant_matches_oracle(pkg, orc, n):
    members, steps, dt = 2, 3, 1e-3
    want = orc.trajectories(n, 1e-3, 0.0, dt, steps, SEED, 0, members, kmax=4)
    with pkg.Ensemble(n, members, "fp32", nu=1e-3, sigma=0.05, amp=0.1, kf=4,
                      seed=SEED) as ens:
        nc, nr = ens.variant_count()
        assert nc > 1 and nr > 1
        shapes = [(c, 0) for c in range(nc)] + [(0, r) for r in range(1, nr)]
        for cv, rv in shapes:
            print(f"n={n} col variant {cv} row variant {rv}", flush=True)
            ens.tuning = (cv, rv)
            assert ens.tuning == (cv, rv)
            ens.step_index = 0
            ens.init_grf()
            ens.step(dt, steps)
            ens.synchronize()
            got = ens.get_state()
            for m in range(members):
                err = np.linalg.norm(got[m] - want[m]) / np.linalg.norm(want[m])
                assert err < 1e-5, (cv, rv, m, err)


def test_fp64_and_small_grids_have_only_the_default(pkg):
    with pkg.Ensemble(256, 1, "fp64") as a, pkg.Ensemble(64, 1, "fp32") as b:
        assert a.variant_count() == (1, 1) and b.variant_count() == (1, 1)
        with pytest.raises(ValueError):
            b.tuning = (1, 0)


def test_autotune_times_variants_and_remembers(pkg, tmp_path, monkeypatch):
    cache = tmp_path / "shapes.txt"
    monkeypatch.setenv("HZL_TUNE_CACHE", str(cache))
    with pkg.Ensemble(256, 64, "fp32", seed=3) as ens:
        ens.init_grf()
        before = ens.get_state()
        col, row = ens.autotune()
        after = ens.get_state()
        cv, rv = ens.tuning
        print("col ms", col, "row ms", row, "choice", cv, rv)
        assert np.all(col > 0) and np.all(row > 0)
        assert 0 <= cv < len(col) and 0 <= rv < len(row)
        assert np.array_equal(before, after)  # states restored
    lines = cache.read_text().split("\n")
    assert any(line.endswith(f" {cv} {rv}") and "/n256/fp32/b64" in line for line in lines)
    with pkg.Ensemble(256, 64, "fp32", seed=3) as again:
        again.autotune(use_cache=True)
        assert again.tuning == (cv, rv)
