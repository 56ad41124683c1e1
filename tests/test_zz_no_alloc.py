"""hzl_ensemble_diagnostics_noalloc (allocation-free variant of
hzl_ensemble_diagnostics, staged through the library's fixed device array)
against the allocating entry point, through the C-ABI: identical results on
two calls in a row, the second on a smaller ensemble, and a clear error
beyond the array's member limit."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
_dp = C.POINTER(C.c_double)


def noalloc(lib, ens):
    out = np.full((ens.members, 6), np.nan)
    rc = lib.hzl_ensemble_diagnostics_noalloc(ens._h, out.ctypes.data_as(_dp))
    return rc, out


def test_noalloc_diagnostics_match_the_allocating_entry_point(pkg):
    lib = pkg.load_library()
    # first call: 12 members at 64^2 FP64; second: 5 members at 32^2 FP32 (smaller)
    for n, members, precision, seed in ((64, 12, "fp64", 3), (32, 5, "fp32", 4)):
        with pkg.Ensemble(n, members, precision, nu=1e-3, sigma=0.05, seed=seed) as ens:
            ens.init_grf()
            ens.step(1e-3, 3)
            ens.synchronize()
            rc, got = noalloc(lib, ens)
            assert rc == 0, lib.hzl_last_error()
            want = ens.diagnostics()
            np.testing.assert_array_equal(got, want)
            assert np.all(got[:, 0] > 0) and np.all(got[:, 5] == 0)
            # again after more steps: the shared array holds the last call's
            # values and every entry is rewritten
            ens.step(1e-3, 2)
            ens.synchronize()
            rc, again = noalloc(lib, ens)
            assert rc == 0
            np.testing.assert_array_equal(again, ens.diagnostics())
            assert not np.array_equal(again, got)


def test_noalloc_diagnostics_reject_too_many_members(pkg):
    lib = pkg.load_library()
    with pkg.Ensemble(16, 65537, "fp64", nu=1e-3, sigma=0.0, seed=1) as ens:
        rc, _ = noalloc(lib, ens)
        assert rc == pkg.engine.HZL_EINVAL
        assert b"at most 65536 members" in lib.hzl_last_error()
