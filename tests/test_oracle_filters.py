"""CPU pins of the filter oracle helpers (tests/oracle_filters.py)."""
from __future__ import annotations

import numpy as np

from oracle_filters import gaspari_cohn, localized_scalar_update, taper_field


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
