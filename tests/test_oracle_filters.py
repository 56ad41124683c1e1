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
