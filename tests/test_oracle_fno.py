"""Pins of the surrogate network reference (tests/oracle_fno.py) on the CPU.

The reference's transforms are explicit DFT sums; they are checked here
against torch.fft and the FP64 solver oracle's inverse transform, and its
autograd gradients against central finite differences, so that a dropped
weight, a sign or a transposed index in the reference fails one of them.
"""
from __future__ import annotations

import math

import numpy as np
import torch

import oracle_fno as F


def real_field(rng, n, kmax):
    """Hermitian n x n spectrum with random modes |kx|, |ky| <= kmax."""
    ph = rng.standard_normal((n, n))
    spec = np.fft.fft2(ph) / (n * n)
    k = np.fft.fftfreq(n, 1.0 / n)
    mask = (np.abs(k)[:, None] <= kmax) & (np.abs(k)[None, :] <= kmax)
    spec = spec * mask
    spec[0, 0] = 0
    return spec


def test_dft_modes_match_torch_fft():
    s = F.Shape(64, 32, 8, 2, 4)
    rng = np.random.default_rng(1)
    v = torch.as_tensor(rng.standard_normal((2, s.G, s.G, 3)))
    X = F.dft_modes(s, v)
    R = torch.fft.fft2(v, dim=(1, 2))  # [s][ky][kx][c]
    ky = F.ky_of(s.KY) % s.G
    want = R[:, ky][:, :, :s.KX].permute(0, 2, 1, 3)
    torch.testing.assert_close(X, want, rtol=0, atol=1e-10)


def test_idft_inverts_dft_on_kept_band():
    """A real field whose modes lie in the kept band is reproduced exactly,
    which fixes the 1/G^2 normalisation and the weights 1 (kx = 0), 2 (kx > 0)."""
    s = F.Shape(64, 32, 8, 2, 4)
    rng = np.random.default_rng(2)
    G = s.G
    y, x = np.meshgrid(np.arange(G), np.arange(G), indexing="ij")
    v = np.zeros((1, G, G, 2))
    for c in range(2):
        for _ in range(5):
            kx, ky = rng.integers(0, s.KX), rng.integers(-s.KY // 2 + 1, s.KY // 2)
            a, b = rng.standard_normal(2)
            v[0, :, :, c] += a * np.cos(2 * np.pi * (kx * x + ky * y) / G) + \
                b * np.sin(2 * np.pi * (kx * x + ky * y) / G)
    vt = torch.as_tensor(v)
    back = F.idft_modes(s, F.dft_modes(s, vt))
    torch.testing.assert_close(back, vt, rtol=0, atol=1e-10)


def test_sample_grid_is_the_physical_field(orc):
    s = F.Shape(64, 32, 8, 2, 4)
    rng = np.random.default_rng(3)
    f = real_field(rng, s.n, s.KS)
    phys = orc.fft2d(f, inverse=True).real
    step = s.n // s.G
    got = F.sample_grid(s, f[None])[0].numpy()
    np.testing.assert_allclose(got, phys[::step, ::step], rtol=0, atol=1e-12)


def test_band_increment_inverts_sampling():
    s = F.Shape(128, 64, 8, 2, 4)
    rng = np.random.default_rng(4)
    f = real_field(rng, s.n, s.KS)
    back = F.band_increment(s, F.sample_grid(s, f[None]).numpy())[0]
    np.testing.assert_allclose(back, f, rtol=0, atol=1e-13)


def test_autograd_matches_finite_differences():
    s = F.Shape(64, 32, 8, 2, 4)
    rng = np.random.default_rng(5)
    fields = np.stack([real_field(rng, s.n, s.KS) for _ in range(2)])
    nxt = fields + 0.1 * np.stack([real_field(rng, s.n, s.KS) for _ in range(2)])
    nu = np.array([1e-3, 2e-3])
    amp = np.array([0.05, 0.4])
    p = 0.3 * rng.standard_normal(s.count())
    loss, grad = F.loss_and_grad(s, p, fields, nxt, nu, 2.0, 3.0, amp)
    for trial in range(3):
        d = rng.standard_normal(s.count())
        h = 1e-5
        lp = float(F.loss(s, torch.as_tensor(p + h * d), fields, nxt, nu, 2.0, 3.0, amp))
        lm = float(F.loss(s, torch.as_tensor(p - h * d), fields, nxt, nu, 2.0, 3.0, amp))
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad @ d) < 1e-6 * max(1.0, abs(fd)), (trial, fd, grad @ d)
    assert math.isfinite(loss) and loss > 0


def test_network_is_translation_equivariant_in_x():
    """Shifting the input along x shifts the output (the only y-dependent
    feature is the forcing pattern, so y shifts break equivariance)."""
    s = F.Shape(64, 32, 8, 2, 4)
    rng = np.random.default_rng(6)
    f = real_field(rng, s.n, s.KS)
    p = torch.as_tensor(0.3 * rng.standard_normal(s.count()))
    shift = 3  # points of the G grid (n / G points of the state grid each)
    k = np.fft.fftfreq(s.n, 1.0 / s.n)
    ph = np.exp(-2j * np.pi * k[None, :] * shift * (s.n // s.G) / s.n)
    g = f * ph
    a = F.forward(s, p, f[None], [1e-3], 1.5)[0].numpy()
    b = F.forward(s, p, g[None], [1e-3], 1.5)[0].numpy()
    np.testing.assert_allclose(np.roll(a, shift, axis=1), b, rtol=0, atol=1e-10)

