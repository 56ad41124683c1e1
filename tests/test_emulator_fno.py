"""CPU checks of the fused surrogate layer passes (csrc/fno_fused.hpp) through
the lane emulator, against numpy in FP64 on the same BF16-rounded operands with
the BF16 roundings the passes make (the x-transformed rows, the spectral
output, the inverse-y rows, the layer output).

The transform matrices are built here from their definitions (the surrogate
stores the same ones, csrc/surrogate.hip FnoConstants):
  Fx  [2kx+ri][x]        (cos, -sin)(2 pi kx x / G)
  Fy2 [2kyi+ro][ri G+y]  [[c, s], [-s, c]],  theta = 2 pi ky y / G
  Gy2 [ro G+y][2kyi+ri]  [[c, -s], [s, c]]
  Gx  [x][2kx+ri]        w / G^2 (cos, -sin),  w = 1 (kx = 0) else 2
A planted indexing mistake in a pass (a transposed operand, a swapped re/im
half, a wrong row block) moves whole rows of the output and fails these
tolerances by orders of magnitude.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from test_emulator_gemm import bf16, gelu

G, C, KX, KY = 64, 32, 16, 32


def constants():
    fx = np.zeros((2 * KX, G))
    gx = np.zeros((G, 2 * KX))
    for kx in range(KX):
        ph = 2 * np.pi * kx * np.arange(G) / G
        w = (1.0 if kx == 0 else 2.0) / (G * G)
        fx[2 * kx], fx[2 * kx + 1] = np.cos(ph), -np.sin(ph)
        gx[:, 2 * kx], gx[:, 2 * kx + 1] = w * np.cos(ph), -w * np.sin(ph)
    fy = np.zeros((2 * KY, 2 * G))
    gy = np.zeros((2 * G, 2 * KY))
    y = np.arange(G)
    for kyi in range(KY):
        ky = kyi if kyi < KY // 2 else kyi - KY
        th = 2 * np.pi * ky * y / G
        c, s = np.cos(th), np.sin(th)
        fy[2 * kyi, :G], fy[2 * kyi, G:] = c, s
        fy[2 * kyi + 1, :G], fy[2 * kyi + 1, G:] = -s, c
        gy[:G, 2 * kyi], gy[:G, 2 * kyi + 1] = c, -s
        gy[G:, 2 * kyi], gy[G:, 2 * kyi + 1] = s, c
    return [bf16(m).astype(np.float64) for m in (fx, fy, gy, gx)]


def close(got, want, rel=1.2e-2):
    """BF16 outputs: elementwise within 2 ulp of BF16 (2^-7) of the larger of
    the value and 5% of the field's scale (a BF16 rounding flip of one
    intermediate moves an output by ~2^-8 of that intermediate)."""
    scale = np.max(np.abs(want))
    err = np.abs(got - want)
    bound = rel * np.maximum(np.abs(want), 0.05 * scale)
    print("max err / scale", float(np.max(err) / scale), "worst ratio",
          float(np.max(err / bound)))
    assert np.all(err <= bound)
    assert np.linalg.norm(got - want) <= 4e-3 * np.linalg.norm(want)


def test_forward_pass_matches_numpy(emu):
    rng = np.random.default_rng(11)
    S = 2
    fx, fy, _, _ = constants()
    v = bf16(rng.standard_normal((S, G, G, C))).astype(np.float64)
    got = emu.fno_forward(v, fx, fy)
    # A[s][kx'][y][c] = sum_x Fx[kx'][x] v[s][y][x][c], stored in BF16
    A = bf16(np.einsum("kx,syxc->skyc", fx, v)).astype(np.float64)
    A2 = A.reshape(S, KX, 2, G, C).reshape(S, KX, 2 * G, C)  # m = ri G + y
    want = bf16(np.einsum("pm,skmc->skpc", fy, A2))
    close(got, want)


@pytest.mark.parametrize("act", [0, 1], ids=["linear", "gelu"])
def test_inverse_pass_matches_numpy(emu, act):
    rng = np.random.default_rng(12 + act)
    S = 2
    _, _, gy, gx = constants()
    U = bf16(rng.standard_normal((S, KX, 2 * KY, C))).astype(np.float64)  # [s][kx][k'][c]
    v = bf16(rng.standard_normal((S, G, G, C))).astype(np.float64)
    W = bf16(rng.standard_normal((C, C)) / math.sqrt(C)).astype(np.float64)
    bias = rng.standard_normal(C).astype(np.float32)
    got = emu.fno_inverse(U, gy, gx, v, W, bias, act)
    # Bi[s][kx'=2kx+ro][y][c] = sum_k' Gy2[ro G + y][k'] U[s][kx][k'][c], in BF16
    bi = np.einsum("mk,sxkc->sxmc", gy, U)  # [s][kx][m = ro G + y][c]
    bi = bf16(bi.reshape(S, KX, 2, G, C).reshape(S, 2 * KX, G, C)).astype(np.float64)
    pre = (np.einsum("xk,skyc->syxc", gx, bi) + np.einsum("syxd,cd->syxc", v, W)
           + bias.astype(np.float64))
    want = bf16(gelu(pre) if act else pre)
    close(got, want)


def lift_case(rng, S, kf=4):
    """Inputs of layer 0's lifted passes and the lifted input v in numpy: the
    grid feature f0 per point (BF16), per-sample viscosities and forcing
    amplitudes whose features f1 = log nu - log 1e-3 and f2 = (amp / 0.1)
    sin(2 pi kf y / G) are rounded to FP32 then BF16 (the feature pass's
    rounding), lift weights, and v = bf16(f0 Lw0 + (f2 Lw2 + (f1 Lw1 + Lb))),
    each fma rounded to FP32."""
    f0 = bf16(rng.standard_normal((S, G, G))).astype(np.float64)
    nu = np.exp(rng.uniform(np.log(5e-4), np.log(2e-3), S))
    amp = rng.uniform(0.05, 0.4, S)
    f1 = bf16(np.float32(np.log(nu) - np.log(1e-3))).astype(np.float64)
    y = np.arange(G)
    f2 = bf16((amp[:, None] / 0.1 * np.sin(2.0 * np.pi * kf * y[None, :] / G)).astype(
        np.float32)).astype(np.float64)
    feat = np.stack([f0, np.broadcast_to(f1[:, None, None], f0.shape),
                     np.broadcast_to(f2[:, :, None], f0.shape)], -1)
    Lw = bf16(rng.standard_normal((C, 3))).astype(np.float64)
    Lb = rng.standard_normal(C).astype(np.float32)
    f32 = lambda x: x.astype(np.float32).astype(np.float64)
    k = f32(f2[:, :, None] * Lw[:, 2] + f32(f1[:, None, None] * Lw[:, 1] + Lb))  # [S][y][c]
    v = bf16(f32(f0[..., None] * Lw[:, 0] + k[:, :, None, :])).astype(np.float64)
    # channels 1 and 2 of feat are not read by the passes: poison them
    feat = feat.copy()
    feat[..., 1:] = np.nan
    return (feat, Lw, Lb, nu, amp, kf), v


def test_lifted_forward_pass_matches_numpy(emu):
    """Layer 0's forward pass lifts its input from the features inside the
    pass (the lift weights and the row constants in LDS): the DFTs of the
    numpy-lifted input."""
    rng = np.random.default_rng(21)
    S = 2
    fx, fy, _, _ = constants()
    lift, v = lift_case(rng, S)
    got = emu.fno_forward(np.zeros((S, G, G, C)), fx, fy, lift=lift)
    A = bf16(np.einsum("kx,syxc->skyc", fx, v)).astype(np.float64)
    A2 = A.reshape(S, KX, 2, G, C).reshape(S, KX, 2 * G, C)
    want = bf16(np.einsum("pm,skmc->skpc", fy, A2))
    close(got, want)


def test_lifted_inverse_pass_matches_numpy(emu):
    """Layer 0's inverse pass with the pointwise term on the lifted input:
    rows of all four row blocks, GELU."""
    rng = np.random.default_rng(22)
    S = 2
    _, _, gy, gx = constants()
    lift, v = lift_case(rng, S)
    U = bf16(rng.standard_normal((S, KX, 2 * KY, C))).astype(np.float64)
    W = bf16(rng.standard_normal((C, C)) / math.sqrt(C)).astype(np.float64)
    bias = rng.standard_normal(C).astype(np.float32)
    got = emu.fno_inverse(U, gy, gx, np.zeros((S, G, G, C)), W, bias, 1, lift=lift)
    bi = np.einsum("mk,sxkc->sxmc", gy, U)
    bi = bf16(bi.reshape(S, KX, 2, G, C).reshape(S, 2 * KX, G, C)).astype(np.float64)
    pre = (np.einsum("xk,skyc->syxc", gx, bi) + np.einsum("syxd,cd->syxc", v, W)
           + bias.astype(np.float64))
    close(got, bf16(gelu(pre)))


def test_head_matches_numpy(emu):
    """Projection C -> 128 -> 1 with GELU, 75 points (a ragged last wavefront)."""
    rng = np.random.default_rng(14)
    M, H = 75, 128
    X = bf16(rng.standard_normal((M, C))).astype(np.float64)
    Q1 = bf16(rng.standard_normal((H, C)) / math.sqrt(C)).astype(np.float64)
    q1 = rng.standard_normal(H).astype(np.float32)
    Q2 = (rng.standard_normal(H) / math.sqrt(H)).astype(np.float32)
    q2 = np.float32(0.3)
    got = emu.fno_head(X, Q1, q1, Q2, q2)
    want = gelu(X @ Q1.T + q1.astype(np.float64)) @ Q2.astype(np.float64) + float(q2)
    # the inference GELU is within 2.6e-5 of the exact one: each output within
    # sum |Q2| 2.6e-5 plus FP32 rounding of the 128-term sum
    err = np.abs(got - want)
    bound = 2.6e-5 * np.sum(np.abs(Q2)) + 1e-5 * np.max(np.abs(want))
    print("head max err", float(np.max(err)), "bound", bound)
    assert np.all(err <= bound)


def test_gelu_sg_matches_exact_erf(emu):
    """The inference GELU x sigma(z(x)) (fitted odd quintic z, argument clamped
    to [-8, 8]) within 2.6e-5 of x Phi(x) over [-9, 9] in FP32 and beyond the
    clamp (|x| up to 1e4, where it is x or 0); exact at 0 and odd-symmetric
    in the sense gelu(x) - gelu(-x) = x."""
    x = np.concatenate([np.linspace(-9.0, 9.0, 200001),
                        [-1e4, -300.0, -20.0, 20.0, 300.0, 1e4]]).astype(np.float32)
    got = emu.gelu_sg(x).astype(np.float64)
    want = gelu(x.astype(np.float64))
    err = np.abs(got - want)
    print("max gelu err", float(np.max(err)))
    assert np.all(err <= 2.6e-5 + 1e-7 * np.abs(want))
    assert emu.gelu_sg(np.zeros(1, np.float32))[0] == 0.0
    xs = np.linspace(0.0, 6.0, 61).astype(np.float32)
    d = emu.gelu_sg(xs).astype(np.float64) - emu.gelu_sg(-xs).astype(np.float64)
    assert np.all(np.abs(d - xs) <= 1e-6 * np.maximum(1.0, xs))


@pytest.mark.parametrize("K", [42, 85], ids=["state128", "state256"])
def test_state_to_grid_matches_numpy(emu, K):
    """The fused state -> grid pass is the product pair of the generic path:
    Bm = g1 st_band^T (band kx <= 31, |ky| <= 31 of the retained half plane),
    out = alpha Bm2 g2^T with Bm2[y][ro 32 + kx] = Bm[2y + ro][kx]."""
    rng = np.random.default_rng(15 + K)
    S, J, KS, JS, KXS = 2, 2 * K + 1, 31, 63, 32
    states = rng.standard_normal((S, K + 1, J, 2)).astype(np.float32)
    g1 = rng.standard_normal((2 * G, 2 * JS)).astype(np.float32)
    g2 = rng.standard_normal((G, 2 * KXS)).astype(np.float32)
    got = emu.state_to_grid(K, states.reshape(S, -1), g1, g2, 0.5)
    band = states[:, :KS + 1, K - KS:K + KS + 1, :].reshape(S, KS + 1, 2 * JS).astype(np.float64)
    bm = np.einsum("rk,sxk->srx", g1.astype(np.float64), band)          # [s][2y+ro][kx]
    bm2 = bm.reshape(S, G, 2 * KXS)                                      # [s][y][ro 32 + kx]
    want = 0.5 * np.einsum("syk,xk->syx", bm2, g2.astype(np.float64))
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print("state->grid rel err", err)
    assert err < 2e-5


def test_grid_to_increment_matches_numpy(emu):
    rng = np.random.default_rng(17)
    S, JS = 2, 63
    grid = rng.standard_normal((S, G, G)).astype(np.float32)
    f1 = rng.standard_normal((64, G)).astype(np.float32)
    f2 = rng.standard_normal((2 * JS, 2 * G)).astype(np.float32)
    got = emu.grid_to_increment(grid, f1, f2)
    dg = np.einsum("rx,syx->sry", f1.astype(np.float64), grid.astype(np.float64))  # [s][2kx+ri][y]
    dg2 = dg.reshape(S, 32, 2 * G)
    want = np.einsum("skm,nm->skn", dg2, f2.astype(np.float64))
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print("grid->increment rel err", err)
    assert err < 2e-5


def test_mixing_pass_matches_numpy(emu):
    """Per-mode mixing: 40 samples (a ragged second sample tile), 3 modes."""
    rng = np.random.default_rng(19)
    S, modes = 40, 3
    V = bf16(rng.standard_normal((S, modes, 2 * C))).astype(np.float64)
    R = bf16(rng.standard_normal((modes, 2 * C, 2 * C)) / 8).astype(np.float64)
    got = emu.fno_mix(V, R)
    want = bf16(np.einsum("smk,mnk->smn", V, R))
    close(got, want)
