"""FP64 CPU reference of the forecast surrogate network (TEST ONLY).

A plain PyTorch (float64, CPU, autograd) statement of the network that
hzl-stochasticflowinference_amd/csrc/surrogate.hip evaluates and trains with
its own MFMA kernels, used to check the GPU forward output and the GPU
gradients.  The transforms are written as explicit DFT sums; the tests pin
them against torch.fft and check the autograd gradients by finite differences.

Parameter vector layout (FP32 on the device, same order here):
  Lw [C][3], Lb [C], then per layer W [C][C], b [C], Rr [KX][KY][C][C],
  Ri [KX][KY][C][C], then Q1 [H][C], q1 [H], Q2 [H], q2 [1].
Spectral modes: kx in [0, KX), ky = kyi for kyi < KY/2 else kyi - KY.
"""
from __future__ import annotations

import math

import numpy as np
import torch

CIN, HIDDEN, KF = 3, 128, 4
AMP_REF = 0.1  # the forcing channel is (amplitude / AMP_REF) sin(KF y)


class Shape:
    def __init__(self, state_n=128, grid=64, width=32, layers=4, modes=16):
        self.n, self.G, self.C, self.L, self.KX = state_n, grid, width, layers, modes
        self.KY, self.H = 2 * modes, HIDDEN
        self.K = (state_n - 1) // 3
        self.KS = min(grid // 2 - 1, self.K)

    def sizes(self):
        C, m = self.C, self.KX * self.KY
        out = [("Lw", (C, CIN)), ("Lb", (C,))]
        for l in range(self.L):
            out += [(f"W{l}", (C, C)), (f"b{l}", (C,)), (f"Rr{l}", (self.KX, self.KY, C, C)),
                    (f"Ri{l}", (self.KX, self.KY, C, C))]
        out += [("Q1", (self.H, C)), ("q1", (self.H,)), ("Q2", (self.H,)), ("q2", (1,))]
        return out

    def count(self):
        return sum(int(np.prod(s)) for _, s in self.sizes())

    def unpack(self, flat):
        out, o = {}, 0
        for name, s in self.sizes():
            k = int(np.prod(s))
            out[name] = flat[o:o + k].reshape(s)
            o += k
        return out


def ky_of(KY):
    return np.array([k if k < KY // 2 else k - KY for k in range(KY)])


def sample_grid(shape: Shape, fields):
    """Grid values of states given as full n x n spectra (reference Field
    layout [ky mod n][kx mod n]): sum over kx in [0, KS] of w_kx Re(sum over
    |ky| <= KS of coef(kx, ky) e^{i(kx x + ky y)}), w_0 = 1, w_kx = 2; the mean
    mode is dropped.  Equals the physical field on the grid for real fields."""
    G, KS, n = shape.G, shape.KS, shape.n
    f = torch.as_tensor(np.asarray(fields), dtype=torch.complex128)
    kx = np.arange(KS + 1)
    ky = np.arange(-KS, KS + 1)
    band = f[:, ky % n][:, :, kx].clone()  # [s][ky][kx]
    band[:, KS, 0] = 0
    g = torch.arange(G, dtype=torch.float64)
    ey = torch.exp(2j * math.pi * torch.as_tensor(ky, dtype=torch.float64)[:, None] * g[None] / G)
    ex = torch.exp(2j * math.pi * torch.as_tensor(kx, dtype=torch.float64)[:, None] * g[None] / G)
    w = torch.where(torch.as_tensor(kx) == 0, 1.0, 2.0).to(torch.float64)
    z = torch.einsum("sab,ay,bx,b->syx", band, ey.to(torch.complex128), ex.to(torch.complex128),
                     w.to(torch.complex128))
    return z.real


def _bases(shape: Shape):
    G, KX, KY = shape.G, shape.KX, shape.KY
    g = torch.arange(G, dtype=torch.float64)
    kx = torch.arange(KX, dtype=torch.float64)
    ky = torch.as_tensor(ky_of(KY), dtype=torch.float64)
    fx = torch.exp(-2j * math.pi * kx[:, None] * g[None] / G)  # [kx][x]
    fy = torch.exp(-2j * math.pi * ky[:, None] * g[None] / G)  # [ky][y]
    return fx, fy


def dft_modes(shape: Shape, v):
    """X[s][kx][kyi][c] = sum_{x,y} v[s][y][x][c] e^{-2 pi i (kx x + ky y)/G}."""
    fx, fy = _bases(shape)
    a = torch.einsum("syxc,ax->syac", v.to(torch.complex128), fx)
    return torch.einsum("syac,by->sabc", a, fy)


def idft_modes(shape: Shape, U):
    """Real field [s][y][x][c] = sum_kx w_kx / G^2 Re(sum_kyi U e^{+i(kx x + ky y)}),
    w_0 = 1, w_kx = 2 (the kept half plane of a real field)."""
    fx, fy = _bases(shape)
    G, KX = shape.G, shape.KX
    w = torch.where(torch.arange(KX) == 0, 1.0, 2.0).to(torch.complex128) / (G * G)
    b = torch.einsum("sabc,by->syac", U, fy.conj())
    return torch.einsum("syac,ax,a->syxc", b, fx.conj(), w).real


def spectral(shape: Shape, v, Rr, Ri):
    """Truncated DFT -> per-mode complex mixing R[kx][ky][co][ci] -> inverse."""
    U = torch.einsum("abij,sabj->sabi", torch.complex(Rr, Ri), dft_modes(shape, v))
    return idft_modes(shape, U)


def forward(shape: Shape, params, fields, nu, in_scale, amp=None):
    """Network output [s][G][G] (float64 tensor); params: flat float64 tensor;
    amp: per-sample Kolmogorov amplitudes (default AMP_REF)."""
    p = shape.unpack(params)
    G = shape.G
    S = len(nu)
    w = sample_grid(shape, fields) * in_scale
    lognu = torch.as_tensor(np.log(np.asarray(nu)) - math.log(1e-3), dtype=torch.float64)
    a = torch.ones(S, dtype=torch.float64) if amp is None else \
        torch.as_tensor(np.asarray(amp, dtype=np.float64) / AMP_REF)
    y = torch.arange(G, dtype=torch.float64)
    pattern = torch.sin(2 * math.pi * KF * y / G)[None, :, None] * a[:, None, None]
    feat = torch.stack([w, lognu[:, None, None].expand(S, G, G), pattern.expand(S, G, G)], -1)
    v = feat @ p["Lw"].T + p["Lb"]
    for l in range(shape.L):
        z = spectral(shape, v, p[f"Rr{l}"], p[f"Ri{l}"]) + v @ p[f"W{l}"].T + p[f"b{l}"]
        v = torch.nn.functional.gelu(z) if l + 1 < shape.L else z
    h = torch.nn.functional.gelu(v @ p["Q1"].T + p["q1"])
    return h @ p["Q2"] + p["q2"][0]


def loss(shape: Shape, params, fields, next_fields, nu, in_scale, out_scale, amp=None):
    out = forward(shape, params, fields, nu, in_scale, amp)
    target = sample_grid(shape, np.asarray(next_fields) - np.asarray(fields)) * out_scale
    return torch.mean((out - target) ** 2)


def loss_and_grad(shape: Shape, params_np, fields, next_fields, nu, in_scale, out_scale,
                  amp=None):
    p = torch.tensor(np.asarray(params_np, dtype=np.float64), requires_grad=True)
    l = loss(shape, p, fields, next_fields, nu, in_scale, out_scale, amp)
    l.backward()
    return float(l.detach()), p.grad.numpy().copy()


def band_increment(shape: Shape, grid):
    """Band modes (kx in [0, KS], |ky| <= KS) of grid values [s][G][G] as the
    forward DFT / G^2, returned as full n x n spectra with the kx > 0 mirror."""
    G, KS, n = shape.G, shape.KS, shape.n
    F = np.fft.fft2(np.asarray(grid)) / (G * G)  # [s][ky][kx]
    out = np.zeros((F.shape[0], n, n), dtype=np.complex128)
    for kx in range(KS + 1):
        for ky in range(-KS, KS + 1):
            if kx == 0 and ky == 0:
                continue
            out[:, ky % n, kx] = F[:, ky % G, kx]
            if kx > 0:
                out[:, (-ky) % n, (-kx) % n] = np.conj(F[:, ky % G, kx])
    return out
