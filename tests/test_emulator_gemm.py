"""CPU checks of the MFMA product body (csrc/gemm_mfma.hpp) through the lane
emulator: operand staging (16-byte loads along k, along rows, and the strided
scalar path), ragged edges, two-level batches, and the epilogue (alpha, beta,
bias, aux, pre-activation store, GELU, BF16 output), against numpy in FP64 on
the same (BF16-rounded) operands.

The emulated instruction follows the lane layout that was measured on gfx950
for v_mfma_f32_32x32x16_bf16 / v_mfma_f32_32x32x2_f32; the same products run
on the GPU in tests/test_gpu_surrogate.py.
"""
from __future__ import annotations

import math
import zlib

import numpy as np
import pytest

ERF = np.vectorize(math.erf)


def bf16(x):
    """Round FP32 to BF16 (nearest even), returned as FP32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def gelu(x):
    return 0.5 * x * (1.0 + ERF(x / math.sqrt(2.0)))


def extent(b1, s1, b2, s2, r, sr, k, sk):
    return (b1 - 1) * s1 + (b2 - 1) * s2 + (r - 1) * sr + (k - 1) * sk + 1


def gather(flat, b1, s1, b2, s2, r, sr, k, sk):
    i = (np.arange(b1)[:, None, None, None] * s1 + np.arange(b2)[None, :, None, None] * s2
         + np.arange(r)[None, None, :, None] * sr + np.arange(k)[None, None, None, :] * sk)
    return flat[i]


def reference(dims, st, A, B, C0, bias, aux, alpha, beta, act):
    M, N, K, n1, n2 = dims
    a = gather(A.astype(np.float64), n1, st[2], n2, st[3], M, st[0], K, st[1])
    b = gather(B.astype(np.float64), n1, st[6], n2, st[7], N, st[4], K, st[5])
    acc = alpha * np.einsum("xymk,xynk->xymn", a, b)
    c_idx = (np.arange(n1)[:, None, None, None] * st[10] + np.arange(n2)[None, :, None, None] * st[11]
             + np.arange(M)[None, None, :, None] * st[8] + np.arange(N)[None, None, None, :] * st[9])
    v = acc + beta * C0.astype(np.float64)[c_idx]
    if bias is not None:
        v = v + bias.astype(np.float64)[None, None, None, :]
    if aux is not None:
        v = v + gather(aux.astype(np.float64), n1, st[14], n2, st[15], M, st[12], N, st[13])
    pre = v.copy()
    if act:
        v = gelu(v)
    return c_idx, v, pre


CASES = {
    # name: (dims, strides a_m a_k a_b1 a_b2 | b_n b_k b_b1 b_b2 | c_m c_n c_b1 c_b2 | x..)
    "k_contiguous_ragged": ((70, 45, 80, 2, 1),
                            (80, 1, 70 * 80, 0, 80, 1, 45 * 80, 0, 45, 1, 70 * 45, 0, 0, 0, 0, 0)),
    "row_contiguous": ((64, 32, 40, 1, 1),
                       (1, 64, 0, 0, 1, 32, 0, 0, 32, 1, 0, 0, 0, 0, 0, 0)),
    "scalar_strided": ((33, 31, 17, 1, 2),
                       (3, 101, 0, 3 * 33 + 101 * 17, 5, 1, 0, 7, 1, 40, 0, 40 * 31, 0, 0, 0, 0)),
    "two_level_batch": ((32, 64, 64, 3, 2),
                        (64, 1, 2 * 64 * 64, 64 * 64, 1, 64, 64 * 64, 0, 64, 1, 2 * 32 * 64,
                         32 * 64, 64, 1, 2 * 32 * 64, 32 * 64)),
}


@pytest.mark.parametrize("bf", [False, True], ids=["f32", "bf16"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_gemm_wave_matches_numpy(emu, case, bf):
    dims, st = CASES[case]
    M, N, K, n1, n2 = dims
    rng = np.random.default_rng(zlib.crc32(case.encode()))
    na = extent(n1, st[2], n2, st[3], M, st[0], K, st[1])
    nb = extent(n1, st[6], n2, st[7], N, st[4], K, st[5])
    nc = extent(n1, st[10], n2, st[11], M, st[8], N, st[9])
    A = rng.standard_normal(na).astype(np.float32)
    B = rng.standard_normal(nb).astype(np.float32)
    C0 = rng.standard_normal(nc).astype(np.float32)
    use_epi = case == "two_level_batch"
    bias = rng.standard_normal(N).astype(np.float32) if use_epi else None
    aux = None
    if use_epi:
        aux = rng.standard_normal(extent(n1, st[14], n2, st[15], M, st[12], N, st[13]))
        aux = aux.astype(np.float32)
    alpha, beta, act = (0.75, 0.5, 1) if use_epi else (1.0, 0.0, 0)
    c_bf16 = bf and use_epi
    got, pre = emu.gemm(bf, dims, st, A, B, bf16(C0) if c_bf16 else C0, bias=bias, aux=aux,
                        want_pre=use_epi, alpha=alpha, beta=beta, act=act, c_bf16=c_bf16)
    r = (lambda x: bf16(x)) if bf else (lambda x: x)
    c_idx, want, want_pre = reference(dims, st, r(A), r(B), bf16(C0) if c_bf16 else C0, bias,
                                      None if aux is None else r(aux), alpha, beta, act)
    # FP32 accumulation over K terms of unit-variance products; BF16 output adds
    # a relative 2^-8 rounding
    tol = 1e-5 * math.sqrt(K) + (8e-3 * np.abs(want) if c_bf16 else 0)
    assert np.all(np.abs(got[c_idx] - want) <= tol)
    if use_epi:
        ptol = 1e-5 * math.sqrt(K) + (8e-3 * np.abs(want_pre) if c_bf16 else 0)
        assert np.all(np.abs(pre[c_idx] - want_pre) <= ptol)
    # entries outside the addressed tiles are untouched
    mask = np.ones(nc, dtype=bool)
    mask[c_idx.ravel()] = False
    np.testing.assert_array_equal(got[mask], (bf16(C0) if c_bf16 else C0)[mask])


@pytest.mark.parametrize("bf", [False, True], ids=["f32", "bf16"])
@pytest.mark.parametrize("M,C", [(96, 32), (45, 64), (32, 8)])
def test_fused_projection_head_matches_numpy(emu, bf, M, C):
    """out[m] = q2 + sum_j Q2[j] gelu(q1[j] + X[m] . Q1[j]) with the hidden layer
    kept in registers and summed across lanes (ragged last wave included)."""
    H = 128
    rng = np.random.default_rng(M * 100 + C)
    X = rng.standard_normal((M, C)).astype(np.float32)
    Q1 = (rng.standard_normal((H, C)) / np.sqrt(C)).astype(np.float32)
    q1 = rng.standard_normal(H).astype(np.float32)
    Q2 = (rng.standard_normal(H) / np.sqrt(H)).astype(np.float32)
    q2 = np.float32(0.25)
    got = emu.proj_head(bf, X, Q1, q1, Q2, q2)
    r = bf16 if bf else (lambda x: x)
    hid = r(X).astype(np.float64) @ r(Q1).astype(np.float64).T + q1
    want = gelu(hid) @ Q2 + q2
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5 * np.sqrt(H))
