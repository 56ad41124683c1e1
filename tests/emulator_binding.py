"""ctypes binding of the CPU lane emulator (tests/emulator/emu.cpp), TEST ONLY."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_SO = os.path.join(HERE, "emulator", "libhzl_emu.so")
_dp = C.POINTER(C.c_double)


class Emulator:
    def __init__(self, path: str = EMU_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run __graft_entry__.build()")
        L = self.lib = C.CDLL(path)
        L.emu_run.restype = C.c_int
        L.emu_run.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_double,
                              C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _dp]
        _fp, _ip, _lp = C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_long)
        L.emu_gemm.restype = C.c_int
        L.emu_gemm.argtypes = [C.c_int, _ip, _lp, _fp, C.c_size_t, _fp, C.c_size_t, _fp,
                               C.c_size_t, _fp, _fp, C.c_size_t, _fp, C.c_float, C.c_float,
                               C.c_int, C.c_int]
        L.emu_fno_forward.restype = C.c_int
        L.emu_fno_forward.argtypes = [C.c_long, _fp, _fp, _fp, _fp, _fp, _fp, _dp, _dp,
                                      C.c_int, _fp]
        L.emu_fno_inverse.restype = C.c_int
        L.emu_fno_inverse.argtypes = [C.c_long, _fp, _fp, _fp, _fp, _fp, _fp, C.c_int, _fp,
                                      _fp, _fp, _dp, _dp, C.c_int, _fp]
        L.emu_fno_head.restype = C.c_int
        L.emu_fno_head.argtypes = [C.c_long, _fp, _fp, _fp, _fp, _fp, _fp]
        L.emu_state_to_grid.restype = C.c_int
        L.emu_state_to_grid.argtypes = [C.c_long, C.c_int, _fp, _fp, _fp, C.c_float, _fp]
        L.emu_grid_to_increment.restype = C.c_int
        L.emu_grid_to_increment.argtypes = [C.c_long, _fp, _fp, _fp, _fp]
        L.emu_fno_mix.restype = C.c_int
        L.emu_fno_mix.argtypes = [C.c_long, C.c_int, _fp, _fp, _fp]
        L.emu_gelu_sg.restype = C.c_int
        L.emu_gelu_sg.argtypes = [C.c_long, _fp, _fp]
        L.emu_proj_head.restype = C.c_int
        L.emu_proj_head.argtypes = [C.c_int, C.c_long, C.c_int, C.c_int, _fp, _fp, _fp, _fp, _fp,
                                    _fp]
        L.emu_run_variant.restype = C.c_int
        L.emu_run_variant.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_int, C.c_uint64, C.c_double, C.c_int, C.c_int, _dp,
                                      _dp, _dp]
        L.emu_calibration.restype = C.c_int
        L.emu_calibration.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_int, C.c_uint64, C.c_double, C.c_int, _dp, _dp, _dp]
        L.emu_ou.restype = C.c_int
        L.emu_ou.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_uint64), _dp,
                             _dp, C.c_int, C.POINTER(C.c_int), _dp, _dp, _dp]
        L.emu_normal_stream.restype = C.c_int
        L.emu_normal_stream.argtypes = [C.c_uint64, C.c_int, C.c_int, C.POINTER(C.c_long), _dp]
        L.emu_fft_lines.restype = C.c_int
        L.emu_fft_lines.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.emu_fft_lines_split.restype = C.c_int
        L.emu_fft_lines_split.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp]

    def gemm(self, bf16, dims, strides, A, B, C0, bias=None, aux=None, want_pre=False,
             alpha=1.0, beta=0.0, act=0, c_bf16=False):
        """gemm_wave on the CPU lanes.  dims (M, N, K, nb1, nb2); strides the 16
        element strides (a_m, a_k, a_b1, a_b2, b_n, b_k, b_b1, b_b2, c_m, c_n,
        c_b1, c_b2, x_m, x_n, x_b1, x_b2).  Returns (C, pre)."""
        fp = C.POINTER(C.c_float)
        f = lambda x: None if x is None else np.ascontiguousarray(x, dtype=np.float32).ravel()
        A, B, aux, bias = f(A), f(B), f(aux), f(bias)
        out = np.ascontiguousarray(C0, dtype=np.float32).ravel().copy()
        pre = np.zeros_like(out) if want_pre else None
        d = np.array(dims, dtype=np.int32)
        st = np.array(strides, dtype=np.int64)
        ptr = lambda x: None if x is None else x.ctypes.data_as(fp)
        rc = self.lib.emu_gemm(int(bf16), d.ctypes.data_as(C.POINTER(C.c_int)),
                               st.ctypes.data_as(C.POINTER(C.c_long)), ptr(A), A.size, ptr(B),
                               B.size, ptr(out), out.size, ptr(bias), ptr(aux),
                               0 if aux is None else aux.size, ptr(pre), alpha, beta, act,
                               int(c_bf16))
        assert rc == 0
        return out, pre

    @staticmethod
    def _f32(x):
        return np.ascontiguousarray(x, dtype=np.float32)

    def _lift(self, lift):
        """lift = (feat, Lw, Lb, nu, amp, kf): the arrays kept alive and the
        argument list (float32 / float64 pointers, kf), or Nones."""
        fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
        if lift is None:
            return (), (None, None, None, None, None, 0)
        feat, Lw, Lb, nu, amp, kf = lift
        f = tuple(self._f32(x) for x in (feat, Lw, Lb))
        d = tuple(np.ascontiguousarray(x, dtype=np.float64) for x in (nu, amp))
        return f + d, tuple(x.ctypes.data_as(fp) for x in f) + tuple(
            x.ctypes.data_as(dp) for x in d) + (int(kf),)

    def fno_forward(self, v, Fx, Fy2, lift=None):
        """fno_forward_body over v [S][64][64][32]: V [S][16][64][32]; with
        lift = (feat [S][64][64][3], Lw [32][3], Lb [32]) the lifted layer-0
        pass (v only gives S; nu, amp [S] and kf give features 1 and 2)."""
        v, Fx, Fy2 = self._f32(v), self._f32(Fx), self._f32(Fy2)
        S = v.shape[0]
        out = np.zeros((S, 16, 64, 32), dtype=np.float32)
        fp = C.POINTER(C.c_float)
        keep, lp = self._lift(lift)
        assert self.lib.emu_fno_forward(S, v.ctypes.data_as(fp), Fx.ctypes.data_as(fp),
                                        Fy2.ctypes.data_as(fp), *lp,
                                        out.ctypes.data_as(fp)) == 0
        return out

    def fno_head(self, X, Q1, q1, Q2, q2):
        """fno_head_wave over X [M][32], Q1 [128][32]: out [M]."""
        X, Q1 = self._f32(X), self._f32(Q1)
        q1, Q2, q2 = self._f32(q1), self._f32(Q2), self._f32(np.atleast_1d(q2))
        out = np.zeros(X.shape[0], dtype=np.float32)
        fp = C.POINTER(C.c_float)
        assert self.lib.emu_fno_head(X.shape[0], X.ctypes.data_as(fp), Q1.ctypes.data_as(fp),
                                     q1.ctypes.data_as(fp), Q2.ctypes.data_as(fp),
                                     q2.ctypes.data_as(fp), out.ctypes.data_as(fp)) == 0
        return out

    def state_to_grid(self, K, states, g1, g2, alpha):
        """fno_state_to_grid_body: states [S][2 (K+1)(2K+1)] FP32 -> [S][64][64]."""
        st, g1, g2 = self._f32(states), self._f32(g1), self._f32(g2)
        S = st.shape[0]
        out = np.zeros((S, 64, 64), dtype=np.float32)
        fp = C.POINTER(C.c_float)
        assert self.lib.emu_state_to_grid(S, K, st.ctypes.data_as(fp), g1.ctypes.data_as(fp),
                                          g2.ctypes.data_as(fp), alpha,
                                          out.ctypes.data_as(fp)) == 0
        return out

    def grid_to_increment(self, grid, f1, f2):
        """fno_grid_to_increment_body: grid [S][64][64] -> incr [S][32][126]."""
        grid, f1, f2 = self._f32(grid), self._f32(f1), self._f32(f2)
        S = grid.shape[0]
        out = np.zeros((S, 32, 126), dtype=np.float32)
        fp = C.POINTER(C.c_float)
        assert self.lib.emu_grid_to_increment(S, grid.ctypes.data_as(fp), f1.ctypes.data_as(fp),
                                              f2.ctypes.data_as(fp),
                                              out.ctypes.data_as(fp)) == 0
        return out

    def fno_mix(self, V, R):
        """fno_mix_body: V [S][modes][64], R [modes][64][64] -> U [S][modes][64]."""
        V, R = self._f32(V), self._f32(R)
        S, modes = V.shape[0], V.shape[1]
        U = np.zeros_like(V)
        fp = C.POINTER(C.c_float)
        assert self.lib.emu_fno_mix(S, modes, V.ctypes.data_as(fp), R.ctypes.data_as(fp),
                                    U.ctypes.data_as(fp)) == 0
        return U

    def gelu_sg(self, x):
        x = self._f32(x)
        y = np.zeros_like(x)
        fp = C.POINTER(C.c_float)
        assert self.lib.emu_gelu_sg(x.size, x.ctypes.data_as(fp), y.ctypes.data_as(fp)) == 0
        return y

    def fno_inverse(self, U, Gy2, Gx, v, W, bias, act, lift=None):
        """fno_inverse_body: U [S][16][64][32], v [S][64][64][32] -> [S][64][64][32];
        lift as in fno_forward."""
        U, Gy2, Gx, v = self._f32(U), self._f32(Gy2), self._f32(Gx), self._f32(v)
        W, bias = self._f32(W), self._f32(bias)
        S = v.shape[0]
        out = np.zeros_like(v)
        fp = C.POINTER(C.c_float)
        keep, lp = self._lift(lift)
        assert self.lib.emu_fno_inverse(S, U.ctypes.data_as(fp), Gy2.ctypes.data_as(fp),
                                        Gx.ctypes.data_as(fp), v.ctypes.data_as(fp),
                                        W.ctypes.data_as(fp), bias.ctypes.data_as(fp), int(act),
                                        *lp, out.ctypes.data_as(fp)) == 0
        return out

    def proj_head(self, bf16, X, Q1, q1, Q2, q2):
        """proj_head_wave on the CPU lanes: X [M][C], Q1 [H][C]; returns out [M]."""
        fp = C.POINTER(C.c_float)
        f = lambda x: np.ascontiguousarray(x, dtype=np.float32)
        X, Q1, q1, Q2, q2 = f(X), f(Q1), f(q1), f(Q2), f(np.atleast_1d(q2))
        M, Cc = X.shape
        out = np.zeros(M, dtype=np.float32)
        rc = self.lib.emu_proj_head(int(bf16), M, Cc, Q1.shape[0], X.ctypes.data_as(fp),
                                    Q1.ctypes.data_as(fp), q1.ctypes.data_as(fp),
                                    Q2.ctypes.data_as(fp), q2.ctypes.data_as(fp),
                                    out.ctypes.data_as(fp))
        assert rc == 0
        return out

    def normal_stream(self, seed, counts, lanes=64):
        """normal_stream_body: the normals of successive calls of counts[c] each."""
        cnt = np.ascontiguousarray(counts, dtype=np.int64)
        out = np.zeros(int(cnt.sum()))
        rc = self.lib.emu_normal_stream(seed, lanes, len(cnt), cnt.ctypes.data_as(C.POINTER(C.c_long)),
                                        out.ctypes.data_as(_dp))
        assert rc == 0
        return out

    def ou(self, npairs, pair_map, keys, sigma, tau, modes, hs):
        """ou_body over members for a sequence of calls (sigma and tau: one
        value per member, or one tau for all); keys is (calls, B)
        (mode 0 / 2 seeds per call).  Returns (F, noise), each
        (calls, B, npairs) complex."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        calls, B = keys.shape
        F = np.zeros((calls, B, npairs), dtype=np.complex128)
        noise = np.zeros_like(F)
        m = np.ascontiguousarray(pair_map, dtype=np.int32)
        mo = np.ascontiguousarray(modes, dtype=np.int32)
        h = np.ascontiguousarray(hs, dtype=np.float64)
        sg = np.ascontiguousarray(sigma, dtype=np.float64)
        tu = np.ascontiguousarray(np.broadcast_to(np.asarray(tau, dtype=np.float64), (B,)))
        rc = self.lib.emu_ou(B, npairs, m.ctypes.data_as(C.POINTER(C.c_int)),
                             keys.ctypes.data_as(C.POINTER(C.c_uint64)), sg.ctypes.data_as(_dp),
                             tu.ctypes.data_as(_dp), calls, mo.ctypes.data_as(C.POINTER(C.c_int)),
                             h.ctypes.data_as(_dp), F.ctypes.data_as(_dp),
                             noise.ctypes.data_as(_dp))
        assert rc == 0
        return F, noise

    def fft_lines(self, x, inverse=False, precision="fp32", elems=None):
        """GroupFFT over each row of x; elems: values per lane (default: the
        grid's shape)."""
        d = np.ascontiguousarray(x, dtype=np.complex128).copy()
        lines, n = d.shape
        p = 0 if precision == "fp32" else 1
        if elems is None:
            rc = self.lib.emu_fft_lines(n, p, lines, int(inverse), d.ctypes.data_as(_dp))
        else:
            rc = self.lib.emu_fft_lines_split(n, elems, p, lines, int(inverse),
                                              d.ctypes.data_as(_dp))
        assert rc == 0, "no such split"
        return d

    def run(self, n, members, precision, dt, steps, nu=1e-3, alpha=0.0, sigma=0.05, amp=0.1,
            kf=4, kmin=1, kmax=4, seed=0, init=None, forcing=None, shared=True):
        """Replay the engine's launch sequence on the CPU; returns (states, diagnostics)."""
        out = np.zeros((members, n, n), dtype=np.complex128)
        diag = np.zeros((members, 6))
        ia = None if init is None else np.ascontiguousarray(init, dtype=np.complex128)
        fa = None if forcing is None else np.ascontiguousarray(forcing, dtype=np.complex128)
        rc = self.lib.emu_run(n, 0 if precision == "fp32" else 1, members, nu, alpha, sigma, amp,
                              kf, kmin, kmax, seed, dt, steps, 1 if ia is None else 0,
                              None if ia is None else ia.ctypes.data_as(_dp),
                              None if fa is None else fa.ctypes.data_as(_dp), int(shared),
                              out.ctypes.data_as(_dp), diag.ctypes.data_as(_dp))
        assert rc == 0
        return out, diag

    def run_variant(self, n, members, col_variant, row_variant, dt, steps, nu=1e-3, alpha=0.0,
                    sigma=0.05, amp=0.1, kf=4, kmin=1, kmax=4, seed=0):
        """FP32 launch-shape variant (one of col / row nonzero) from GRF fields."""
        out = np.zeros((members, n, n), dtype=np.complex128)
        diag = np.zeros((members, 6))
        rc = self.lib.emu_run_variant(n, col_variant, row_variant, members, nu, alpha, sigma,
                                      amp, kf, kmin, kmax, seed, dt, steps, 1, None,
                                      out.ctypes.data_as(_dp), diag.ctypes.data_as(_dp))
        assert rc == 0, "no such variant"
        return out, diag

    def calibration(self, kind, state, aux=None, nu=0.01, alpha=0.1, precision="fp64", dt=0.0,
                    steps=0, sigma=0.0, amp=0.0, kf=4, kmin=1, kmax=4, seed=0):
        """kind "drift", "tangent" (along aux), "adjoint" (of aux) or "ssprk3"
        (`steps` steps of dt) for members state[m] (reference Fields)."""
        code = {"drift": 0, "tangent": 1, "adjoint": 2, "ssprk3": 3}[kind]
        st = np.ascontiguousarray(state, dtype=np.complex128)
        members, n = st.shape[0], st.shape[-1]
        ax = None if aux is None else np.ascontiguousarray(aux, dtype=np.complex128)
        out = np.zeros_like(st)
        rc = self.lib.emu_calibration(n, 0 if precision == "fp32" else 1, code, members, nu,
                                      alpha, sigma, amp, kf, kmin, kmax, seed, dt, steps,
                                      st.ctypes.data_as(_dp),
                                      None if ax is None else ax.ctypes.data_as(_dp),
                                      out.ctypes.data_as(_dp))
        assert rc == 0
        return out
