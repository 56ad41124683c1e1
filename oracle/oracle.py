"""ctypes bindings of the FP64 oracle (liboracle.so) and the reference build
(_ref/libhzlref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never by the product path.

Complex fields are numpy complex128 arrays of shape (n, n) indexed [y, x], the
reference's Field layout (include/hzl/fft.hpp:7-9).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhzlref.so")

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_ubyte)
_ip = C.POINTER(C.c_int)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


class Oracle:
    """FP64 C restatement of the reference algorithms (oracle/hzl_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.lib = C.CDLL(path)
        L.orc_wave_number.restype = C.c_int
        L.orc_wave_number.argtypes = [C.c_long, C.c_long]
        L.orc_fft2d.restype = C.c_int
        L.orc_fft2d.argtypes = [C.c_long, _dp, C.c_int]
        L.orc_active_mask.argtypes = [C.c_long, _u8p]
        L.orc_rhs.restype = C.c_int
        L.orc_rhs.argtypes = [C.c_long, C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_ssprk3_steps.restype = C.c_int
        L.orc_ssprk3_steps.argtypes = [C.c_long, C.c_double, C.c_double, C.c_double, C.c_long,
                                       _dp, _dp, C.c_long]
        L.orc_linearized_drift.restype = C.c_int
        L.orc_linearized_drift.argtypes = [C.c_long, C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_adjoint_drift.restype = C.c_int
        L.orc_adjoint_drift.argtypes = [C.c_long, C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_rk4_steps.restype = C.c_int
        L.orc_rk4_steps.argtypes = [C.c_long, C.c_double, C.c_double, C.c_double, C.c_long, _dp,
                                    _dp, C.c_long]
        L.orc_diagnostics.restype = C.c_int
        L.orc_diagnostics.argtypes = [C.c_long, C.c_double, _dp, _dp, _dp]
        L.orc_energy_spectrum.argtypes = [C.c_long, _dp, _dp]
        L.orc_stream_seed.restype = C.c_uint64
        L.orc_stream_seed.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_mt_size.restype = C.c_size_t
        L.orc_mt_seed.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_mt_next.restype = C.c_uint64
        L.orc_mt_next.argtypes = [C.c_void_p]
        L.orc_normal.restype = C.c_double
        L.orc_normal.argtypes = [C.c_void_p]
        L.orc_mt_uniform.restype = C.c_double
        L.orc_mt_uniform.argtypes = [C.c_void_p]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32)]
        L.orc_philox_normal2.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, _dp]
        L.orc_shell_modes.restype = C.c_int
        L.orc_shell_modes.argtypes = [C.c_long, C.c_int, C.c_int, _ip, _ip]
        L.orc_step_forcing.restype = C.c_int
        L.orc_step_forcing.argtypes = [C.c_long, C.c_double, C.c_int, C.c_double, C.c_int,
                                       C.c_int, C.c_double, C.c_uint64, C.c_uint64, _dp]
        L.orc_ou_pairs.restype = C.c_int
        L.orc_ou_pairs.argtypes = [C.c_long, C.c_int, C.c_int, C.POINTER(C.c_long)]
        L.orc_ou_initialize.argtypes = [C.c_void_p, C.c_long, C.c_int, C.POINTER(C.c_long),
                                        C.c_double, C.c_double, _dp]
        L.orc_ou_transition.argtypes = [C.c_void_p, C.c_long, C.c_int, C.POINTER(C.c_long),
                                        C.c_double, C.c_double, C.c_double, _dp]
        L.orc_run_ou.restype = C.c_int
        L.orc_run_ou.argtypes = [C.c_long, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.c_int, C.c_uint64, C.c_double, C.c_long, _dp, _dp]
        L.orc_grf.restype = C.c_int
        L.orc_grf.argtypes = [C.c_long, C.c_uint64, _dp]
        L.orc_member_trajectory.restype = C.c_int
        L.orc_member_trajectory.argtypes = [C.c_long, C.c_double, C.c_double, C.c_double,
                                            C.c_long, C.c_uint64, C.c_double, C.c_int,
                                            C.c_double, C.c_int, C.c_int, C.c_uint64, C.c_int,
                                            _dp]
        L.orc_trajectories.restype = C.c_int
        L.orc_trajectories.argtypes = [C.c_long, C.c_double, C.c_double, C.c_double, C.c_long,
                                       C.c_double, C.c_int, C.c_double, C.c_int, C.c_int,
                                       C.c_uint64, C.c_long, C.c_long, C.c_int, _dp]
        L.orc_ensemble.restype = C.c_int
        L.orc_ensemble.argtypes = [C.c_long, C.c_double, C.c_double, C.c_double, C.c_long,
                                   C.c_double, C.c_int, C.c_double, C.c_int, C.c_int,
                                   C.c_uint64, C.c_long, C.c_int, _dp]
        L.orc_pf_logweight.restype = C.c_double
        L.orc_pf_logweight.argtypes = [C.c_double, C.c_long, _dp, _dp, _dp]
        L.orc_pf_normalize.restype = C.c_double
        L.orc_pf_normalize.argtypes = [C.c_long, _dp, _dp, _dp, _dp]
        L.orc_systematic_resample.argtypes = [C.c_long, _dp, C.c_double, _ip]
        L.orc_enkf_scalar_update.argtypes = [C.c_long, C.c_long, _dp, _dp, C.c_double,
                                             C.c_double, _dp]

    # -- transforms ---------------------------------------------------------
    def fft2d(self, field, inverse=False):
        a = _c(field).copy()
        n = a.shape[0]
        if self.lib.orc_fft2d(n, _ptr(a), int(inverse)):
            raise ValueError("grid must be a power of two between 8 and 4096")
        return a

    def wave_number(self, i, n):
        return self.lib.orc_wave_number(i, n)

    def active_mask(self, n):
        m = np.zeros((n, n), dtype=np.uint8)
        self.lib.orc_active_mask(n, m.ctypes.data_as(_u8p))
        return m.astype(bool)

    # -- solver -------------------------------------------------------------
    def rhs(self, state, nu, alpha, forcing=None):
        w = _c(state)
        n = w.shape[0]
        out = np.zeros_like(w)
        F = _c(forcing) if forcing is not None else None
        rc = self.lib.orc_rhs(n, nu, alpha, _ptr(w), _ptr(F) if F is not None else None,
                              _ptr(out))
        if rc:
            raise ValueError("bad grid")
        return out

    def ssprk3(self, state, nu, alpha, dt, steps, forcing=None):
        """`steps` SSPRK(3,3) steps with one held forcing field (src/solver.cpp:235-246)."""
        w = _c(state).copy()
        n = w.shape[0]
        F = _c(forcing) if forcing is not None else None
        if self.lib.orc_ssprk3_steps(n, nu, alpha, dt, steps, _ptr(w),
                                     _ptr(F) if F is not None else None, 0):
            raise ValueError("bad grid")
        return w

    def linearized_drift(self, state, direction, nu, alpha):
        """Solver::linearized_drift (src/solver.cpp:264-290)."""
        w, d = _c(state), _c(direction)
        out = np.zeros_like(w)
        if self.lib.orc_linearized_drift(w.shape[0], nu, alpha, _ptr(w), _ptr(d), _ptr(out)):
            raise ValueError("bad grid")
        return out

    def adjoint_drift(self, state, cotangent, nu, alpha):
        """Solver::adjoint_drift (src/solver.cpp:291-323)."""
        w, c = _c(state), _c(cotangent)
        out = np.zeros_like(w)
        if self.lib.orc_adjoint_drift(w.shape[0], nu, alpha, _ptr(w), _ptr(c), _ptr(out)):
            raise ValueError("bad grid")
        return out

    def rk4(self, state, nu, alpha, dt, steps, forcing=None, per_step=False):
        """`steps` RK4 steps; forcing is one (n,n) field held for every step or,
        with per_step, an array (steps, n, n)."""
        w = _c(state).copy()
        n = w.shape[0]
        F = _c(forcing) if forcing is not None else None
        stride = n * n if per_step else 0
        rc = self.lib.orc_rk4_steps(n, nu, alpha, dt, steps, _ptr(w),
                                    _ptr(F) if F is not None else None, stride)
        if rc:
            raise ValueError("bad grid")
        return w

    # -- OU shell forcing (src/solver.cpp:67-119) -----------------------------
    def ou_pairs(self, n, kmin, kmax):
        """Grid indices y*n + x of the first mode of every forcing pair, in the
        reference's order."""
        count = self.lib.orc_ou_pairs(n, kmin, kmax, None)
        first = (C.c_long * count)()
        self.lib.orc_ou_pairs(n, kmin, kmax, first)
        return np.array(first[:], dtype=np.int64)

    class OuProcess:
        """One member's OU forcing: mt19937_64(seed) NormalStream and the
        n x n forcing field; the constructor draws the stationary start."""

        def __init__(self, orc, n, sigma, tau, kmin, kmax, seed):
            self.orc, self.n, self.sigma, self.tau = orc, n, sigma, tau
            first = orc.ou_pairs(n, kmin, kmax)
            self.np = len(first)
            self._first = (C.c_long * self.np)(*first.tolist())
            self.state = orc.mt(seed)
            self.F = np.zeros((n, n), dtype=np.complex128)
            orc.lib.orc_ou_initialize(self.state, n, self.np, self._first, sigma, tau,
                                      _ptr(self.F))

        def transition(self, dt):
            self.orc.lib.orc_ou_transition(self.state, self.n, self.np, self._first, self.sigma,
                                           self.tau, dt, _ptr(self.F))

        def reseed(self, seed):
            """Solver::reseed_future: a new stream, the coefficients kept."""
            self.state = self.orc.mt(seed)

    def ou_process(self, n, sigma, tau, kmin, kmax, seed):
        return Oracle.OuProcess(self, n, sigma, tau, kmin, kmax, seed)

    def run_ou(self, state, nu, alpha, sigma, tau, kmin, kmax, seed, dt, steps):
        """Solver(n, p, seed) with OU forcing from `state`: `steps` RK4 steps;
        returns (state, forcing)."""
        w = _c(state).copy()
        n = w.shape[0]
        F = np.zeros_like(w)
        if self.lib.orc_run_ou(n, nu, alpha, sigma, tau, kmin, kmax, seed, dt, steps, _ptr(w),
                               _ptr(F)):
            raise ValueError("bad grid or tau")
        return w, F

    def diagnostics(self, state, nu, forcing=None):
        w = _c(state)
        out = np.zeros(5)
        F = _c(forcing) if forcing is not None else None
        self.lib.orc_diagnostics(w.shape[0], nu, _ptr(w), _ptr(F) if F is not None else None,
                                 _ptr(out))
        return dict(zip(("energy", "enstrophy", "dissipation", "injection", "cfl_rate"), out))

    def energy_spectrum(self, state):
        w = _c(state)
        bins = np.zeros(w.shape[0])
        self.lib.orc_energy_spectrum(w.shape[0], _ptr(w), _ptr(bins))
        return bins

    # -- streams ------------------------------------------------------------
    def stream_seed(self, base, index):
        return self.lib.orc_stream_seed(base, index)

    def mt(self, seed):
        buf = C.create_string_buffer(self.lib.orc_mt_size())
        self.lib.orc_mt_seed(buf, seed)
        return buf

    def mt_next(self, state):
        return self.lib.orc_mt_next(state)

    def normal(self, state):
        return self.lib.orc_normal(state)

    def mt_uniform(self, state):
        return self.lib.orc_mt_uniform(state)

    def philox(self, ctr, key):
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        o = (C.c_uint32 * 4)()
        self.lib.orc_philox4x32_10(c, k, o)
        return list(o)

    def philox_normal2(self, key, c0, c12, tag):
        g = np.zeros(2)
        self.lib.orc_philox_normal2(key, c0, c12, tag, _ptr(g))
        return g

    def shell_modes(self, n, kmin, kmax):
        count = self.lib.orc_shell_modes(n, kmin, kmax, None, None)
        kx = np.zeros(count, dtype=np.int32)
        ky = np.zeros(count, dtype=np.int32)
        self.lib.orc_shell_modes(n, kmin, kmax, kx.ctypes.data_as(_ip), ky.ctypes.data_as(_ip))
        return kx, ky

    def step_forcing(self, n, amp, kf, sigma, kmin, kmax, dt, member_key, step):
        F = np.zeros((n, n), dtype=np.complex128)
        self.lib.orc_step_forcing(n, amp, kf, sigma, kmin, kmax, dt, member_key, step, _ptr(F))
        return F

    def grf(self, n, member_key):
        w = np.zeros((n, n), dtype=np.complex128)
        if self.lib.orc_grf(n, member_key, _ptr(w)):
            raise ValueError("bad grid")
        return w

    def trajectory(self, n, nu, alpha, dt, steps, member_key, amp=0.1, kf=4, sigma=0.05,
                   kmin=1, kmax=4, step0=0, initial=None):
        if initial is None:
            w = np.zeros((n, n), dtype=np.complex128)
            grf = 1
        else:
            w = _c(initial).copy()
            grf = 0
        rc = self.lib.orc_member_trajectory(n, nu, alpha, dt, steps, step0, amp, kf, sigma, kmin,
                                            kmax, member_key, grf, _ptr(w))
        if rc:
            raise ValueError("bad grid")
        return w

    def trajectories(self, n, nu, alpha, dt, steps, seed, first, count, threads=None,
                     amp=0.1, kf=4, sigma=0.05, kmin=1, kmax=4):
        """Final FP64 states of members first..first+count-1 (GRF start)."""
        threads = threads or min(count, os.cpu_count() or 1, 16)
        out = np.zeros((count, n, n), dtype=np.complex128)
        rc = self.lib.orc_trajectories(n, nu, alpha, dt, steps, amp, kf, sigma, kmin, kmax, seed,
                                       first, count, threads, _ptr(out))
        if rc:
            raise ValueError("bad grid")
        return out

    def ensemble(self, n, nu, alpha, dt, steps, members, threads, seed, amp=0.1, kf=4,
                 sigma=0.05, kmin=1, kmax=4):
        e = np.zeros(members)
        rc = self.lib.orc_ensemble(n, nu, alpha, dt, steps, amp, kf, sigma, kmin, kmax, seed,
                                   members, threads, _ptr(e))
        if rc:
            raise ValueError("ensemble failed")
        return e

    # -- filters ------------------------------------------------------------
    def pf_logweight(self, prev_weight, pred, value, sd):
        pred = np.ascontiguousarray(pred, dtype=np.float64)
        value = np.ascontiguousarray(value, dtype=np.float64)
        sd = np.ascontiguousarray(sd, dtype=np.float64)
        return self.lib.orc_pf_logweight(prev_weight, len(pred), _ptr(pred), _ptr(value),
                                         _ptr(sd))

    def pf_normalize(self, logs):
        logs = np.ascontiguousarray(logs, dtype=np.float64)
        w = np.zeros_like(logs)
        ess = np.zeros(1)
        mw = np.zeros(1)
        inc = self.lib.orc_pf_normalize(len(logs), _ptr(logs), _ptr(w), _ptr(ess), _ptr(mw))
        return inc, w, float(ess[0]), float(mw[0])

    def systematic_resample(self, weights, u):
        w = np.ascontiguousarray(weights, dtype=np.float64)
        a = np.zeros(len(w), dtype=np.int32)
        self.lib.orc_systematic_resample(len(w), _ptr(w), u, a.ctypes.data_as(_ip))
        return a

    def enkf_scalar_update(self, states, h, y, sd, eps):
        x = np.ascontiguousarray(states, dtype=np.complex128).copy()
        h = np.ascontiguousarray(h, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        m = x.shape[0]
        self.lib.orc_enkf_scalar_update(m, x.size // m, _ptr(x), _ptr(h), y, sd, _ptr(eps))
        return x


class Reference:
    """The unmodified reference CPU library (oracle/_ref/libhzlref.so)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref`")
        L = self.lib = C.CDLL(path)
        L.ref_last_error.restype = C.c_char_p
        L.ref_fft2d.restype = C.c_int
        L.ref_fft2d.argtypes = [C.c_size_t, _dp, C.c_int]
        L.ref_wave_number.restype = C.c_int
        L.ref_wave_number.argtypes = [C.c_size_t, C.c_size_t]
        L.ref_stream_seed.restype = C.c_uint64
        L.ref_stream_seed.argtypes = [C.c_uint64, C.c_uint64]
        L.ref_normals.argtypes = [C.c_uint64, C.c_size_t, _dp]
        L.ref_run.restype = C.c_int
        L.ref_run.argtypes = [C.c_size_t, C.c_double, C.c_double, C.c_double, C.c_size_t,
                              C.c_int, C.c_int, _dp, _dp, C.c_size_t, _dp]
        L.ref_initialize_physical.restype = C.c_int
        L.ref_initialize_physical.argtypes = [C.c_size_t, _dp, _dp]
        L.ref_run_ou.restype = C.c_int
        L.ref_run_ou.argtypes = [C.c_size_t, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.c_int, C.c_uint64, C.c_double, C.c_size_t, _dp, _dp,
                                 _dp]
        L.ref_initialize.restype = C.c_int
        L.ref_initialize.argtypes = [C.c_size_t, C.c_char_p, _dp]
        L.ref_drift.restype = C.c_int
        L.ref_drift.argtypes = [C.c_size_t, C.c_double, C.c_double, _dp, _dp]
        L.ref_diagnostics.restype = C.c_int
        L.ref_diagnostics.argtypes = [C.c_size_t, C.c_double, C.c_double, C.c_int, C.c_int, _dp,
                                      _dp, _dp]
        L.ref_energy_spectrum.restype = C.c_int
        L.ref_energy_spectrum.argtypes = [C.c_size_t, _dp, _dp]
        L.ref_ensemble.restype = C.c_int
        L.ref_ensemble.argtypes = [C.c_size_t, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_int, C.c_int, C.c_double, C.c_size_t, C.c_size_t,
                                   C.c_size_t, C.c_uint64, _dp]
        L.ref_synthetic.restype = C.c_void_p
        L.ref_synthetic.argtypes = [C.c_size_t, C.c_double, C.c_size_t, C.c_size_t, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                    C.c_double, C.c_uint64]
        L.ref_dataset_free.argtypes = [C.c_void_p]
        L.ref_dataset_size.restype = C.c_size_t
        L.ref_dataset_size.argtypes = [C.c_void_p]
        L.ref_dataset_rows.argtypes = [C.c_void_p, _dp]
        L.ref_bootstrap_filter.restype = C.c_int
        L.ref_bootstrap_filter.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_size_t, C.c_uint64, _dp, _dp,
                                           C.c_size_t, _dp]
        L.ref_negative_log_likelihood.restype = C.c_int
        L.ref_negative_log_likelihood.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
        L.ref_dataset_save.restype = C.c_int
        L.ref_dataset_save.argtypes = [C.c_void_p, C.c_char_p]
        L.ref_enkf.restype = C.c_int
        L.ref_enkf.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_size_t, C.c_uint64, C.c_double, C.c_double, _dp, C.c_size_t,
                               _dp, C.POINTER(C.c_size_t)]

    def _check(self, rc):
        if rc:
            raise RuntimeError(self.lib.ref_last_error().decode())

    def fft2d(self, field, inverse=False):
        a = _c(field).copy()
        self._check(self.lib.ref_fft2d(a.shape[0], _ptr(a), int(inverse)))
        return a

    def stream_seed(self, base, index):
        return self.lib.ref_stream_seed(base, index)

    def normals(self, seed, count):
        out = np.zeros(count)
        self.lib.ref_normals(seed, count, _ptr(out))
        return out

    def run_ou(self, initial, nu, alpha, sigma, tau, kmin, kmax, seed, dt, steps):
        """hzl::Solver(n, p, seed) with OU forcing (sigma > 0) from `initial`
        (its constructor's forcing kept): `steps` steps; (state, forcing)."""
        w = _c(initial)
        n = w.shape[0]
        out_w, out_F = np.zeros_like(w), np.zeros_like(w)
        self._check(self.lib.ref_run_ou(n, nu, alpha, sigma, tau, kmin, kmax, seed, dt, steps,
                                        _ptr(w), _ptr(out_w), _ptr(out_F)))
        return out_w, out_F

    def run(self, initial, nu, alpha, dt, steps, forcing=None, per_step=False, kmin=1, kmax=4):
        w = _c(initial)
        n = w.shape[0]
        out = np.zeros_like(w)
        F = _c(forcing) if forcing is not None else None
        self._check(self.lib.ref_run(n, nu, alpha, dt, steps, kmin, kmax, _ptr(w),
                                     _ptr(F) if F is not None else None,
                                     n * n if per_step else 0, _ptr(out)))
        return out

    def initialize(self, n, name):
        out = np.zeros((n, n), dtype=np.complex128)
        self._check(self.lib.ref_initialize(n, name.encode(), _ptr(out)))
        return out

    def run_ssprk3(self, initial, nu, alpha, dt, steps, forcing=None, kmin=1, kmax=4):
        """Reference Solver::step(dt, TimeIntegrator::ssprk3) with a held forcing."""
        L = self.lib
        L.ref_run_ssprk3.restype = C.c_int
        L.ref_run_ssprk3.argtypes = [C.c_size_t, C.c_double, C.c_double, C.c_double, C.c_size_t,
                                     C.c_int, C.c_int, _dp, _dp, _dp]
        w = _c(initial)
        out = np.zeros_like(w)
        F = _c(forcing) if forcing is not None else None
        self._check(L.ref_run_ssprk3(w.shape[0], nu, alpha, dt, steps, kmin, kmax, _ptr(w),
                                     _ptr(F) if F is not None else None, _ptr(out)))
        return out

    def linearized_drift(self, state, direction, nu, alpha):
        """Reference Solver::linearized_drift."""
        L = self.lib
        L.ref_linearized_drift.restype = C.c_int
        L.ref_linearized_drift.argtypes = [C.c_size_t, C.c_double, C.c_double, _dp, _dp, _dp]
        w, d = _c(state), _c(direction)
        out = np.zeros_like(w)
        self._check(L.ref_linearized_drift(w.shape[0], nu, alpha, _ptr(w), _ptr(d), _ptr(out)))
        return out

    def adjoint_drift(self, state, cotangent, nu, alpha):
        """Reference Solver::adjoint_drift."""
        L = self.lib
        L.ref_adjoint_drift.restype = C.c_int
        L.ref_adjoint_drift.argtypes = [C.c_size_t, C.c_double, C.c_double, _dp, _dp, _dp]
        w, c = _c(state), _c(cotangent)
        out = np.zeros_like(w)
        self._check(L.ref_adjoint_drift(w.shape[0], nu, alpha, _ptr(w), _ptr(c), _ptr(out)))
        return out

    def initialize_physical(self, physical):
        p = np.ascontiguousarray(physical, dtype=np.float64)
        out = np.zeros(p.shape, dtype=np.complex128)
        self._check(self.lib.ref_initialize_physical(p.shape[0], _ptr(p), _ptr(out)))
        return out

    def drift(self, state, nu, alpha):
        w = _c(state)
        out = np.zeros_like(w)
        self._check(self.lib.ref_drift(w.shape[0], nu, alpha, _ptr(w), _ptr(out)))
        return out

    def diagnostics(self, state, nu, alpha=0.0, forcing=None, kmin=1, kmax=4):
        w = _c(state)
        out = np.zeros(5)
        F = _c(forcing) if forcing is not None else None
        self._check(self.lib.ref_diagnostics(w.shape[0], nu, alpha, kmin, kmax, _ptr(w),
                                             _ptr(F) if F is not None else None, _ptr(out)))
        return dict(zip(("energy", "enstrophy", "dissipation", "injection", "cfl_rate"), out))

    def energy_spectrum(self, state):
        w = _c(state)
        bins = np.zeros(w.shape[0])
        self._check(self.lib.ref_energy_spectrum(w.shape[0], _ptr(w), _ptr(bins)))
        return bins

    def ensemble(self, n, nu, alpha, sigma, tau, kmin, kmax, dt, steps, count, workers, seed):
        e = np.zeros(count)
        self._check(self.lib.ref_ensemble(n, nu, alpha, sigma, tau, kmin, kmax, dt, steps, count,
                                          workers, seed, _ptr(e)))
        return e

    def synthetic(self, n, dt, steps, every, nu, alpha, sigma, tau, kmin, kmax, noise, seed):
        h = self.lib.ref_synthetic(n, dt, steps, every, nu, alpha, sigma, tau, kmin, kmax, noise,
                                   seed)
        if not h:
            raise RuntimeError(self.lib.ref_last_error().decode())
        rows = np.zeros((self.lib.ref_dataset_size(h), 5))
        self.lib.ref_dataset_rows(h, _ptr(rows))
        return h, rows

    def dataset_free(self, h):
        self.lib.ref_dataset_free(h)

    def bootstrap_filter(self, h, nu, alpha, sigma, tau, particles, seed, means=False, n=16,
                         max_records=4096):
        summary = np.zeros(3)
        records = np.zeros((max_records, 3))
        mean = np.zeros((max_records, n, n)) if means else None
        self._check(self.lib.ref_bootstrap_filter(h, nu, alpha, sigma, tau, particles, seed,
                                                  _ptr(summary), _ptr(records), max_records,
                                                  _ptr(mean) if means else None))
        k = int(summary[2])
        return dict(log_likelihood=summary[0], resampling_events=int(summary[1]),
                    records=records[:k], means=mean[:k] if means else None)

    def particle_metropolis(self, h, q0, scale, count, particles, seed):
        """hzl::particle_metropolis: rows (q[4], log_likelihood, log_posterior,
        likelihood_seed, accepted)."""
        L = self.lib
        L.ref_particle_metropolis.restype = C.c_int
        L.ref_particle_metropolis.argtypes = [C.c_void_p, _dp, C.c_double, C.c_size_t,
                                              C.c_size_t, C.c_uint64, _dp]
        q = np.ascontiguousarray(q0, dtype=np.float64)
        out = np.zeros((count, 8))
        self._check(L.ref_particle_metropolis(h, _ptr(q), scale, count, particles, seed,
                                              _ptr(out)))
        seeds = out[:, 6].copy().view(np.uint64)
        return [dict(q=out[i, :4].copy(), log_likelihood=out[i, 4], log_posterior=out[i, 5],
                     likelihood_seed=int(seeds[i]), accepted=bool(out[i, 7]))
                for i in range(count)]

    def checkpoint_ou(self, n, nu, alpha, sigma, tau, kmin, kmax, seed, dt, before, after,
                      path_before, path_after):
        """Reference Solver with OU forcing from multimode: checkpoints after
        `before` and after `before + after` steps."""
        L = self.lib
        L.ref_checkpoint_ou.restype = C.c_int
        L.ref_checkpoint_ou.argtypes = [C.c_size_t, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int, C.c_int, C.c_uint64, C.c_double,
                                        C.c_size_t, C.c_size_t, C.c_char_p, C.c_char_p]
        self._check(L.ref_checkpoint_ou(n, nu, alpha, sigma, tau, kmin, kmax, seed, dt, before,
                                        after, str(path_before).encode(),
                                        str(path_after).encode()))

    def negative_log_likelihood(self, h, log_nu, log_alpha):
        out = np.zeros(1)
        self._check(self.lib.ref_negative_log_likelihood(h, log_nu, log_alpha, _ptr(out)))
        return float(out[0])

    def likelihood_gradient(self, h, log_nu, log_alpha):
        """hzl::likelihood_gradient: (value, gradient[2])."""
        L = self.lib
        L.ref_likelihood_gradient.restype = C.c_int
        L.ref_likelihood_gradient.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
        out = np.zeros(3)
        self._check(L.ref_likelihood_gradient(h, log_nu, log_alpha, _ptr(out)))
        return float(out[0]), out[1:3].copy()

    def step_tangent(self, state, tangents, nu, alpha, dt, steps):
        """`steps` reference Solver::step_tangent steps: (state, tangents[2])."""
        L = self.lib
        L.ref_step_tangent.restype = C.c_int
        L.ref_step_tangent.argtypes = [C.c_size_t, C.c_double, C.c_double, C.c_double,
                                       C.c_size_t, _dp, _dp, _dp, _dp]
        w = _c(state)
        t = np.ascontiguousarray(tangents, dtype=np.complex128)
        wo, to = np.zeros_like(w), np.zeros_like(t)
        self._check(L.ref_step_tangent(w.shape[0], nu, alpha, dt, steps, _ptr(w), _ptr(t),
                                       _ptr(wo), _ptr(to)))
        return wo, to

    def likelihood_adjoint(self, h, log_nu, log_alpha, stride):
        """hzl::likelihood_adjoint: dict(value, gradient, checkpoint_fields,
        peak_block_fields, replayed_steps)."""
        L = self.lib
        L.ref_likelihood_adjoint.restype = C.c_int
        L.ref_likelihood_adjoint.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_size_t, _dp]
        out = np.zeros(6)
        self._check(L.ref_likelihood_adjoint(h, log_nu, log_alpha, stride, _ptr(out)))
        return dict(value=float(out[0]), gradient=out[1:3].copy(),
                    checkpoint_fields=int(out[3]), peak_block_fields=int(out[4]),
                    replayed_steps=int(out[5]))

    def dataset_save(self, h, path):
        self._check(self.lib.ref_dataset_save(h, str(path).encode()))

    def checkpoint_write(self, path, n, nu, alpha, sigma, tau, kmin, kmax, seed, initial, dt,
                         steps):
        """Solver(n, p, seed).initialize(initial), `steps` steps, Solver::save."""
        L = self.lib
        L.ref_checkpoint_write.restype = C.c_int
        L.ref_checkpoint_write.argtypes = [C.c_size_t, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_int, C.c_int, C.c_uint64,
                                           C.c_char_p, C.c_double, C.c_size_t, C.c_char_p]
        self._check(L.ref_checkpoint_write(n, nu, alpha, sigma, tau, kmin, kmax, seed,
                                           initial.encode(), dt, steps, str(path).encode()))

    def checkpoint_fields(self, path, n):
        """Solver::load(path): (coefficients(), forcing()) as n x n arrays."""
        L = self.lib
        L.ref_checkpoint_fields.restype = C.c_int
        L.ref_checkpoint_fields.argtypes = [C.c_char_p, _dp, _dp]
        w = np.zeros((n, n), dtype=np.complex128)
        f = np.zeros((n, n), dtype=np.complex128)
        self._check(L.ref_checkpoint_fields(str(path).encode(), _ptr(w), _ptr(f)))
        return w, f

    def checkpoint_diagnostics(self, path):
        """Solver::load(path).diagnostics(): dict(time, energy, enstrophy,
        dissipation, injection, cfl_rate)."""
        L = self.lib
        L.ref_checkpoint_diagnostics.restype = C.c_int
        L.ref_checkpoint_diagnostics.argtypes = [C.c_char_p, _dp]
        out = np.zeros(6)
        self._check(L.ref_checkpoint_diagnostics(str(path).encode(), _ptr(out)))
        return dict(zip(("time", "energy", "enstrophy", "dissipation", "injection",
                         "cfl_rate"), map(float, out)))

    def enkf(self, h, nu, alpha, sigma, tau, members, seed, inflation=1.0, radius=0.0, n=16,
             max_records=4096):
        records = np.zeros((max_records, 3))
        means = np.zeros((max_records, n, n))
        count = C.c_size_t(0)
        self._check(self.lib.ref_enkf(h, nu, alpha, sigma, tau, members, seed, inflation, radius,
                                      _ptr(records), max_records, _ptr(means), C.byref(count)))
        k = count.value
        return dict(records=records[:k], means=means[:k])
