"""Python binding of the C-ABI (include/hzl_amd.h) over ctypes.

Mirrors the reference's solver interfaces for tests and benchmarks:
  Ensemble      -- a batch of hzl::Solver / hzl::CudaSolver members
                   (include/hzl/solver.hpp:61-117, include/hzl/cuda_solver.hpp:7-22)
There is no CPU fallback: constructing an Ensemble without the built HIP
library, or without a GPU, raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libhzl_amd.so")

HZL_OK, HZL_EINVAL, HZL_ERUNTIME = 0, 1, 2
FP32, FP64, BF16 = 0, 1, 2
ENKF_SERIAL, ENKF_BATCH = 0, 1
ALL = C.c_size_t(-1).value  # (size_t)-1: every member

# Every entry point declared in include/hzl_amd.h.
EXPORTS = (
    "hzl_last_error", "hzl_ensemble_create", "hzl_ensemble_destroy",
    "hzl_ensemble_set_member_parameters", "hzl_ensemble_set_state", "hzl_ensemble_get_state",
    "hzl_ensemble_init_grf", "hzl_ensemble_step", "hzl_ensemble_set_graph",
    "hzl_ensemble_step_integrator", "hzl_ensemble_set_forcing_process",
    "hzl_ensemble_get_forcing", "hzl_ensemble_reseed_future", "hzl_ensemble_set_forcing_state",
    "hzl_ensemble_get_forcing_state", "hzl_ensemble_drift", "hzl_ensemble_linearized_drift",
    "hzl_ensemble_adjoint_drift", "hzl_likelihood_adjoint", "hzl_ensemble_step_tangent",
    "hzl_likelihood_gradient",
    "hzl_ensemble_variant_count", "hzl_ensemble_set_tuning", "hzl_ensemble_get_tuning",
    "hzl_ensemble_autotune",
    "hzl_ensemble_step_forced",
    "hzl_ensemble_synchronize", "hzl_ensemble_diagnostics", "hzl_ensemble_profile_step",
    "hzl_ensemble_step_index",
    "hzl_ensemble_device_state", "hzl_ensemble_stream",
    "hzl_pf_begin", "hzl_pf_assimilate", "hzl_pf_result",
    "hzl_pf_begin_filters", "hzl_ensemble_fill_state", "hzl_ensemble_set_step_index",
    "hzl_enkf_begin", "hzl_enkf_assimilate", "hzl_ensemble_mean_spread", "hzl_pmcmc_run",
    "hzl_parameters_validate", "hzl_stream_seed", "hzl_ensemble_set_member_tau",
    "hzl_ensemble_energy_spectrum", "hzl_normal_draws", "hzl_stochastic_log_prior",
    "hzl_pf_log_likelihoods",
    "hzl_ensemble_diagnostics_noalloc",
    "hzl_pf_joint_create", "hzl_pf_joint_destroy", "hzl_pf_joint_begin",
    "hzl_pf_joint_assimilate", "hzl_pf_joint_result",
    "hzl_surrogate_create", "hzl_surrogate_destroy", "hzl_surrogate_param_count",
    "hzl_surrogate_get_params", "hzl_surrogate_set_params", "hzl_surrogate_set_scales",
    "hzl_surrogate_get_scales", "hzl_surrogate_forward", "hzl_surrogate_loss_grad",
    "hzl_surrogate_collect", "hzl_surrogate_dataset_size", "hzl_surrogate_fit",
    "hzl_surrogate_evaluate", "hzl_surrogate_error_estimate",
    "hzl_surrogate_set_error_estimate", "hzl_surrogate_spread",
    "hzl_surrogate_forecast_or_solve", "hzl_surrogate_set_graph", "hzl_surrogate_set_fused",
    "hzl_surrogate_forecast",
    "hzl_surrogate_judge",
    "hzl_surrogate_judge_start",
    "hzl_surrogate_judge_poll",
)


class Parameters(C.Structure):
    """hzl_parameters; first six fields mirror hzl::Parameters (include/hzl/solver.hpp:15-23)."""
    _fields_ = [("viscosity", C.c_double), ("friction", C.c_double),
                ("forcing_sigma", C.c_double), ("forcing_tau", C.c_double),
                ("forcing_min", C.c_int), ("forcing_max", C.c_int),
                ("kolmogorov_amplitude", C.c_double), ("kolmogorov_wavenumber", C.c_int)]


PRIOR_LOG_UNIFORM, PRIOR_STOCHASTIC = 0, 1


class PmcmcOptions(C.Structure):
    """hzl_pmcmc_options (include/hzl_amd.h)."""
    _fields_ = [("chains", C.c_size_t), ("samples", C.c_size_t),
                ("proposal_scale", C.c_double), ("log_lower", C.c_double * 2),
                ("log_upper", C.c_double * 2), ("friction", C.c_double),
                ("forcing_sigma", C.c_double), ("seed", C.c_uint64), ("prior", C.c_int),
                ("chain_seeds", C.POINTER(C.c_uint64)), ("surrogate", C.c_void_p),
                ("surrogate_precision", C.c_int)]


class HzlError(RuntimeError):
    pass


class HzlInvalidArgument(HzlError, ValueError):
    """std::invalid_argument of the reference."""


_lib = None
_dp = C.POINTER(C.c_double)


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libhzl_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} not built: run `python hzl-stochasticflowinference_amd/build.py engine`")
    L = C.CDLL(path)
    L.hzl_last_error.restype = C.c_char_p
    vp = C.c_void_p
    L.hzl_ensemble_create.argtypes = [C.POINTER(vp), C.c_size_t, C.c_size_t, C.c_int,
                                      C.POINTER(Parameters), C.c_uint64]
    L.hzl_ensemble_destroy.argtypes = [vp]
    L.hzl_ensemble_set_member_parameters.argtypes = [vp, C.c_size_t, C.c_double, C.c_double,
                                                     C.c_double, C.c_double]
    L.hzl_ensemble_set_state.argtypes = [vp, C.c_size_t, _dp]
    L.hzl_ensemble_get_state.argtypes = [vp, C.c_size_t, _dp]
    L.hzl_ensemble_init_grf.argtypes = [vp]
    L.hzl_ensemble_step.argtypes = [vp, C.c_double, C.c_size_t]
    ipp = C.POINTER(C.c_int)
    L.hzl_ensemble_set_graph.argtypes = [vp, C.c_int]
    L.hzl_ensemble_step_integrator.argtypes = [vp, C.c_double, C.c_size_t, C.c_int]
    L.hzl_ensemble_set_forcing_process.argtypes = [vp, C.c_int]
    L.hzl_ensemble_get_forcing.argtypes = [vp, C.c_size_t, _dp]
    L.hzl_ensemble_reseed_future.argtypes = [vp, C.c_size_t, C.c_uint64]
    u64p = C.POINTER(C.c_uint64)
    L.hzl_ensemble_set_forcing_state.argtypes = [vp, C.c_size_t, u64p, _dp]
    L.hzl_ensemble_get_forcing_state.argtypes = [vp, C.c_size_t, u64p, _dp]
    L.hzl_ensemble_drift.argtypes = [vp, _dp]
    L.hzl_ensemble_linearized_drift.argtypes = [vp, _dp, _dp]
    L.hzl_ensemble_adjoint_drift.argtypes = [vp, _dp, _dp]
    szp = C.POINTER(C.c_size_t)
    L.hzl_ensemble_step_tangent.argtypes = [vp, C.c_double, C.c_size_t, _dp, _dp]
    L.hzl_likelihood_gradient.argtypes = [vp, _dp, C.c_double, C.c_size_t, szp, ipp, ipp, _dp,
                                          _dp, _dp, _dp, _dp]
    L.hzl_likelihood_adjoint.argtypes = [vp, _dp, C.c_double, C.c_size_t, szp, ipp, ipp, _dp, _dp,
                                         _dp, C.c_size_t, _dp, _dp, szp]
    L.hzl_ensemble_variant_count.argtypes = [vp, ipp, ipp]
    L.hzl_ensemble_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    L.hzl_ensemble_get_tuning.argtypes = [vp, ipp, ipp]
    L.hzl_ensemble_autotune.argtypes = [vp, C.c_int, _dp, _dp]
    L.hzl_ensemble_step_forced.argtypes = [vp, C.c_double, _dp, C.c_int]
    L.hzl_ensemble_synchronize.argtypes = [vp]
    L.hzl_ensemble_diagnostics.argtypes = [vp, _dp]
    L.hzl_ensemble_profile_step.argtypes = [vp, C.c_double, C.POINTER(C.c_float)]
    L.hzl_ensemble_step_index.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.hzl_ensemble_device_state.argtypes = [vp, C.POINTER(vp), C.POINTER(C.c_size_t)]
    L.hzl_ensemble_stream.argtypes = [vp, C.POINTER(vp)]
    ip = C.POINTER(C.c_int)
    L.hzl_pf_begin.argtypes = [vp, C.c_uint64, C.c_size_t]
    L.hzl_pf_assimilate.argtypes = [vp, C.c_double, C.c_size_t, C.c_size_t, ip, ip, _dp, _dp,
                                    C.c_int, _dp, _dp]
    L.hzl_pf_result.argtypes = [vp, _dp, C.POINTER(C.c_size_t)]
    L.hzl_pf_begin_filters.argtypes = [vp, C.POINTER(C.c_uint64), C.c_size_t, C.c_size_t]
    L.hzl_ensemble_fill_state.argtypes = [vp, _dp]
    L.hzl_ensemble_set_step_index.argtypes = [vp, C.c_uint64]
    L.hzl_enkf_begin.argtypes = [vp, C.c_uint64, C.c_int]
    L.hzl_enkf_assimilate.argtypes = [vp, C.c_double, C.c_size_t, C.c_size_t, ip, ip, _dp, _dp,
                                      C.c_double, C.c_double, _dp, _dp]
    L.hzl_ensemble_mean_spread.argtypes = [vp, _dp, _dp]
    L.hzl_pf_joint_create.argtypes = [C.POINTER(vp), C.POINTER(vp), C.c_size_t]
    L.hzl_pf_joint_destroy.argtypes = [vp]
    L.hzl_pf_joint_begin.argtypes = [vp, C.c_uint64, C.c_size_t]
    L.hzl_pf_joint_assimilate.argtypes = [vp, C.c_double, C.c_size_t, C.c_size_t, ip, ip, _dp,
                                          _dp, C.c_int, _dp]
    L.hzl_pf_joint_result.argtypes = [vp, _dp, C.POINTER(C.c_size_t)]
    L.hzl_pmcmc_run.argtypes = [vp, C.POINTER(PmcmcOptions), C.c_double, C.c_size_t, ip, ip, ip,
                                _dp, _dp, _dp, _dp, _dp, _dp, _dp, ip, u64p]
    L.hzl_pf_log_likelihoods.argtypes = [vp, C.POINTER(PmcmcOptions), C.c_double, C.c_size_t,
                                         ip, ip, ip, _dp, _dp, _dp, C.c_size_t, _dp, u64p, _dp]
    L.hzl_parameters_validate.argtypes = [C.c_size_t, C.POINTER(Parameters)]
    L.hzl_stream_seed.restype = C.c_uint64
    L.hzl_stream_seed.argtypes = [C.c_uint64, C.c_uint64]
    L.hzl_ensemble_set_member_tau.argtypes = [vp, C.c_size_t, C.c_double]
    L.hzl_ensemble_energy_spectrum.argtypes = [vp, _dp]
    L.hzl_ensemble_diagnostics_noalloc.argtypes = [vp, _dp]
    L.hzl_normal_draws.argtypes = [C.c_uint64, C.c_size_t, _dp]
    L.hzl_stochastic_log_prior.argtypes = [_dp, C.c_size_t, _dp]
    fp, sz = C.POINTER(C.c_float), C.c_size_t
    L.hzl_surrogate_create.argtypes = [C.POINTER(vp), sz, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_uint64]
    L.hzl_surrogate_destroy.argtypes = [vp]
    L.hzl_surrogate_param_count.argtypes = [vp, C.POINTER(sz)]
    L.hzl_surrogate_get_params.argtypes = [vp, fp]
    L.hzl_surrogate_set_params.argtypes = [vp, fp]
    L.hzl_surrogate_set_scales.argtypes = [vp, C.c_double, C.c_double]
    L.hzl_surrogate_get_scales.argtypes = [vp, _dp]
    L.hzl_surrogate_forward.argtypes = [vp, C.c_int, sz, _dp, _dp, _dp, _dp]
    L.hzl_surrogate_loss_grad.argtypes = [vp, sz, _dp, _dp, _dp, _dp, _dp, fp]
    L.hzl_surrogate_judge.argtypes = [vp, vp, vp, C.c_double, sz, C.c_int, _dp]
    L.hzl_surrogate_judge_start.argtypes = [vp, vp, vp, C.c_double, sz, C.c_int]
    L.hzl_surrogate_judge_poll.argtypes = [vp, C.c_int, C.POINTER(C.c_int), _dp]
    L.hzl_surrogate_collect.argtypes = [vp, vp, C.c_double, sz, sz, C.c_int]
    L.hzl_surrogate_dataset_size.argtypes = [vp, C.POINTER(sz)]
    L.hzl_surrogate_fit.argtypes = [vp, sz, sz, sz, sz, C.c_double, C.c_uint64, _dp]
    L.hzl_surrogate_evaluate.argtypes = [vp, C.c_int, sz, sz, _dp]
    L.hzl_surrogate_forecast.argtypes = [vp, vp, C.c_double, sz, C.c_int]
    L.hzl_surrogate_error_estimate.argtypes = [vp, _dp]
    L.hzl_surrogate_set_graph.argtypes = [vp, C.c_int]
    L.hzl_surrogate_set_fused.argtypes = [vp, C.c_int]
    L.hzl_surrogate_set_error_estimate.argtypes = [vp, C.c_double]
    L.hzl_surrogate_spread.argtypes = [vp, vp, _dp]
    L.hzl_surrogate_forecast_or_solve.argtypes = [vp, vp, C.c_double, sz, C.c_int, C.c_double,
                                                  C.POINTER(C.c_int), _dp]
    for name in EXPORTS[1:]:
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc == HZL_OK:
        return
    msg = _lib.hzl_last_error().decode()
    if rc == HZL_EINVAL:
        raise HzlInvalidArgument(msg)
    raise HzlError(msg)


def _fields(a, count: int, n: int) -> np.ndarray:
    arr = np.ascontiguousarray(a, dtype=np.complex128)
    if arr.size != count * n * n:
        raise ValueError(f"expected {count} field(s) of {n}x{n}")
    return arr


class Ensemble:
    """A batch of independent members on one n x n grid (one engine handle)."""

    def __init__(self, n: int, members: int, precision: str = "fp32", *, nu: float = 1e-3,
                 alpha: float = 0.0, sigma: float = 0.05, tau: float = 0.5, kmin: int = 1,
                 kmax: int = 4, amp: float = 0.1, kf: int = 4, seed: int = 0):
        L = load_library()
        self.n, self.members = int(n), int(members)
        self.precision = {"fp32": FP32, "fp64": FP64}[precision]
        p = Parameters(nu, alpha, sigma, tau, kmin, kmax, amp, kf)
        h = C.c_void_p()
        _check(L.hzl_ensemble_create(C.byref(h), self.n, self.members, self.precision,
                                     C.byref(p), seed))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.hzl_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_member_parameters(self, m, nu, alpha, sigma, amp):
        _check(_lib.hzl_ensemble_set_member_parameters(self._h, m, nu, alpha, sigma, amp))

    def set_state(self, fields, member=None):
        if member is None:
            a = _fields(fields, self.members, self.n)
            _check(_lib.hzl_ensemble_set_state(self._h, ALL, a.ctypes.data_as(_dp)))
        else:
            a = _fields(fields, 1, self.n)
            _check(_lib.hzl_ensemble_set_state(self._h, member, a.ctypes.data_as(_dp)))

    def get_state(self, member=None) -> np.ndarray:
        if member is None:
            out = np.zeros((self.members, self.n, self.n), dtype=np.complex128)
            _check(_lib.hzl_ensemble_get_state(self._h, ALL, out.ctypes.data_as(_dp)))
        else:
            out = np.zeros((self.n, self.n), dtype=np.complex128)
            _check(_lib.hzl_ensemble_get_state(self._h, member, out.ctypes.data_as(_dp)))
        return out

    def init_grf(self):
        _check(_lib.hzl_ensemble_init_grf(self._h))

    def step(self, dt: float, steps: int = 1, integrator: str = "rk4"):
        """`steps` steps of every member; integrator "rk4" or "ssprk3"."""
        if integrator == "rk4":
            _check(_lib.hzl_ensemble_step(self._h, dt, steps))
        else:
            code = {"ssprk3": 1}.get(integrator)
            if code is None:
                raise HzlInvalidArgument("integrator must be rk4 or ssprk3")
            _check(_lib.hzl_ensemble_step_integrator(self._h, dt, steps, code))

    def set_forcing_process(self, process: str):
        """Shell forcing process: "white" (default, white in time) or "ou", the
        reference's Ornstein-Uhlenbeck process with an mt19937_64 NormalStream
        per member (ShellForcing, src/solver.cpp:84-119); selecting "ou" draws
        every member's stationary start from its stream seed."""
        code = {"white": 0, "ou": 1}.get(process)
        if code is None:
            raise HzlInvalidArgument("forcing process must be white or ou")
        _check(_lib.hzl_ensemble_set_forcing_process(self._h, code))

    def reseed_future(self, member: int, seed: int):
        """Solver::reseed_future: member's future innovations from `seed`."""
        _check(_lib.hzl_ensemble_reseed_future(self._h, member, C.c_uint64(seed)))

    def set_forcing_state(self, member: int, engine, forcing):
        """Member's OU stream (313 words: state, position) and forcing field."""
        words = np.ascontiguousarray(engine, dtype=np.uint64)
        if words.size != 313:
            raise HzlInvalidArgument("the stream is 312 words and the position")
        F = np.ascontiguousarray(forcing, dtype=np.complex128).reshape(self.n, self.n)
        _check(_lib.hzl_ensemble_set_forcing_state(
            self._h, member, words.ctypes.data_as(C.POINTER(C.c_uint64)),
            F.ctypes.data_as(_dp)))

    def get_forcing_state(self, member: int):
        """(stream words[313], forcing (n, n)) of a member (OU process)."""
        words = np.zeros(313, dtype=np.uint64)
        F = np.zeros((self.n, self.n), dtype=np.complex128)
        _check(_lib.hzl_ensemble_get_forcing_state(
            self._h, member, words.ctypes.data_as(C.POINTER(C.c_uint64)),
            F.ctypes.data_as(_dp)))
        return words, F

    def set_member_seeds(self, seeds):
        """Stream seed of every member (as Solver(n, p, seeds[m]))."""
        seeds = list(seeds)
        if len(seeds) != self.members:
            raise HzlInvalidArgument("one seed per member")
        for m, s in enumerate(seeds):
            self.reseed_future(m, int(s))

    def get_forcing(self, member=None) -> np.ndarray:
        """Current OU forcing coefficients (Solver::forcing()): (n, n) spectral
        for one member, (members, n, n) when member is None."""
        if member is None:
            out = np.zeros((self.members, self.n, self.n), dtype=np.complex128)
            member = ALL
        else:
            out = np.zeros((self.n, self.n), dtype=np.complex128)
        _check(_lib.hzl_ensemble_get_forcing(self._h, member, out.ctypes.data_as(_dp)))
        return out

    def step_tangent(self, dt: float, steps: int, tangents) -> np.ndarray:
        """Solver::step_tangent for every member: tangents (members, 2, n, n)
        in, advanced sensitivities out (the states advance too)."""
        t = np.ascontiguousarray(tangents, dtype=np.complex128).reshape(
            self.members, 2, self.n, self.n).copy()
        _check(_lib.hzl_ensemble_step_tangent(self._h, dt, steps, t.ctypes.data_as(_dp),
                                              t.ctypes.data_as(_dp)))
        return t

    def drift(self) -> np.ndarray:
        """Solver::drift of every member's state: (members, n, n) spectral."""
        out = np.zeros((self.members, self.n, self.n), dtype=np.complex128)
        _check(_lib.hzl_ensemble_drift(self._h, out.ctypes.data_as(_dp)))
        return out

    def linearized_drift(self, directions) -> np.ndarray:
        """Solver::linearized_drift of every member along directions[m]."""
        d = _fields(directions, self.members, self.n)
        out = np.zeros((self.members, self.n, self.n), dtype=np.complex128)
        _check(_lib.hzl_ensemble_linearized_drift(self._h, d.ctypes.data_as(_dp),
                                                  out.ctypes.data_as(_dp)))
        return out

    def adjoint_drift(self, cotangents) -> np.ndarray:
        """Solver::adjoint_drift of every member applied to cotangents[m]."""
        c = _fields(cotangents, self.members, self.n)
        out = np.zeros((self.members, self.n, self.n), dtype=np.complex128)
        _check(_lib.hzl_ensemble_adjoint_drift(self._h, c.ctypes.data_as(_dp),
                                               out.ctypes.data_as(_dp)))
        return out

    def set_graph(self, enabled: bool):
        """Replay steps and filter cycles from captured HIP graphs (default on)."""
        _check(_lib.hzl_ensemble_set_graph(self._h, int(bool(enabled))))

    # -- launch-shape variants (autotuner) ----------------------------------
    def variant_count(self):
        c, r = C.c_int(), C.c_int()
        _check(_lib.hzl_ensemble_variant_count(self._h, C.byref(c), C.byref(r)))
        return c.value, r.value

    @property
    def tuning(self):
        c, r = C.c_int(), C.c_int()
        _check(_lib.hzl_ensemble_get_tuning(self._h, C.byref(c), C.byref(r)))
        return c.value, r.value

    @tuning.setter
    def tuning(self, cr):
        _check(_lib.hzl_ensemble_set_tuning(self._h, int(cr[0]), int(cr[1])))

    def autotune(self, use_cache: bool = False):
        """Time every launch-shape variant now and keep the fastest; returns
        (column ms per step per variant, row ms per step per variant).  With
        use_cache a remembered choice is applied instead (timings all zero)."""
        nc, nr = self.variant_count()
        col, row = np.zeros(nc), np.zeros(nr)
        _check(_lib.hzl_ensemble_autotune(self._h, int(use_cache), col.ctypes.data_as(_dp),
                                          row.ctypes.data_as(_dp)))
        return col, row

    def step_forced(self, dt: float, forcing, shared: bool = True):
        a = _fields(forcing, 1 if shared else self.members, self.n)
        _check(_lib.hzl_ensemble_step_forced(self._h, dt, a.ctypes.data_as(_dp), int(shared)))

    def synchronize(self):
        _check(_lib.hzl_ensemble_synchronize(self._h))

    def diagnostics(self) -> np.ndarray:
        out = np.zeros((self.members, 6))
        _check(_lib.hzl_ensemble_diagnostics(self._h, out.ctypes.data_as(_dp)))
        return out

    def energy_spectrum(self) -> np.ndarray:
        """Shell energy spectrum of every member, (members, n) FP64
        (Solver::energy_spectrum, src/solver.cpp:426-433)."""
        out = np.zeros((self.members, self.n))
        _check(_lib.hzl_ensemble_energy_spectrum(self._h, out.ctypes.data_as(_dp)))
        return out

    def set_member_tau(self, member: int, tau: float) -> None:
        """Per-member OU time constant."""
        _check(_lib.hzl_ensemble_set_member_tau(self._h, member, tau))

    def profile_step(self, dt: float) -> np.ndarray:
        """Durations (ms) of the 8 launches of one step: row1, col1, ..., row4, col4."""
        out = np.zeros(8, dtype=np.float32)
        _check(_lib.hzl_ensemble_profile_step(self._h, dt,
                                              out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    @property
    def step_index(self) -> int:
        v = C.c_uint64()
        _check(_lib.hzl_ensemble_step_index(self._h, C.byref(v)))
        return v.value

    def device_state(self):
        p, s = C.c_void_p(), C.c_size_t()
        _check(_lib.hzl_ensemble_device_state(self._h, C.byref(p), C.byref(s)))
        return p.value, s.value

    @property
    def stream(self) -> int:
        p = C.c_void_p()
        _check(_lib.hzl_ensemble_stream(self._h, C.byref(p)))
        return p.value

    # -- particle filter (members are particles) ---------------------------
    def pf_begin(self, seed: int, max_groups: int = 4096):
        _check(_lib.hzl_pf_begin(self._h, seed, max_groups))
        self._filters = 1

    def pf_begin_filters(self, seeds, max_groups: int = 4096):
        s = np.ascontiguousarray(seeds, dtype=np.uint64)
        _check(_lib.hzl_pf_begin_filters(self._h, s.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         len(s), max_groups))
        self._filters = len(s)

    def fill_state(self, field):
        a = _fields(field, 1, self.n)
        _check(_lib.hzl_ensemble_fill_state(self._h, a.ctypes.data_as(_dp)))

    @step_index.setter
    def step_index(self, step: int):
        _check(_lib.hzl_ensemble_set_step_index(self._h, step))

    def pf_assimilate(self, dt, steps, x, y, value, sd, is_last, mean=False):
        """Returns (records, mean); records has 4 entries per filter."""
        x = np.ascontiguousarray(x, dtype=np.int32)
        y = np.ascontiguousarray(y, dtype=np.int32)
        value = np.ascontiguousarray(value, dtype=np.float64)
        sd = np.ascontiguousarray(sd, dtype=np.float64)
        rec = np.zeros(4 * getattr(self, "_filters", 1))
        # every entry is written (retained modes, their conjugates, zeros elsewhere)
        mf = np.empty((self.n, self.n), dtype=np.complex128) if mean else None
        ip = C.POINTER(C.c_int)
        _check(_lib.hzl_pf_assimilate(self._h, dt, steps, len(x), x.ctypes.data_as(ip),
                                      y.ctypes.data_as(ip), value.ctypes.data_as(_dp),
                                      sd.ctypes.data_as(_dp), int(is_last),
                                      rec.ctypes.data_as(_dp),
                                      mf.ctypes.data_as(_dp) if mean else None))
        return rec, mf

    # -- ensemble Kalman filter (members are the ensemble) ----------------
    def enkf_begin(self, seed: int, mode: str = "serial"):
        _check(_lib.hzl_enkf_begin(self._h, seed, {"serial": ENKF_SERIAL,
                                                   "batch": ENKF_BATCH}[mode]))

    def enkf_assimilate(self, dt, steps, x, y, value, sd, inflation=1.0, radius=0.0,
                        mean=False):
        x = np.ascontiguousarray(x, dtype=np.int32)
        y = np.ascontiguousarray(y, dtype=np.int32)
        value = np.ascontiguousarray(value, dtype=np.float64)
        sd = np.ascontiguousarray(sd, dtype=np.float64)
        rec = np.zeros(2)
        mf = np.zeros((self.n, self.n), dtype=np.complex128) if mean else None
        ip = C.POINTER(C.c_int)
        _check(_lib.hzl_enkf_assimilate(self._h, dt, steps, len(x), x.ctypes.data_as(ip),
                                        y.ctypes.data_as(ip), value.ctypes.data_as(_dp),
                                        sd.ctypes.data_as(_dp), inflation, radius,
                                        rec.ctypes.data_as(_dp),
                                        mf.ctypes.data_as(_dp) if mean else None))
        return rec, mf

    def mean_spread(self, mean=True):
        sp = C.c_double()
        mf = np.zeros((self.n, self.n), dtype=np.complex128) if mean else None
        _check(_lib.hzl_ensemble_mean_spread(self._h, C.byref(sp),
                                             mf.ctypes.data_as(_dp) if mean else None))
        return sp.value, mf

    def pf_result(self):
        """(log-likelihood, resampling events); arrays with several filters."""
        F = getattr(self, "_filters", 1)
        ll = np.zeros(F)
        ev = np.zeros(F, dtype=np.uint64)
        _check(_lib.hzl_pf_result(self._h, ll.ctypes.data_as(_dp),
                                  ev.ctypes.data_as(C.POINTER(C.c_size_t))))
        if F == 1:
            return float(ll[0]), int(ev[0])
        return ll, ev.astype(np.int64)


OBSERVATIONS_COLUMNS = "step,x,y,vorticity,standard_deviation"


def load_observations(path: str):
    """Read an HZL_OBSERVATIONS_V2 (or V1) CSV file (src/inference.cpp:226-255).
    Returns (meta, rows): meta dict(n, dt, stochastic, forcing_min,
    forcing_max), rows an (m, 5) array of (step, x, y, value, sd)."""
    with open(path) as f:
        lines = f.read().split("\n")
    head = lines[0].split(",")
    if head[0] not in ("HZL_OBSERVATIONS_V1", "HZL_OBSERVATIONS_V2") or len(head) < 3:
        raise HzlError("invalid dataset header")
    meta = dict(n=int(head[1]), dt=float(head[2]), stochastic=False, forcing_min=2,
                forcing_max=3)
    if head[0] == "HZL_OBSERVATIONS_V2":
        if len(head) != 6:
            raise HzlError("invalid stochastic dataset metadata")
        meta.update(stochastic=bool(int(head[3])), forcing_min=int(head[4]),
                    forcing_max=int(head[5]))
    if len(lines) < 2 or lines[1] != OBSERVATIONS_COLUMNS:
        raise HzlError("invalid dataset columns")
    rows = []
    for line in lines[2:]:
        if not line:
            continue
        parts = line.split(",")
        if len(parts) != 5:
            raise HzlError("invalid observation row")
        rows.append((int(parts[0]), int(parts[1]), int(parts[2]), float(parts[3]),
                     float(parts[4])))
    rows = np.array(rows, dtype=np.float64).reshape(-1, 5)
    validate_observations(meta, rows)
    return meta, rows


def validate_observations(meta, rows):
    """Dataset::validate (src/inference.cpp:197-213)."""
    n = meta["n"]
    if (n < 16 or n & (n - 1) or not np.isfinite(meta["dt"]) or meta["dt"] <= 0
            or len(rows) == 0 or meta["forcing_min"] < 1
            or meta["forcing_max"] < meta["forcing_min"] or 3 * meta["forcing_max"] >= n):
        raise HzlInvalidArgument("invalid observation dataset")
    if (np.any(np.diff(rows[:, 0]) < 0) or np.any(rows[:, 1] >= n) or np.any(rows[:, 2] >= n)
            or not np.all(np.isfinite(rows[:, 3:])) or np.any(rows[:, 4] <= 0)):
        raise HzlInvalidArgument("observations must be time ordered, finite, in bounds, and "
                                 "have positive uncertainty")


def save_observations(path: str, meta, rows):
    """Write an HZL_OBSERVATIONS_V2 CSV file with 17 significant digits
    (src/inference.cpp:214-225)."""
    rows = np.asarray(rows, dtype=np.float64)
    validate_observations(meta, rows)
    g = lambda v: "%.17g" % v
    out = [f"HZL_OBSERVATIONS_V2,{meta['n']},{g(meta['dt'])},{int(meta['stochastic'])},"
           f"{meta['forcing_min']},{meta['forcing_max']}", OBSERVATIONS_COLUMNS]
    out += [f"{int(r[0])},{int(r[1])},{int(r[2])},{g(r[3])},{g(r[4])}" for r in rows]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


CHECKPOINT_MAGIC = "HZL_CHECKPOINT_V1"
_ENGINE_WORDS = 312  # std::mt19937_64 state words (then its position)


def _retained_mask(n: int) -> np.ndarray:
    k = np.fft.fftfreq(n, 1.0 / n).astype(np.int64)
    kx, ky = np.meshgrid(k, k)  # [y][x]
    m = (3 * np.abs(kx) < n) & (3 * np.abs(ky) < n)
    m[0, 0] = False
    return m


def load_checkpoint(path: str) -> dict:
    """Read an HZL_CHECKPOINT_V1 file of Solver::save (src/solver.cpp:434-491)
    with Solver::load's checks.  Returns dict(n, time, parameters (the six
    Parameters fields), random_engine (313 words), random_has_spare,
    random_spare, vorticity, forcing (n x n spectral, [y][x]))."""
    with open(path) as f:
        tok = f.read().split()
    try:
        if tok[0] != CHECKPOINT_MAGIC:
            raise ValueError
        n, t = int(tok[1]), float(tok[2])
        par = dict(viscosity=float(tok[3]), friction=float(tok[4]), forcing_sigma=float(tok[5]),
                   forcing_tau=float(tok[6]), forcing_min=int(tok[7]), forcing_max=int(tok[8]))
        if not np.isfinite(t) or t < 0:
            raise ValueError
    except (ValueError, IndexError):
        raise HzlError("invalid checkpoint header") from None
    if n < 8 or n > 4096 or n & (n - 1):
        raise HzlInvalidArgument("grid must be a power of two between 8 and 4096")
    if (not all(np.isfinite([par["viscosity"], par["friction"], par["forcing_sigma"],
                             par["forcing_tau"]]))
            or par["viscosity"] < 0 or par["friction"] < 0 or par["forcing_sigma"] < 0
            or par["forcing_tau"] <= 0):
        raise HzlInvalidArgument("nu, alpha, sigma must be finite and nonnegative; "
                                 "tau must be positive")
    if par["forcing_min"] < 1 or par["forcing_max"] < par["forcing_min"] or \
            3 * par["forcing_max"] >= n:
        raise HzlInvalidArgument("forcing shell must lie strictly inside the dealiased grid")
    at = 9
    try:
        engine = [int(w) for w in tok[at:at + _ENGINE_WORDS + 1]]
        spare_flag, spare = int(tok[at + _ENGINE_WORDS + 1]), float(tok[at + _ENGINE_WORDS + 2])
        if (len(engine) != _ENGINE_WORDS + 1 or engine[-1] > _ENGINE_WORDS
                or any(w < 0 or w >= 1 << 64 for w in engine) or spare_flag not in (0, 1)
                or not np.isfinite(spare)):
            raise ValueError
    except (ValueError, IndexError):
        raise HzlError("invalid checkpoint random stream") from None
    at += _ENGINE_WORDS + 3
    body = tok[at:]
    if len(body) < 4 * n * n:
        raise HzlError("invalid checkpoint field")
    try:
        vals = np.array([float(x) for x in body[:4 * n * n]]).reshape(n * n, 4)
    except ValueError:
        raise HzlError("invalid checkpoint field") from None
    if not np.all(np.isfinite(vals)):
        raise HzlError("invalid checkpoint field")
    w = (vals[:, 0] + 1j * vals[:, 1]).reshape(n, n)
    F = (vals[:, 2] + 1j * vals[:, 3]).reshape(n, n)
    mask = _retained_mask(n)
    k = np.fft.fftfreq(n, 1.0 / n).astype(np.int64)
    kx, ky = np.meshgrid(k, k)
    k2 = kx * kx + ky * ky
    shell = mask & (k2 >= par["forcing_min"] ** 2) & (k2 <= par["forcing_max"] ** 2)
    partner = lambda a: np.roll(a[::-1, ::-1], 1, axis=(0, 1))  # a[(-y) % n, (-x) % n]
    if (np.any(w[~mask] != 0) or np.any(F[~shell] != 0)
            or np.any(np.abs(w - np.conj(partner(w))) > 1e-10 * (1 + np.abs(w)))
            or np.any(F != np.conj(partner(F)))):
        raise HzlError("checkpoint violates spectral constraints")
    if len(body) > 4 * n * n:
        raise HzlError("unexpected checkpoint data")
    return dict(n=n, time=t, parameters=par, random_engine=engine,
                random_has_spare=bool(spare_flag), random_spare=spare, vorticity=w, forcing=F)


def save_checkpoint(path: str, c: dict) -> None:
    """Write an HZL_CHECKPOINT_V1 file byte for byte as Solver::save does
    (17 significant digits, src/solver.cpp:434-449)."""
    g = lambda v: "%.17g" % v
    n, p = int(c["n"]), c["parameters"]
    w = np.asarray(c["vorticity"], dtype=np.complex128).reshape(n * n)
    F = np.asarray(c["forcing"], dtype=np.complex128).reshape(n * n)
    engine = [int(x) for x in c["random_engine"]]
    if len(engine) != _ENGINE_WORDS + 1:
        raise HzlInvalidArgument("checkpoint random stream needs 313 words")
    out = [CHECKPOINT_MAGIC, f"{n} {g(c['time'])}",
           f"{g(p['viscosity'])} {g(p['friction'])} {g(p['forcing_sigma'])} "
           f"{g(p['forcing_tau'])} {int(p['forcing_min'])} {int(p['forcing_max'])}",
           " ".join(str(x) for x in engine),
           f"{int(bool(c['random_has_spare']))} {g(c['random_spare'])}"]
    out += [f"{g(a.real)} {g(a.imag)} {g(b.real)} {g(b.imag)}" for a, b in zip(w, F)]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def _observation_arrays(observations):
    obs = np.asarray(observations, dtype=np.float64)
    if obs.ndim != 2 or obs.shape[1] != 5 or len(obs) == 0:
        raise ValueError("observations must be rows of (step, x, y, value, sd)")
    return (np.ascontiguousarray(obs[:, 0], dtype=np.uint64).astype(np.uintp),
            np.ascontiguousarray(obs[:, 1], dtype=np.int32),
            np.ascontiguousarray(obs[:, 2], dtype=np.int32),
            np.ascontiguousarray(obs[:, 3]), np.ascontiguousarray(obs[:, 4]))


def likelihood_gradient(ensemble: Ensemble, initial, observations, dt: float, log_params):
    """Batched likelihood_gradient (src/inference.cpp:340-382): the objective
    of likelihood_adjoint with its gradient by forward sensitivities.
    Returns (values [members], gradients [members, 2])."""
    steps, x, y, v, sd = _observation_arrays(observations)
    q = np.ascontiguousarray(log_params, dtype=np.float64).reshape(ensemble.members, 2)
    init = _fields(initial, 1, ensemble.n)
    values = np.zeros(ensemble.members)
    grads = np.zeros((ensemble.members, 2))
    ip, szp = C.POINTER(C.c_int), C.POINTER(C.c_size_t)
    _check(_lib.hzl_likelihood_gradient(ensemble._h, init.ctypes.data_as(_dp), dt, len(x),
                                        steps.ctypes.data_as(szp), x.ctypes.data_as(ip),
                                        y.ctypes.data_as(ip), v.ctypes.data_as(_dp),
                                        sd.ctypes.data_as(_dp), q.ctypes.data_as(_dp),
                                        values.ctypes.data_as(_dp), grads.ctypes.data_as(_dp)))
    return values, grads


def likelihood_adjoint(ensemble: Ensemble, initial, observations, dt: float, log_params,
                       stride: int = 20):
    """Batched likelihood_adjoint (src/adjoint.cpp:43-161): member m evaluates
    the deterministic negative log-likelihood of `observations` (rows of
    step, x, y, value, sd) and its gradient at log_params[m] = (log nu,
    log alpha), from `initial` (one spectral field).  Returns (values
    [members], gradients [members, 2], counts (checkpoint fields, peak block
    fields, replayed steps))."""
    obs = np.asarray(observations, dtype=np.float64)
    if obs.ndim != 2 or obs.shape[1] != 5 or len(obs) == 0:
        raise ValueError("observations must be rows of (step, x, y, value, sd)")
    q = np.ascontiguousarray(log_params, dtype=np.float64).reshape(ensemble.members, 2)
    init = _fields(initial, 1, ensemble.n)
    steps = np.ascontiguousarray(obs[:, 0], dtype=np.uint64).astype(np.uintp)
    x = np.ascontiguousarray(obs[:, 1], dtype=np.int32)
    y = np.ascontiguousarray(obs[:, 2], dtype=np.int32)
    v = np.ascontiguousarray(obs[:, 3])
    sd = np.ascontiguousarray(obs[:, 4])
    values = np.zeros(ensemble.members)
    grads = np.zeros((ensemble.members, 2))
    counts = np.zeros(3, dtype=np.uintp)
    ip, szp = C.POINTER(C.c_int), C.POINTER(C.c_size_t)
    _check(_lib.hzl_likelihood_adjoint(ensemble._h, init.ctypes.data_as(_dp), dt, len(obs),
                                       steps.ctypes.data_as(szp), x.ctypes.data_as(ip),
                                       y.ctypes.data_as(ip), v.ctypes.data_as(_dp),
                                       sd.ctypes.data_as(_dp), q.ctypes.data_as(_dp), stride,
                                       values.ctypes.data_as(_dp), grads.ctypes.data_as(_dp),
                                       counts.ctypes.data_as(szp)))
    return values, grads, tuple(int(c) for c in counts)


def bootstrap_filter(particles: Ensemble, observations, dt: float, seed: int,
                     retain_states: bool = True):
    """Bootstrap particle filter over an initialised ensemble.

    Mirrors hzl::bootstrap_filter (src/filter.cpp:38-126): `observations` rows are
    (step, x, y, value, sd), time ordered; rows sharing a step form one group.
    Returns dict(log_likelihood, resampling_events, history) with history
    records (step, increment, ess, max_weight, resampled, mean_field).
    """
    obs = np.asarray(observations, dtype=np.float64)
    if obs.ndim != 2 or obs.shape[1] != 5 or len(obs) == 0:
        raise ValueError("observations must be rows of (step, x, y, value, sd)")
    if np.any(np.diff(obs[:, 0]) < 0):
        raise ValueError("observations must be time ordered")
    steps = obs[:, 0].astype(np.int64)
    groups = np.unique(steps)
    particles.pf_begin(seed, max_groups=len(groups) + 1)
    history, current = [], 0
    for gi, target in enumerate(groups):
        sel = obs[steps == target]
        rec, mean = particles.pf_assimilate(dt, int(target - current), sel[:, 1], sel[:, 2],
                                            sel[:, 3], sel[:, 4], gi == len(groups) - 1,
                                            mean=retain_states)
        if not np.isfinite(rec[0]):
            break  # every weight -inf: no record for this group (src/filter.cpp:76-80)
        history.append(dict(step=int(target), increment=rec[0], ess=rec[1],
                            max_weight=rec[2], resampled=bool(rec[3]), mean_field=mean))
        current = int(target)
    ll, events = particles.pf_result()
    return dict(log_likelihood=ll, resampling_events=events, history=history)


def joint_bootstrap_filter(ensembles, observations, dt: float, seed: int):
    """One bootstrap particle filter over the members of several ensembles of
    different grid sizes (hzl_pf_joint_*); observation grid indices on the
    coarsest grid.  Returns dict(log_likelihood, resampling_events, history)."""
    L = load_library()
    obs = np.asarray(observations, dtype=np.float64)
    if obs.ndim != 2 or obs.shape[1] != 5 or len(obs) == 0:
        raise ValueError("observations must be rows of (step, x, y, value, sd)")
    handles = (C.c_void_p * len(ensembles))(*[e._h.value for e in ensembles])
    h = C.c_void_p()
    _check(L.hzl_pf_joint_create(C.byref(h), handles, len(ensembles)))
    try:
        steps = obs[:, 0].astype(np.int64)
        groups = np.unique(steps)
        _check(L.hzl_pf_joint_begin(h, seed, len(groups) + 1))
        ip = C.POINTER(C.c_int)
        history, current = [], 0
        for gi, target in enumerate(groups):
            sel = obs[steps == target]
            x = np.ascontiguousarray(sel[:, 1], dtype=np.int32)
            y = np.ascontiguousarray(sel[:, 2], dtype=np.int32)
            v = np.ascontiguousarray(sel[:, 3])
            sd = np.ascontiguousarray(sel[:, 4])
            rec = np.zeros(4)
            _check(L.hzl_pf_joint_assimilate(h, dt, int(target - current), len(sel),
                                             x.ctypes.data_as(ip), y.ctypes.data_as(ip),
                                             v.ctypes.data_as(_dp), sd.ctypes.data_as(_dp),
                                             int(gi == len(groups) - 1),
                                             rec.ctypes.data_as(_dp)))
            if not np.isfinite(rec[0]):
                break  # every weight -inf (src/filter.cpp:76-80)
            history.append(dict(step=int(target), increment=rec[0], ess=rec[1],
                                max_weight=rec[2], resampled=bool(rec[3])))
            current = int(target)
        ll, ev = C.c_double(), C.c_size_t()
        _check(L.hzl_pf_joint_result(h, C.byref(ll), C.byref(ev)))
    finally:
        L.hzl_pf_joint_destroy(h)
    return dict(log_likelihood=ll.value, resampling_events=ev.value, history=history)


def ensemble_kalman_filter(members: Ensemble, observations, dt: float, seed: int,
                           mode: str = "serial", inflation: float = 1.0,
                           localization_radius: float = 0.0, forecast_steps: int = 0,
                           forecast_every: int = 1, retain_states: bool = True):
    """Perturbed-observation EnKF over an initialised ensemble.

    Mirrors hzl::ensemble_kalman_filter (src/filter.cpp:204-356): rows
    (step, x, y, value, sd) grouped by step; per group a record dict(step,
    forecast_sensor_rmse, posterior_spread, mean_field, is_forecast), then
    `forecast_steps` free steps recorded every `forecast_every` (and the last).
    mode "serial" follows the reference's scalar order; "batch" assimilates
    each group at once on MFMA (FP32 ensembles).
    """
    obs = np.asarray(observations, dtype=np.float64)
    if obs.ndim != 2 or obs.shape[1] != 5 or len(obs) == 0:
        raise ValueError("observations must be rows of (step, x, y, value, sd)")
    if np.any(np.diff(obs[:, 0]) < 0):
        raise ValueError("observations must be time ordered")
    if forecast_every <= 0:
        raise HzlInvalidArgument("EnKF forecast output interval must be positive")
    members.enkf_begin(seed, mode)
    steps = obs[:, 0].astype(np.int64)
    history, current = [], 0
    for target in np.unique(steps):
        sel = obs[steps == target]
        rec, mean = members.enkf_assimilate(dt, int(target - current), sel[:, 1], sel[:, 2],
                                            sel[:, 3], sel[:, 4], inflation,
                                            localization_radius, mean=retain_states)
        history.append(dict(step=int(target), forecast_sensor_rmse=rec[0],
                            posterior_spread=rec[1], mean_field=mean, is_forecast=False))
        current = int(target)
    for horizon in range(1, forecast_steps + 1):
        members.step(dt, 1)
        if horizon % forecast_every != 0 and horizon != forecast_steps:
            continue
        sp, mean = members.mean_spread(mean=retain_states)
        history.append(dict(step=current + horizon, forecast_sensor_rmse=float("nan"),
                            posterior_spread=sp, mean_field=mean, is_forecast=True))
    return history


def particle_metropolis(particles: Ensemble, observations, dt: float, initial_field,
                        initial_q, samples: int, proposal_scale: float, seed: int,
                        nu_bounds=(2e-4, 5e-3), amp_bounds=(0.02, 0.5), friction: float = 0.0,
                        forcing_sigma: float = 0.05, surrogate: "Surrogate | None" = None,
                        surrogate_precision: str = "bf16"):
    """Batched pseudo-marginal random-walk Metropolis over q = (log nu,
    log amplitude), one chain per filter of members / chains particles
    (reference particle_metropolis, src/filter.cpp:166-203, with a
    log-uniform prior).  initial_q: (chains, 2).  Returns dict(q (chains,
    samples, 2), log_likelihood, log_density, accepted (chains, samples))."""
    obs = np.asarray(observations, dtype=np.float64)
    q0 = np.ascontiguousarray(initial_q, dtype=np.float64).reshape(-1, 2)
    chains = q0.shape[0]
    o = PmcmcOptions(chains, samples, proposal_scale,
                     (C.c_double * 2)(np.log(nu_bounds[0]), np.log(amp_bounds[0])),
                     (C.c_double * 2)(np.log(nu_bounds[1]), np.log(amp_bounds[1])),
                     friction, forcing_sigma, seed, PRIOR_LOG_UNIFORM, None,
                     surrogate._h.value if surrogate is not None else None,
                     _SUR_PRECISION[surrogate_precision])
    step = np.ascontiguousarray(obs[:, 0], dtype=np.int32)
    x = np.ascontiguousarray(obs[:, 1], dtype=np.int32)
    y = np.ascontiguousarray(obs[:, 2], dtype=np.int32)
    val = np.ascontiguousarray(obs[:, 3])
    sd = np.ascontiguousarray(obs[:, 4])
    f = _fields(initial_field, 1, particles.n)
    q = np.zeros((chains, samples, 2))
    ll = np.zeros((chains, samples))
    dens = np.zeros((chains, samples))
    acc = np.zeros((chains, samples), dtype=np.int32)
    ip = C.POINTER(C.c_int)
    _check(_lib.hzl_pmcmc_run(particles._h, C.byref(o), dt, len(obs), step.ctypes.data_as(ip),
                              x.ctypes.data_as(ip), y.ctypes.data_as(ip),
                              val.ctypes.data_as(_dp), sd.ctypes.data_as(_dp),
                              f.ctypes.data_as(_dp), q0.ctypes.data_as(_dp),
                              q.ctypes.data_as(_dp), ll.ctypes.data_as(_dp),
                              dens.ctypes.data_as(_dp), acc.ctypes.data_as(ip), None))
    particles._filters = chains
    return dict(q=q, log_likelihood=ll, log_density=dens, accepted=acc.astype(bool))


def pf_log_likelihoods(particles: Ensemble, observations, dt: float, initial_field, q, seeds,
                       nu_bounds=(2e-4, 5e-3), amp_bounds=(0.02, 0.5), friction: float = 0.0,
                       forcing_sigma: float = 0.05, surrogate: "Surrogate | None" = None,
                       surrogate_precision: str = "bf16") -> np.ndarray:
    """Log-likelihood estimates of len(q) independent bootstrap filters over
    the ensemble's members (members / len(q) particles each): filter f at
    q[f] = (log nu, log amplitude) with filter seed seeds[f], its particles
    restarting from initial_field at step 0 (hzl_pf_log_likelihoods; the
    likelihood the pMCMC chains evaluate, src/filter.cpp:141-149).  -inf
    outside the log-uniform prior box.  With a surrogate, every interval's
    forecast is the network's."""
    obs = np.asarray(observations, dtype=np.float64)
    qs = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 2)
    sd_ = np.ascontiguousarray(seeds, dtype=np.uint64)
    F = qs.shape[0]
    if sd_.shape != (F,):
        raise ValueError("one filter seed per parameter point")
    o = PmcmcOptions(F, 0, 0.0,
                     (C.c_double * 2)(np.log(nu_bounds[0]), np.log(amp_bounds[0])),
                     (C.c_double * 2)(np.log(nu_bounds[1]), np.log(amp_bounds[1])),
                     friction, forcing_sigma, 0, PRIOR_LOG_UNIFORM, None,
                     surrogate._h.value if surrogate is not None else None,
                     _SUR_PRECISION[surrogate_precision])
    ip = C.POINTER(C.c_int)
    step = np.ascontiguousarray(obs[:, 0], dtype=np.int32)
    x = np.ascontiguousarray(obs[:, 1], dtype=np.int32)
    y = np.ascontiguousarray(obs[:, 2], dtype=np.int32)
    val = np.ascontiguousarray(obs[:, 3])
    sd = np.ascontiguousarray(obs[:, 4])
    f = _fields(initial_field, 1, particles.n)
    ll = np.zeros(F)
    _check(_lib.hzl_pf_log_likelihoods(particles._h, C.byref(o), dt, len(obs),
                                       step.ctypes.data_as(ip), x.ctypes.data_as(ip),
                                       y.ctypes.data_as(ip), val.ctypes.data_as(_dp),
                                       sd.ctypes.data_as(_dp), f.ctypes.data_as(_dp), F,
                                       qs.ctypes.data_as(_dp),
                                       sd_.ctypes.data_as(C.POINTER(C.c_uint64)),
                                       ll.ctypes.data_as(_dp)))
    particles._filters = F
    return ll


_SUR_PRECISION = {"fp32": FP32, "bf16": BF16}


class Surrogate:
    """Learned forecast surrogate (spectral neural operator) on one handle.

    No reference counterpart: it replaces the particle forecast of the filter
    loops (src/filter.cpp:46-48).  See include/hzl_amd.h for every call.
    """

    def __init__(self, state_n: int = 128, grid: int = 64, width: int = 32, layers: int = 4,
                 modes: int = 16, seed: int = 0):
        L = load_library()
        h = C.c_void_p()
        _check(L.hzl_surrogate_create(C.byref(h), state_n, grid, width, layers, modes, seed))
        self._h = h
        self.state_n, self.grid = int(state_n), int(grid)
        self.width, self.layers, self.modes = int(width), int(layers), int(modes)

    def close(self):
        if getattr(self, "_h", None):
            _lib.hzl_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def param_count(self) -> int:
        v = C.c_size_t()
        _check(_lib.hzl_surrogate_param_count(self._h, C.byref(v)))
        return v.value

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.param_count, dtype=np.float32)
        _check(_lib.hzl_surrogate_get_params(self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        if a.size != self.param_count:
            raise ValueError("parameter vector has the wrong length")
        _check(_lib.hzl_surrogate_set_params(self._h, a.ctypes.data_as(C.POINTER(C.c_float))))

    @property
    def scales(self):
        out = np.zeros(2)
        _check(_lib.hzl_surrogate_get_scales(self._h, out.ctypes.data_as(_dp)))
        return float(out[0]), float(out[1])

    def set_scales(self, input_scale: float, output_scale: float):
        _check(_lib.hzl_surrogate_set_scales(self._h, input_scale, output_scale))

    def forward(self, fields, nu, precision: str = "fp32", amp=None) -> np.ndarray:
        """Network output per sample; amp: Kolmogorov amplitudes (default 0.1)."""
        f = np.ascontiguousarray(fields, dtype=np.complex128)
        S = f.shape[0]
        nu = np.ascontiguousarray(np.broadcast_to(nu, (S,)), dtype=np.float64)
        a = None if amp is None else np.ascontiguousarray(np.broadcast_to(amp, (S,)),
                                                          dtype=np.float64)
        _fields(f, S, self.state_n)
        out = np.zeros((S, self.grid, self.grid))
        _check(_lib.hzl_surrogate_forward(self._h, _SUR_PRECISION[precision], S,
                                          f.ctypes.data_as(_dp), nu.ctypes.data_as(_dp),
                                          None if a is None else a.ctypes.data_as(_dp),
                                          out.ctypes.data_as(_dp)))
        return out

    def judge(self, ensemble: Ensemble, judge: Ensemble, dt: float, steps: int,
              precision: str = "bf16") -> float:
        """Live error estimate: the first judge.members members forecast by the
        network and by the solver (noise-free); sets and returns the RMS."""
        v = C.c_double()
        _check(_lib.hzl_surrogate_judge(self._h, ensemble._h, judge._h, dt, steps,
                                        _SUR_PRECISION[precision], C.byref(v)))
        return v.value

    def judge_start(self, ensemble: Ensemble, judge: Ensemble, dt: float, steps: int,
                    precision: str = "bf16") -> None:
        """The same judge, enqueued on the judge ensemble's stream without
        waiting (the ensemble's stream waits only for the copy of the judged
        states); judge_poll folds its result into the error estimate."""
        _check(_lib.hzl_surrogate_judge_start(self._h, ensemble._h, judge._h, dt, steps,
                                              _SUR_PRECISION[precision]))

    def judge_poll(self, wait: bool = False):
        """(finished, error_estimate): finished is True when this call folded a
        judge in flight into the estimate (wait=True blocks until it ends)."""
        f, v = C.c_int(), C.c_double()
        _check(_lib.hzl_surrogate_judge_poll(self._h, int(wait), C.byref(f), C.byref(v)))
        return bool(f.value), v.value

    def loss_grad(self, fields, next_fields, nu, amp=None):
        f = np.ascontiguousarray(fields, dtype=np.complex128)
        g = np.ascontiguousarray(next_fields, dtype=np.complex128)
        S = f.shape[0]
        _fields(f, S, self.state_n)
        _fields(g, S, self.state_n)
        nu = np.ascontiguousarray(np.broadcast_to(nu, (S,)), dtype=np.float64)
        a = None if amp is None else np.ascontiguousarray(np.broadcast_to(amp, (S,)),
                                                          dtype=np.float64)
        loss = C.c_double()
        grad = np.zeros(self.param_count, dtype=np.float32)
        _check(_lib.hzl_surrogate_loss_grad(self._h, S, f.ctypes.data_as(_dp),
                                            g.ctypes.data_as(_dp), nu.ctypes.data_as(_dp),
                                            None if a is None else a.ctypes.data_as(_dp),
                                            C.byref(loss), grad.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, grad

    def collect(self, ensemble: Ensemble, dt: float, horizon_steps: int, windows: int,
                noise_free_targets: bool = True):
        _check(_lib.hzl_surrogate_collect(self._h, ensemble._h, dt, horizon_steps, windows,
                                          int(noise_free_targets)))

    @property
    def dataset_size(self) -> int:
        v = C.c_size_t()
        _check(_lib.hzl_surrogate_dataset_size(self._h, C.byref(v)))
        return v.value

    def fit(self, first: int, count: int, epochs: int, batch: int, lr: float, seed: int = 0):
        out = np.zeros(epochs)
        _check(_lib.hzl_surrogate_fit(self._h, first, count, epochs, batch, lr, seed,
                                      out.ctypes.data_as(_dp)))
        return out

    def evaluate(self, first: int, count: int, precision: str = "fp32") -> float:
        v = C.c_double()
        _check(_lib.hzl_surrogate_evaluate(self._h, _SUR_PRECISION[precision], first, count,
                                           C.byref(v)))
        return v.value

    def forecast(self, ensemble: Ensemble, dt: float, steps: int, precision: str = "bf16"):
        _check(_lib.hzl_surrogate_forecast(self._h, ensemble._h, dt, steps,
                                           _SUR_PRECISION[precision]))

    def set_graph(self, enabled: bool):
        """Replay forecasts from a captured HIP graph (default) or launch each kernel."""
        _check(_lib.hzl_surrogate_set_graph(self._h, int(enabled)))

    def set_fused(self, enabled: bool):
        """BF16 inference through the fused layer passes (default) or the
        training path's batched products (hzl_surrogate_set_fused)."""
        _check(_lib.hzl_surrogate_set_fused(self._h, int(enabled)))

    @property
    def error_estimate(self) -> float:
        v = C.c_double()
        _check(_lib.hzl_surrogate_error_estimate(self._h, C.byref(v)))
        return v.value

    @error_estimate.setter
    def error_estimate(self, rms: float):
        _check(_lib.hzl_surrogate_set_error_estimate(self._h, rms))

    def spread(self, ensemble: Ensemble) -> float:
        v = C.c_double()
        _check(_lib.hzl_surrogate_spread(self._h, ensemble._h, C.byref(v)))
        return v.value

    def forecast_or_solve(self, ensemble: Ensemble, dt: float, steps: int, threshold: float,
                          precision: str = "bf16"):
        """Surrogate forecast, or the plain solver when error > threshold x spread.
        Returns (used_solver, spread)."""
        used, sp = C.c_int(), C.c_double()
        _check(_lib.hzl_surrogate_forecast_or_solve(self._h, ensemble._h, dt, steps,
                                                    _SUR_PRECISION[precision], threshold,
                                                    C.byref(used), C.byref(sp)))
        return bool(used.value), sp.value


def surrogate_filter(particles: Ensemble, surrogate: Surrogate, observations, dt: float,
                     seed: int, precision: str = "bf16", retain_states: bool = True,
                     fallback_threshold: float | None = None, judge: Ensemble | None = None,
                     judge_every: int = 5, judge_async: bool = False):
    """Bootstrap particle filter whose forecast between observation groups is
    one surrogate evaluation (every group interval must equal the surrogate's
    horizon); weighting and resampling as in bootstrap_filter.  With
    fallback_threshold, a group whose surrogate error estimate exceeds
    threshold x the ensemble spread is forecast by the plain solver instead
    (history records `solver`, `spread` and `error_estimate`).  With a judge
    ensemble, every judge_every-th group first re-forecasts judge.members
    particles with the solver and the network (Surrogate.judge) and uses their
    measured RMS difference as the error estimate.  With judge_async, only the
    first group's judge is waited for; later ones run on the judge ensemble's
    stream beside the filter's cycles and each group uses the latest finished
    judge (history records `judged` on the groups whose estimate is new)."""
    obs = np.asarray(observations, dtype=np.float64)
    if obs.ndim != 2 or obs.shape[1] != 5 or len(obs) == 0:
        raise ValueError("observations must be rows of (step, x, y, value, sd)")
    if np.any(np.diff(obs[:, 0]) < 0):
        raise ValueError("observations must be time ordered")
    steps = obs[:, 0].astype(np.int64)
    groups = np.unique(steps)
    particles.pf_begin(seed, max_groups=len(groups) + 1)
    history, current = [], 0
    for gi, target in enumerate(groups):
        sel = obs[steps == target]
        solver, spread, judged = False, None, False
        if judge is not None and judge_async and gi > 0:
            # the previous judge has had judge_every groups to finish
            judged = surrogate.judge_poll(wait=gi % judge_every == 0)[0]
            if gi % judge_every == 0:
                surrogate.judge_start(particles, judge, dt, int(target - current), precision)
        elif judge is not None and gi % judge_every == 0:
            surrogate.judge(particles, judge, dt, int(target - current), precision)
            judged = True
        if fallback_threshold is None:
            surrogate.forecast(particles, dt, int(target - current), precision)
        else:
            solver, spread = surrogate.forecast_or_solve(particles, dt, int(target - current),
                                                         fallback_threshold, precision)
        rec, mean = particles.pf_assimilate(dt, 0, sel[:, 1], sel[:, 2], sel[:, 3], sel[:, 4],
                                            gi == len(groups) - 1, mean=retain_states)
        if not np.isfinite(rec[0]):
            break  # every weight -inf (src/filter.cpp:76-80)
        history.append(dict(step=int(target), increment=rec[0], ess=rec[1],
                            max_weight=rec[2], resampled=bool(rec[3]), mean_field=mean,
                            solver=solver, spread=spread, judged=judged,
                            error_estimate=surrogate.error_estimate))
        current = int(target)
    if judge is not None and judge_async:
        surrogate.judge_poll(wait=True)
    ll, events = particles.pf_result()
    return dict(log_likelihood=ll, resampling_events=events, history=history)
