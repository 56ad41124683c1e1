"""MI355X (gfx950) batched vorticity ensemble engine: Python host bindings.

The product is the C-ABI library libhzl_amd.so (include/hzl_amd.h); this
package only binds it.  The directory name is not a Python identifier, so it is
loaded with `load_package()` from __graft_entry__.py (or importlib) under the
module name `hzl_sfi_amd`.
"""
from .engine import (ALL, BF16, EXPORTS, FP32, FP64, Ensemble, HzlError, HzlInvalidArgument,
                     Parameters, Surrogate, bootstrap_filter, ensemble_kalman_filter,
                     joint_bootstrap_filter, likelihood_adjoint, likelihood_gradient,
                     load_checkpoint, load_library,
                     load_observations,
                     particle_metropolis, pf_log_likelihoods, save_checkpoint,
                     save_observations, surrogate_filter, validate_observations)

__all__ = ["ALL", "BF16", "EXPORTS", "FP32", "FP64", "Ensemble", "HzlError",
           "HzlInvalidArgument", "Parameters", "Surrogate", "bootstrap_filter",
           "ensemble_kalman_filter", "joint_bootstrap_filter", "likelihood_adjoint", "likelihood_gradient",
           "load_checkpoint",
           "load_library", "load_observations", "particle_metropolis", "pf_log_likelihoods",
           "save_checkpoint",
           "save_observations", "surrogate_filter", "validate_observations"]
