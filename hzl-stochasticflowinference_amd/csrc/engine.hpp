// engine.hpp -- internal interface between the engine's translation units.
#pragma once
#include "kernels.hpp"
#include <cstddef>
#include <cstdint>
#include <hip/hip_runtime.h>
#include <memory>

namespace hzl_amd {

// Execution context of device kernels (the CPU lane emulator has its own).
struct DevCtx {
    __host__ __device__ int tid() const {
#if __HIP_DEVICE_COMPILE__
        return (int)threadIdx.x;
#else
        return 0;
#endif
    }
    __host__ __device__ void sync() const {
#if __HIP_DEVICE_COMPILE__
        __syncthreads();
#endif
    }
    // Wave-local ordering of LDS traffic: a compiler barrier plus wavefront-
    // scope fences (the hardware completes one wavefront's LDS ops in order).
    __host__ __device__ void wsync() const {
#if __HIP_DEVICE_COMPILE__
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#endif
    }
    // value of v held by lane (lane ^ mask) of the wavefront
    __host__ __device__ float shfl_xor(float v, int mask, int) const {
#if __HIP_DEVICE_COMPILE__
        return __shfl_xor(v, mask);
#else
        (void)mask;
        return v;
#endif
    }
    // host-only placeholders so host-side template checks of HD bodies that
    // emulate matrix instructions on the CPU compile; never called
    template <class A, class B, class C> void mfma_bf16(const A &, const B &, C &, int) const {}
    template <class A, class B, class C> void mfma_f32(const A &, const B &, C &, int) const {}
};

// Raw device pointers and sizes of an ensemble, for modules that operate on
// its members in place (FP32 engines only where noted).
struct EngineViews {
    int n, K, B, precision; // precision: 0 FP32, 1 FP64
    void *w0;               // retained half-plane states [B][K+1][2K+1] complex
    double *nu, *alpha, *sigma, *amp;
    uint64_t *keys;
    const int *pair_of;
    int kmin, kmax, kf;
    hipStream_t stream;
    // OU forcing process (null / 0 with the white process): coefficients
    // [B][npairs] in the reference's pair order and, per pair, the half-plane
    // offset c (2K+1) + ky + K of its canonical mode (~offset when the pair's
    // first mode is the conjugate of the stored one)
    cx<double> *ou_F = nullptr;
    int npairs = 0;
    const int *ou_off = nullptr;
};

// State that another module keeps per ensemble handle (owned by the handle).
struct ModuleState {
    virtual ~ModuleState() = default;
};

class EngineBase {
  public:
    virtual ~EngineBase() = default;
    std::unique_ptr<ModuleState> enkf; // ensemble Kalman filter state, if started
    virtual void set_member_parameters(size_t m, double nu, double alpha, double sigma,
                                       double amp) = 0;
    // The host copies of a member's parameters (the values the device holds
    // once the next launch has uploaded them).
    virtual void get_member_parameters(size_t m, double *nu, double *alpha, double *sigma,
                                       double *amp) const = 0;
    // Per-member OU time constant (the creation parameters' forcing_tau until set).
    virtual void set_member_tau(size_t m, double tau) = 0;
    virtual void set_state(size_t m, const double *field) = 0;
    virtual void get_state(size_t m, double *field) = 0;
    virtual void init_grf() = 0;
    virtual void step(double dt, size_t steps) = 0;
    virtual void set_graph(bool on) = 0;
    // SSPRK3 stepping and the calibration drifts (batched over members)
    virtual void step_integrator(double dt, size_t steps, int integrator) = 0;
    // Forcing process of the shell noise: HZL_FORCING_WHITE (the engine's
    // white-in-time noise) or HZL_FORCING_OU (the reference's OU process,
    // mt19937_64 streams per member); get_forcing: current OU coefficients.
    virtual void set_forcing_process(int process) = 0;
    virtual int forcing_process() const = 0;
    virtual void get_forcing(size_t m, double *field) = 0;
    virtual void reseed_future(size_t m, uint64_t seed) = 0;
    virtual void set_forcing_state(size_t m, const uint64_t *engine, const double *field) = 0;
    virtual void get_forcing_state(size_t m, uint64_t *engine, double *field) = 0;
    virtual void drift(double *out) = 0;
    virtual void linearized_drift(const double *directions, double *out) = 0;
    virtual void adjoint_drift(const double *cotangents, double *out) = 0;
    virtual void step_tangent(double dt, size_t steps, const double *tangents_in,
                              double *tangents_out) = 0;
    virtual void likelihood_gradient(const double *initial, double dt, size_t nobs,
                                     const size_t *ostep, const int *ox, const int *oy,
                                     const double *oval, const double *osd,
                                     const double *log_params, double *objective,
                                     double *gradient) = 0;
    virtual void likelihood_adjoint(const double *initial, double dt, size_t nobs,
                                    const size_t *ostep, const int *ox, const int *oy,
                                    const double *oval, const double *osd,
                                    const double *log_params, size_t stride, double *objective,
                                    double *gradient, size_t *counts) = 0;
    // launch-shape variants of the RK4 kernels (variant 0: the default)
    virtual int col_variants() const = 0;
    virtual int row_variants() const = 0;
    virtual void set_tuning(int col_variant, int row_variant) = 0;
    virtual void get_tuning(int *col_variant, int *row_variant) const = 0;
    virtual void tune(bool use_cache, double *col_ms, double *row_ms) = 0;
    virtual void step_forced(double dt, const double *forcing, bool shared) = 0;
    virtual void synchronize() = 0;
    virtual void diagnostics(double *out) = 0;
    // diagnostics through caller-owned device staging (B x 6 doubles)
    virtual void diagnostics_with(double *out, double *staging) = 0;
    virtual size_t members() const = 0;
    virtual int device() const = 0; // the device the handle's buffers live on
    virtual void energy_spectrum(double *out) = 0; // [B][n] shell bins
    virtual void profile_step(double dt, float *ms) = 0;
    virtual void pf_begin(uint64_t seed, size_t max_groups) = 0;
    virtual void pf_assimilate(double dt, size_t steps, size_t nobs, const int *x, const int *y,
                               const double *value, const double *sd, bool is_last,
                               double *record, double *mean_field) = 0;
    virtual void pf_result(double *log_likelihood, size_t *resampling_events) = 0;
    virtual void pf_begin_filters(const uint64_t *seeds, size_t filters, size_t max_groups,
                                  bool reset_keys) = 0;
    virtual size_t pf_filters() const = 0;
    virtual void pf_logweights(size_t nobs, const int *x, const int *y, const double *value,
                               const double *sd, const double *wprev, double *logw) = 0;
    // The same in two parts for a joint filter's captured cycle: check and
    // stage one observation group (eager), then launch the log-weight kernel
    // on the current mixed fields (capturable).
    virtual void pf_stage_logweights(size_t nobs, const int *x, const int *y,
                                     const double *value, const double *sd) = 0;
    virtual void pf_enqueue_logweights(size_t nobs, const double *wprev, double *logw) = 0;
    virtual void fill_state(const double *field) = 0;
    virtual void set_step_index(uint64_t step) = 0;
    virtual uint64_t step_index() const = 0;
    virtual void device_state(void **ptr, size_t *stride) = 0;
    virtual void *stream() = 0;
    // device views for other engine modules (surrogate forecast, EnKF)
    virtual EngineViews views() = 0;
    virtual void invalidate_mixed() = 0;
    // RK4 steps inside another module's launch sequence (which that module may
    // capture as a HIP graph): prepare_steps(dt) runs eagerly (time-step
    // checks, parameters, and the device step counter set to the current step
    // index); enqueue_steps(dt, steps) launches the mixed fields of the current
    // state and `steps` RK4 steps whose noise reads the step index from the
    // device counter, so a replay continues the stream; commit_steps(dt, steps)
    // then advances the host's step index.
    virtual void prepare_steps(double dt) = 0;
    virtual void enqueue_steps(double dt, size_t steps) = 0;
    virtual void commit_steps(double dt, size_t steps) = 0;
    // Captured launch sequences stay valid while this value is unchanged (it
    // moves whenever kernels, launch shapes, buffers or the forcing process
    // change), and only while graph replay is enabled.
    virtual uint64_t launch_epoch() const = 0;
    virtual bool graphs_enabled() const = 0;
    virtual void advance_counter(uint64_t steps) = 0;
    virtual void rewind_counter(uint64_t steps) = 0;
};

} // namespace hzl_amd

#include "hzl_amd.h"
#include <stdexcept>
#include <string>
#include <cstdlib>
#include <cstring>
namespace hzl_amd {
// Default of graph replay for new handles: on, unless the environment sets
// HZL_GRAPHS=0 (launch-by-launch, e.g. under a kernel tracer).
inline bool graphs_by_default() {
    const char *v = std::getenv("HZL_GRAPHS");
    return !(v && std::strcmp(v, "0") == 0);
}

// Engine behind a C-ABI handle (throws std::invalid_argument on null).
EngineBase &engine_of(hzl_ensemble_t h);

// Message of the last failed C-ABI call on the calling thread.
std::string &error_slot();

// Advance every member of `eng` over steps * dt with a trained surrogate
// network (csrc/surrogate.hip; as hzl_surrogate_forecast).
void surrogate_forecast(hzl_surrogate_t s, EngineBase &eng, double dt, size_t steps,
                        int precision);

// Run f, mapping exceptions to C-ABI codes: std::invalid_argument ->
// HZL_EINVAL, anything else -> HZL_ERUNTIME (message kept in error_slot()).
template <class F> int guarded(F &&f) {
    try {
        f();
        error_slot().clear();
        return HZL_OK;
    } catch (const std::invalid_argument &e) {
        error_slot() = e.what();
        return HZL_EINVAL;
    } catch (const std::exception &e) {
        error_slot() = e.what();
        return HZL_ERUNTIME;
    } catch (...) {
        error_slot() = "unknown error";
        return HZL_ERUNTIME;
    }
}
} // namespace hzl_amd
