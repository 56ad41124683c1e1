/*
 * hzl_amd.h -- C-ABI of the MI355X batched vorticity ensemble engine.
 *
 * Drop-in boundary for the solver hot path of pvasbob/HZL-stochasticFlowInference.
 * The reference exposes that path as C++ classes (no FFI of its own):
 *   hzl::CudaSolver  (include/hzl/cuda_solver.hpp:7-22, src/cuda_solver.cu:86-235)
 *   hzl::Solver      (include/hzl/solver.hpp:61-117, src/solver.cpp:120-254)
 *   hzl::ensemble    (include/hzl/inference.hpp:84-85, src/inference.cpp:555-590)
 * Each entry point below names the reference member it replaces.  A C++ layer
 * that restores those exact class signatures on top of this ABI lives in
 * hzl-stochasticflowinference_amd/csrc/host/hzl_gpu.{hpp,cpp} (libhzl_gpu.so).
 *
 * Conventions
 *   - every function returns HZL_OK (0) or an error code; the message of the
 *     last failure on the calling thread is returned by hzl_last_error().
 *     HZL_EINVAL corresponds to the reference's std::invalid_argument,
 *     HZL_ERUNTIME to std::runtime_error (src/cuda_solver.cu:11-19).
 *   - host pointers are borrowed for the duration of the call only.
 *   - a handle owns its device memory and one HIP stream; it is not re-entrant
 *     (one host thread at a time, like one CudaSolver object).
 *   - "field" arguments use the reference Field layout: n*n complex values as
 *     interleaved doubles (re, im), row-major index y*n + x, forward-normalised
 *     coefficients (physical(x) = sum_k coef(k) exp(i k.x); include/hzl/fft.hpp:7-9).
 *     The engine stores the retained 2/3 half plane internally in FP32 or FP64.
 */
#ifndef HZL_AMD_H
#define HZL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HZL_OK 0
#define HZL_EINVAL 1
#define HZL_ERUNTIME 2

#define HZL_FP32 0
#define HZL_FP64 1
#define HZL_BF16 2 /* surrogate inference operands only */

/* time integrators (reference hzl::TimeIntegrator, include/hzl/solver.hpp:12) */
#define HZL_RK4 0
#define HZL_SSPRK3 1
/* Forcing processes of the shell noise (hzl_ensemble_set_forcing_process). */
#define HZL_FORCING_WHITE 0
#define HZL_FORCING_OU 1

typedef struct hzl_ensemble *hzl_ensemble_t;
typedef struct hzl_surrogate *hzl_surrogate_t;
typedef struct hzl_pf_joint *hzl_pf_joint_t;

/* Model parameters.  The first six fields mirror hzl::Parameters
 * (include/hzl/solver.hpp:15-23).  The shell forcing acts on
 * forcing_min <= |k| <= forcing_max.  Under the default white-in-time process
 * its per-component increment variance is forcing_sigma^2 dt / 2 and
 * forcing_tau is not used; under the OU process (hzl_ensemble_set_forcing_process)
 * forcing_tau is every member's time constant until hzl_ensemble_set_member_tau
 * changes it.  A deterministic Kolmogorov term
 * kolmogorov_amplitude * sin(kolmogorov_wavenumber * y) is added. */
typedef struct hzl_parameters {
    double viscosity;
    double friction;
    double forcing_sigma;
    double forcing_tau;
    int forcing_min;
    int forcing_max;
    double kolmogorov_amplitude;
    int kolmogorov_wavenumber;
} hzl_parameters;

/* Message of the last failed call on this thread ("" if none). */
const char *hzl_last_error(void);

/* Check the model fields of params (the first six) for an n x n grid, with
 * the reference's rules and messages (HZL_EINVAL): nu, alpha, sigma finite and
 * nonnegative, tau finite and positive, 1 <= forcing_min <= forcing_max and
 * 3 forcing_max < n.  The grid itself is not checked.
 * Replaces: Parameters::validate (src/solver.cpp:31-40). */
int hzl_parameters_validate(size_t n, const hzl_parameters *params);

/* SplitMix64 mixing of (base, index + 1): the seed of stream `index` derived
 * from `base`.  Replaces: stream_seed (src/solver.cpp:61-66). */
uint64_t hzl_stream_seed(uint64_t base, uint64_t index);

/* Ensemble of `members` independent members on an n x n grid, n a power of
 * two in [16, 512].  Member m's random streams are keyed by
 * stream_seed(seed, m) (src/solver.cpp:61-66).  State starts at zero.
 * Replaces: CudaSolver::CudaSolver (src/cuda_solver.cu:126-171) batched over
 * members; validation as Parameters::validate (src/solver.cpp:31-40). */
int hzl_ensemble_create(hzl_ensemble_t *out, size_t n, size_t members, int precision,
                        const hzl_parameters *params, uint64_t seed);

/* Replaces: CudaSolver::~CudaSolver. */
int hzl_ensemble_destroy(hzl_ensemble_t h);

/* Per-member overrides of nu, alpha, sigma and Kolmogorov amplitude (used by
 * parameter inference where every member carries its own parameters). */
int hzl_ensemble_set_member_parameters(hzl_ensemble_t h, size_t member, double viscosity,
                                       double friction, double forcing_sigma,
                                       double kolmogorov_amplitude);

/* Per-member OU time constant (> 0, finite); used by the OU forcing process
 * from the next transition on (and by the stationary start that
 * hzl_pf_begin_filters draws).  Lets every member carry its own
 * hzl::Parameters, as the reference's per-proposal Solvers do
 * (src/filter.cpp:15-24,141-149). */
int hzl_ensemble_set_member_tau(hzl_ensemble_t h, size_t member, double tau);

/* Upload one member's spectral state (reference Field layout); non-retained
 * modes are dropped as the CudaSolver constructor masks its upload
 * (src/cuda_solver.cu:160-167).  member == (size_t)-1 uploads all members from
 * `field` holding members consecutive fields. */
int hzl_ensemble_set_state(hzl_ensemble_t h, size_t member, const double *field);

/* Download one member's state (or all with member == (size_t)-1).
 * Replaces: CudaSolver::coefficients (src/cuda_solver.cu:222-235), including
 * its non-finite check. */
int hzl_ensemble_get_state(hzl_ensemble_t h, size_t member, double *field);

/* Seeded Gaussian random field initial state for every member:
 * E(k) ~ k^4 exp(-(k/4)^2), enstrophy (A/2) sum |w|^2 = 1. */
int hzl_ensemble_init_grf(hzl_ensemble_t h);

/* `steps` RK4 steps of every member with the engine's own forcing (generated
 * on the device from (seed, member, step) counters).  Replaces:
 * CudaSolver::step(double dt) (src/cuda_solver.cu:172-183) and the per-member
 * loop of hzl::ensemble (src/inference.cpp:572-575).  Non-finite or diffusion-
 * unstable dt fails immediately (src/cuda_solver.cu:203-208); the CFL limit
 * dt * max(|u|+|v|) n / 2 pi <= 0.5 (src/solver.cpp:204-216) is evaluated on
 * the device at every step and reported by hzl_ensemble_synchronize. */
int hzl_ensemble_step(hzl_ensemble_t h, double dt, size_t steps);

/* `steps` steps of every member with the given integrator: HZL_RK4 as
 * hzl_ensemble_step, or HZL_SSPRK3, the Shu-Osher SSPRK(3,3) scheme with the
 * forcing held through its three stages.  Replaces: Solver::step(dt,
 * TimeIntegrator) (src/solver.cpp:217-254), batched; same time-step checks. */
int hzl_ensemble_step_integrator(hzl_ensemble_t h, double dt, size_t steps, int integrator);

/* Forcing process of the shell noise.  HZL_FORCING_WHITE (the default): white
 * in time, variance sigma^2 / (2 dt) per component held through the step.
 * HZL_FORCING_OU: the reference's Ornstein-Uhlenbeck shell forcing -- per member
 * an mt19937_64 NormalStream seeded with the member's stream seed
 * (stream_seed(seed, m) at creation; stream_seed(seed_f, i) after
 * hzl_pf_begin_filters), a stationary start with variance sigma^2 tau / 4 per
 * component drawn when the process is selected (and again when
 * hzl_pf_begin_filters resets the seeds), an exact half-step transition before
 * and after every step with the midpoint held through the stages, and
 * particle-filter children that keep their ancestor's coefficients and restart
 * their stream from stream_seed(seed, (g+1) P + i).  sigma and tau are per
 * member (hzl_ensemble_set_member_parameters, hzl_ensemble_set_member_tau;
 * the creation parameters' values until set).  Replaces: ShellForcing /
 * initialize_ou / transition_ou / Solver::reseed_future (src/solver.cpp:84-119,
 * 174-176,217-254; include/hzl/solver.hpp:73-75). */
int hzl_ensemble_set_forcing_process(hzl_ensemble_t h, int process);

/* Current OU forcing coefficients of member m (or of every member when
 * m == (size_t)-1) as reference Fields (n*n interleaved doubles, zero off the
 * shell pairs).  Replaces: Solver::forcing() (include/hzl/solver.hpp:83-85).
 * HZL_EINVAL unless the OU process is selected. */
int hzl_ensemble_get_forcing(hzl_ensemble_t h, size_t m, double *field);

/* Future innovations of member m come from stream seed `seed`: the member's
 * Philox key (white forcing) or a fresh mt19937_64(seed) NormalStream with the
 * OU coefficients kept (OU forcing).  Replaces: Solver::reseed_future
 * (include/hzl/solver.hpp:73-75); with it a member reproduces a reference
 * Solver(n, p, seed) whose forcing process is selected afterwards. */
int hzl_ensemble_reseed_future(hzl_ensemble_t h, size_t m, uint64_t seed);

/* Member m's OU forcing state: its mt19937_64 stream as 313 words (the 312
 * state words, then the position, the text order of std::mt19937_64 and of
 * HZL_CHECKPOINT_V1 files) and its forcing as a reference Field (n*n
 * interleaved doubles; the shell pairs are read, each pair's first mode).  A
 * stream taken from a checkpoint must not hold a cached second normal (the
 * reference's OU draws never leave one).  With set, a member resumes a
 * reference Solver's stochastic run draw for draw.  Replaces: the random stream
 * and forcing of Solver::load / Solver::save (src/solver.cpp:434-491).
 * HZL_EINVAL unless the OU process is selected. */
int hzl_ensemble_set_forcing_state(hzl_ensemble_t h, size_t m, const uint64_t *engine,
                                   const double *field);
int hzl_ensemble_get_forcing_state(hzl_ensemble_t h, size_t m, uint64_t *engine,
                                   double *field);

/* Calibration drifts of every member's current state, batched: out,
 * directions and cotangents hold one reference Field per member (members
 * n*n interleaved doubles; input modes outside the dealiasing mask are
 * dropped, as hzl_ensemble_set_state does).  Each member uses its own
 * viscosity and friction.
 *   hzl_ensemble_drift             Solver::drift (src/solver.cpp:255-263): the
 *                                  right-hand side without forcing
 *   hzl_ensemble_linearized_drift  Solver::linearized_drift
 *                                  (src/solver.cpp:264-290): its tangent along
 *                                  directions[m]
 *   hzl_ensemble_adjoint_drift     Solver::adjoint_drift (src/solver.cpp:291-323):
 *                                  the tangent's adjoint applied to
 *                                  cotangents[m], for Re sum conj(a) b */
int hzl_ensemble_drift(hzl_ensemble_t h, double *out);
int hzl_ensemble_linearized_drift(hzl_ensemble_t h, const double *directions, double *out);
int hzl_ensemble_adjoint_drift(hzl_ensemble_t h, const double *cotangents, double *out);

/* `steps` RK4 steps of every member together with its sensitivities to
 * (log viscosity, log friction): tangents_in / tangents_out hold per member
 * two reference Fields [member][2][n*n] (may alias).  Members must have no
 * forcing (sigma = 0, Kolmogorov amplitude 0), as the reference requires.
 * Replaces: Solver::step_tangent (src/solver.cpp:324-380), batched. */
int hzl_ensemble_step_tangent(hzl_ensemble_t h, double dt, size_t steps,
                              const double *tangents_in, double *tangents_out);

/* The same objective as hzl_likelihood_adjoint below, with its gradient by
 * forward sensitivities (step_tangent from zero tangents) instead of the
 * adjoint; member m at log_params[2m..2m+1].
 * Replaces: likelihood_gradient (src/inference.cpp:340-382), batched. */
int hzl_likelihood_gradient(hzl_ensemble_t h, const double *initial_field, double dt,
                            size_t nobs, const size_t *steps, const int *x, const int *y,
                            const double *value, const double *sd, const double *log_params,
                            double *objective, double *gradient);

/* Deterministic negative log-likelihood and its gradient with respect to
 * (log viscosity, log friction) by the discrete adjoint of RK4 with state
 * checkpoints every `stride` steps, for every member's own parameter vector
 * at once: member m runs nu = exp(log_params[2m]), alpha =
 * exp(log_params[2m + 1]) from initial_field (one reference Field) with no
 * forcing; observations are nobs rows (steps ascending, grid x, y, values,
 * standard deviations), as Dataset::observations.  objective[m] receives
 * sum 0.5 ((w(x, y) - value) / sd)^2, gradient[2m..2m+1] its gradient;
 * objective[m] = +inf and gradient 0 where the reference returns +inf
 * (parameters outside [-12, 0], or dt beyond the explicit diffusion limit).
 * counts (optional, 3): checkpoint fields, peak block fields, replayed steps.
 * The members' states, parameters and step counter are overwritten.
 * Replaces: likelihood_adjoint (src/adjoint.cpp:43-161), batched over
 * parameter vectors. */
int hzl_likelihood_adjoint(hzl_ensemble_t h, const double *initial_field, double dt,
                           size_t nobs, const size_t *steps, const int *x, const int *y,
                           const double *value, const double *sd, const double *log_params,
                           size_t stride, double *objective, double *gradient, size_t *counts);

/* Replay launch sequences from captured HIP graphs (default on; a handle
 * created while the environment holds HZL_GRAPHS=0 starts with replay off,
 * as does a surrogate's forecast graph).  A
 * hzl_ensemble_step of 2..200 steps (longer calls in chunks of 200), and a
 * whole particle-filter cycle of hzl_pf_assimilate with 1..200 steps
 * (forecast, log-weights, normalisation, resampling copies, which are
 * device-side no-ops when the filter keeps its particles), is recorded on
 * its second consecutive call with the same dt and sizes and replayed
 * afterwards, with the step index, group and last-group flag read from
 * device memory.  Results are identical to the launch-by-launch path (off).
 * No counterpart in the reference, which launches every kernel from the host
 * (src/cuda_solver.cu:172-218). */
int hzl_ensemble_set_graph(hzl_ensemble_t h, int enabled);

/* Launch-shape autotuner.  FP32 ensembles at 128^2, 256^2 and 512^2 can run
 * their column and row kernels in several compiled shapes (values per lane,
 * i.e. the T x E split of each transform; workgroup size; register budget);
 * variant 0 is the default shape and the only one in FP64 and at other
 * grids.  hzl_ensemble_create of an FP32 ensemble with members * n^2 >= 2^22
 * applies the shapes remembered for (device architecture, variant table,
 * n, member-count bucket) in the process or in the cache file
 * ($HZL_TUNE_CACHE, else $HOME/.cache/hzl_amd/launch_shapes.txt), else times
 * the variants on the zero initial state and remembers the fastest
 * (environment HZL_AUTOTUNE=0: keep the defaults).  hzl_ensemble_autotune
 * times them now and restores the states (use_cache != 0: apply a
 * remembered choice instead, if there is one); col_ms / row_ms (optional,
 * one entry per variant) receive the per-step kernel milliseconds of a
 * timing run and are left unchanged otherwise.  Variants
 * differ only in FP32 rounding.  No counterpart in the reference (cuFFT picks
 * its own plans, src/cuda_solver.cu:153-157). */
int hzl_ensemble_variant_count(hzl_ensemble_t h, int *col_variants, int *row_variants);
int hzl_ensemble_set_tuning(hzl_ensemble_t h, int col_variant, int row_variant);
int hzl_ensemble_get_tuning(hzl_ensemble_t h, int *col_variant, int *row_variant);
int hzl_ensemble_autotune(hzl_ensemble_t h, int use_cache, double *col_ms, double *row_ms);

/* One RK4 step with host-provided midpoint forcing held through the four
 * stages: `forcing` holds one field (shared != 0) or one per member.  Checks
 * finiteness and conjugate symmetry to 1e-12 as CudaSolver::step(dt, F)
 * (src/cuda_solver.cu:184-202). */
int hzl_ensemble_step_forced(hzl_ensemble_t h, double dt, const double *forcing, int shared);

/* Wait for queued work, then report a CFL violation or a non-finite state
 * seen by any step since the previous call.  Replaces: CudaSolver::synchronize
 * (src/cuda_solver.cu:219-221). */
int hzl_ensemble_synchronize(hzl_ensemble_t h);

/* Per-member diagnostics, 6 doubles each: energy, enstrophy, dissipation,
 * injection (against the forcing of the last step), cfl_rate of the current
 * state (max over the grid of (|u|+|v|) n / 2 pi), 0.  FP64 reductions on the
 * device.
 * Replaces: Solver::diagnostics (src/solver.cpp:381-397). */
int hzl_ensemble_diagnostics(hzl_ensemble_t h, double *out);

/* hzl_ensemble_diagnostics without any device allocation: the same kernels
 * and the same out[6 x members], staged through a fixed device array the
 * library holds on every GPU it runs on (3 MiB, for the life of the process)
 * instead of a buffer allocated per call.  At most 65536 members (HZL_EINVAL
 * beyond).  Calls on one GPU share that array: calls that overlap on the same
 * GPU (from several threads or handles) race and must be serialised by the
 * caller.  No counterpart in the reference; for callers that cannot allocate
 * device memory (tight serving or sampling loops, allocation-free harnesses). */
int hzl_ensemble_diagnostics_noalloc(hzl_ensemble_t h, double *out);

/* Shell energy spectrum of every member's current state: out[m * n + s]
 * receives bin s of member m, the sum of 2 pi^2 |w_k|^2 / |k|^2 over the modes
 * with floor(|k| + 1/2) = s (FP64 on the device, fixed summation order).
 * Replaces: Solver::energy_spectrum (src/solver.cpp:426-433), batched. */
int hzl_ensemble_energy_spectrum(hzl_ensemble_t h, double *out);

/* One engine-forcing RK4 step (as hzl_ensemble_step with steps = 1) with HIP
 * events around each launch; kernel_ms[8] receives the durations of
 * row(stage 1), col(stage 1), ..., row(stage 4), col(stage 4) in milliseconds.
 * Measurement-harness counterpart of src/cuda_ensemble_benchmark.cpp:73-79. */
int hzl_ensemble_profile_step(hzl_ensemble_t h, double dt, float *kernel_ms);

/* ------------------------------------------------------- particle filter --
 * Bootstrap particle filter over the ensemble's members as particles
 * (reference bootstrap_filter, src/filter.cpp:38-126; include/hzl/filter.hpp:17-23).
 * Members keep the keys stream_seed(seed_of_create, i); after a resampling
 * event at observation group g, child i is reseeded with
 * stream_seed(seed, (g + 1) * particles + i) (src/filter.cpp:112-115).
 * Systematic resampling uses one uniform per event from
 * mt19937_64(stream_seed(seed, UINT64_MAX)) (src/filter.cpp:54,102-110). */

/* Start a filter: weights 1/N, log-likelihood 0, for at most max_groups
 * observation groups that may resample (every group but the last):
 * hzl_pf_assimilate of a further non-final group fails with HZL_EINVAL. */
int hzl_pf_begin(hzl_ensemble_t h, uint64_t seed, size_t max_groups);

/* Advance every particle `steps` RK4 steps (engine forcing), then assimilate
 * one observation group (grid points x, y; values; standard deviations):
 * FP64 Gaussian log-weights, log-sum-exp increment, ESS, maximum weight, and
 * systematic resampling when ESS < N/2 unless is_last.  record[4 x filters]
 * receives per filter (increment, ESS, max weight, resampled 0/1);
 * mean_field (optional, single filter only, reference Field layout) the
 * weighted mean state with this group's weights. */
int hzl_pf_assimilate(hzl_ensemble_t h, double dt, size_t steps, size_t nobs, const int *x,
                      const int *y, const double *value, const double *sd, int is_last,
                      double *record, double *mean_field);

/* Accumulated log-likelihood (sum of increments; -inf once every weight was
 * -inf, src/filter.cpp:76-80) and number of resampling events, one entry per
 * filter. */
int hzl_pf_result(hzl_ensemble_t h, double *log_likelihood, size_t *resampling_events);

/* `filters` independent particle filters over contiguous member segments of
 * members / filters particles (one filter per chain of a batched particle
 * MCMC, src/filter.cpp:166-203).  Filter f behaves as hzl_pf_begin(seeds[f])
 * on its own particles, whose noise keys restart at stream_seed(seeds[f], i)
 * (the reference's per-particle Solver(n, p, stream_seed(seed, i)),
 * src/filter.cpp:44-48). */
int hzl_pf_begin_filters(hzl_ensemble_t h, const uint64_t *seeds, size_t filters,
                         size_t max_groups);

/* One bootstrap particle filter over the members of up to 4 ensembles of
 * different grid sizes and one precision (e.g. 256 particles at 256^2 plus
 * 32 at 512^2).  Each ensemble advances on its own stream; weights are
 * normalised over the union; a child copies the retained modes its grid
 * shares with its ancestor's (truncation / zero padding).  Observation grid
 * indices refer to the coarsest grid.  Particle i of the union starts on
 * noise key stream_seed(seed, i); everything else as hzl_pf_begin /
 * hzl_pf_assimilate (record[4]) / hzl_pf_result.  The ensembles must not be
 * used for other work while the filter runs, and run the white forcing
 * process (HZL_EINVAL at begin for an OU ensemble). */
int hzl_pf_joint_create(hzl_pf_joint_t *out, const hzl_ensemble_t *ensembles, size_t count);
int hzl_pf_joint_destroy(hzl_pf_joint_t h);
int hzl_pf_joint_begin(hzl_pf_joint_t h, uint64_t seed, size_t max_groups);
int hzl_pf_joint_assimilate(hzl_pf_joint_t h, double dt, size_t steps, size_t nobs, const int *x,
                            const int *y, const double *value, const double *sd, int is_last,
                            double *record);
int hzl_pf_joint_result(hzl_pf_joint_t h, double *log_likelihood, size_t *resampling_events);

/* Upload one state (reference Field layout) into every member. */
int hzl_ensemble_fill_state(hzl_ensemble_t h, const double *field);

/* Set the step counter (the noise counter of the next step), e.g. back to 0
 * when a likelihood evaluation restarts from the initial state. */
int hzl_ensemble_set_step_index(hzl_ensemble_t h, uint64_t step);

/* ------------------------------------------------------------------------
 * Ensemble Kalman filter (members are the ensemble), reference
 * ensemble_kalman_filter (src/filter.cpp:204-356).
 * ------------------------------------------------------------------------ */
#define HZL_ENKF_SERIAL 0 /* scalar observations in order, as the reference */
#define HZL_ENKF_BATCH 1  /* all observations of a group at once (MFMA, FP32) */

/* Start a filter: observation perturbations from the reference NormalStream
 * seeded stream_seed(seed, members) (src/filter.cpp:224).  With the OU forcing
 * process the serial analysis carries each member's forcing as a latent state
 * (inflated and updated with its own covariance, src/filter.cpp:237-321);
 * HZL_ENKF_BATCH needs the white forcing process (HZL_EINVAL otherwise). */
int hzl_enkf_begin(hzl_ensemble_t h, uint64_t seed, int mode);

/* Advance every member `steps` RK4 steps (engine forcing), apply the
 * multiplicative anomaly inflation (>= 1, src/filter.cpp:237-253), then
 * assimilate one observation group.  record[2] receives (forecast sensor
 * RMSE of the ensemble mean before the analysis, posterior spread
 * sqrt(sum |w - mean|^2 / (M - 1)) over the full field); mean_field
 * (optional, reference Field layout) the analysis mean.  With
 * localization_radius > 0 (serial mode) every covariance is tapered by the
 * Gaspari-Cohn function of half-support radius grid cells around its
 * observation (src/filter.cpp:25-36,290-308). */
int hzl_enkf_assimilate(hzl_ensemble_t h, double dt, size_t steps, size_t nobs, const int *x,
                        const int *y, const double *value, const double *sd, double inflation,
                        double localization_radius, double *record, double *mean_field);

/* Ensemble mean (reference Field layout, optional) and spread, as recorded
 * by the reference's free-forecast phase (src/filter.cpp:333-355). */
int hzl_ensemble_mean_spread(hzl_ensemble_t h, double *spread, double *mean_field);

/* ------------------------------------------------------------------------
 * Batched particle-marginal Metropolis-Hastings, reference
 * particle_metropolis (src/filter.cpp:166-203), one chain per filter segment,
 * and batched particle likelihoods (particle_log_likelihood,
 * src/filter.cpp:141-149).
 * ------------------------------------------------------------------------ */
/* Parameter vectors and their priors. */
#define HZL_PRIOR_LOG_UNIFORM 0 /* q = (log nu, log Kolmogorov amplitude), uniform in
                                   [log_lower, log_upper]; alpha and sigma fixed */
#define HZL_PRIOR_STOCHASTIC 1  /* q = (log nu, log alpha, log sigma, log tau) with the
                                   reference's stochastic_log_prior
                                   (src/filter.cpp:127-140); no Kolmogorov term */

typedef struct hzl_pmcmc_options {
    size_t chains;         /* members / chains particles per chain's filter */
    size_t samples;        /* Metropolis steps per chain */
    double proposal_scale; /* random-walk step on every component of q */
    double log_lower[2];   /* HZL_PRIOR_LOG_UNIFORM: bounds of log nu, log amplitude */
    double log_upper[2];
    double friction;       /* HZL_PRIOR_LOG_UNIFORM: fixed alpha and forcing sigma */
    double forcing_sigma;
    uint64_t seed;         /* chain c's stream seed is stream_seed(seed, c) ... */
    int prior;             /* HZL_PRIOR_* */
    const uint64_t *chain_seeds; /* ... or chain_seeds[c] when not null */
    hzl_surrogate_t surrogate;   /* optional: forecast every observation interval with
                                    this trained network instead of the solver (white
                                    forcing; every interval must equal its horizon) */
    int surrogate_precision;     /* HZL_BF16 or HZL_FP32 operands of that forecast */
} hzl_pmcmc_options;

/* The first `count` draws of the reference's NormalStream(seed): mt19937_64,
 * open-interval uniforms ((x >> 11) + 0.5) 2^-53, Box-Muller returning
 * r cos and then the kept r sin.  Replaces: NormalStream (src/solver.cpp:41-54)
 * where a caller consumes one stream in order (synthetic observations). */
int hzl_normal_draws(uint64_t seed, size_t count, double *out);

/* Log prior density of the reference's stochastic parameter vector
 * (log nu, log alpha, log sigma, log tau): -inf outside its bounds.  HZL_EINVAL
 * with the reference's message unless dims == 4.
 * Replaces: stochastic_log_prior (src/filter.cpp:127-140). */
int hzl_stochastic_log_prior(const double *q, size_t dims, double *out);

/* Likelihood evaluations with options->surrogate set forecast each interval
 * with the network for the whole run: a pseudo-marginal chain must target one
 * likelihood, so the forecast model is chosen per run (judge it first with
 * hzl_surrogate_judge), never switched between evaluations. */

/* Log-likelihood estimates of `filters` independent bootstrap filters over
 * the ensemble's members (members / filters particles each): filter f runs
 * at parameter vector q[f * dims ..] (dims 2 or 4 by options->prior; the
 * options' chains, samples, proposal and seed fields are not used) with
 * filter seed seeds[f], its particles restarting from initial_field at step 0
 * on noise keys stream_seed(seeds[f], i).  -inf where q lies outside the
 * prior or beyond the explicit-step limit dt (2 nu K^2 + alpha) <= 2.5.
 * Observation rows must be time ordered from step 0.  With the OU forcing
 * process selected this is the reference's particle_log_likelihood for every
 * filter at once (src/filter.cpp:141-149; bootstrap_filter(.., false)). */
int hzl_pf_log_likelihoods(hzl_ensemble_t particles, const hzl_pmcmc_options *options,
                           double dt, size_t nobs, const int *step, const int *x, const int *y,
                           const double *value, const double *sd, const double *initial_field,
                           size_t filters, const double *q, const uint64_t *seeds,
                           double *log_likelihood);

/* Run every chain `samples` steps over one window of observations (rows time
 * ordered from step 0), every likelihood evaluation restarting all
 * particles from initial_field at step 0.  initial_q[dims x chains] holds
 * each chain's starting point.  Outputs per chain c and sample s at index
 * k = c * samples + s: samples_q[dims x ...] the chain state after the step,
 * samples_ll its retained likelihood estimate, samples_density (optional)
 * log-likelihood + log-prior, samples_accepted (optional) 0/1, samples_seed
 * (optional) the filter seed of the retained estimate.  With one chain,
 * chain_seeds = {seed}, HZL_PRIOR_STOCHASTIC and the OU forcing process, this
 * is the reference's particle_metropolis(data, q, scale, samples, particles,
 * seed) draw for draw.  HZL_EINVAL "initial particle posterior estimate is
 * not finite" when a chain's starting density is not finite. */
int hzl_pmcmc_run(hzl_ensemble_t particles, const hzl_pmcmc_options *options, double dt,
                  size_t nobs, const int *step, const int *x, const int *y, const double *value,
                  const double *sd, const double *initial_field, const double *initial_q,
                  double *samples_q, double *samples_ll, double *samples_density,
                  int *samples_accepted, uint64_t *samples_seed);

/* Number of steps taken so far (the noise counter of the next step). */
int hzl_ensemble_step_index(hzl_ensemble_t h, uint64_t *out);

/* Device pointer of member 0's retained half-plane state ([c][j], K = (n-1)/3,
 * c in [0,K], j = ky + K in [0,2K], complex float or double) and the byte
 * stride between members; for zero-copy consumers on the same device. */
int hzl_ensemble_device_state(hzl_ensemble_t h, void **ptr, size_t *member_stride_bytes);

/* The handle's HIP stream (hipStream_t) for callers that order their own work
 * after the engine's or time it with events. */
int hzl_ensemble_stream(hzl_ensemble_t h, void **stream);

/* ------------------------------------------------------------------------
 * Learned forecast surrogate (spectral neural operator).  No counterpart in
 * the reference: it replaces the particle forecast of the filter loops
 * (src/filter.cpp:46-48 advances every particle with the solver) by one
 * network evaluation per observation interval.  The network sees the state
 * on an internal grid x grid lattice (exact truncated DFT of the band
 * |kx|, |ky| <= grid/2 - 1), log viscosity and the member's Kolmogorov
 * forcing (amplitude / 0.1 times its pattern), and
 * predicts the increment over a fixed horizon of `horizon_steps` solver
 * steps; modes outside the band decay linearly, and the shell forcing adds
 * white noise of variance forcing_sigma^2 h / 2 per component.
 * ------------------------------------------------------------------------ */

/* state_n: the ensemble grid it will serve (64, 128 or 256); grid 32 or 64;
 * width (channels, multiple of 8, <= 64); layers 1..8; modes kept per axis
 * (kx in [0, modes), ky in [-modes, modes)), modes <= grid / 4.  Weights are
 * drawn from mt19937_64(seed). */
int hzl_surrogate_create(hzl_surrogate_t *out, size_t state_n, int grid, int width, int layers,
                         int modes, uint64_t seed);
int hzl_surrogate_destroy(hzl_surrogate_t h);

/* The FP32 parameter vector (layout documented in csrc/surrogate.hip). */
int hzl_surrogate_param_count(hzl_surrogate_t h, size_t *count);
int hzl_surrogate_get_params(hzl_surrogate_t h, float *params);
int hzl_surrogate_set_params(hzl_surrogate_t h, const float *params);

/* Input scale (grid w is multiplied by it) and output scale (the network
 * predicts increment * output_scale).  Set from the first collected window
 * unless given here. */
int hzl_surrogate_set_scales(hzl_surrogate_t h, double input_scale, double output_scale);
int hzl_surrogate_get_scales(hzl_surrogate_t h, double *scales);

/* Network output (count x grid x grid doubles) for `count` states in the
 * reference Field layout with per-sample viscosities and Kolmogorov
 * amplitudes (amplitude may be null: 0.1 for every sample); precision
 * HZL_FP32 or HZL_BF16 operands. */
int hzl_surrogate_forward(hzl_surrogate_t h, int precision, size_t count, const double *fields,
                          const double *viscosity, const double *amplitude, double *grid_out);

/* Mean squared error between the network output and the scaled grid
 * increment next - fields, and its gradient with respect to every parameter
 * (FP32 operands). */
int hzl_surrogate_loss_grad(hzl_surrogate_t h, size_t count, const double *fields,
                            const double *next_fields, const double *viscosity,
                            const double *amplitude, double *loss, float *grad);

/* Append windows x members training pairs from an FP32 ensemble of the
 * surrogate's state_n: each window copies the states as inputs and advances
 * the ensemble horizon_steps steps of dt with its own forcing.  Targets are
 * the advanced states, or with noise_free_targets the same inputs advanced
 * without the white shell forcing (the deterministic propagator, which the
 * forecast then complements with the forcing noise). */
int hzl_surrogate_collect(hzl_surrogate_t h, hzl_ensemble_t ensemble, double dt,
                          size_t horizon_steps, size_t windows, int noise_free_targets);
int hzl_surrogate_dataset_size(hzl_surrogate_t h, size_t *pairs);

/* Adam on pairs [first, first + count): `epochs` passes of shuffled
 * minibatches (mt19937_64(seed)), cosine learning rate from learning_rate to
 * learning_rate / 20; epoch_loss (optional) receives each epoch's mean loss. */
int hzl_surrogate_fit(hzl_surrogate_t h, size_t first, size_t count, size_t epochs, size_t batch,
                      double learning_rate, uint64_t seed, double *epoch_loss);

/* Mean loss over pairs [first, first + count), forward only.  Also sets the
 * error estimate: the grid RMS error of the predicted increment against the
 * solver's, sqrt(loss) / output_scale, in vorticity units. */
int hzl_surrogate_evaluate(hzl_surrogate_t h, int precision, size_t first, size_t count,
                           double *loss);

/* Replay forecasts from a captured HIP graph (default on): the second
 * forecast of an ensemble with unchanged scales, horizon and buffers records
 * the launch sequence, later ones replay it with the current step counter.
 * Off: every forecast launches its kernels one by one. */
int hzl_surrogate_set_graph(hzl_surrogate_t h, int enabled);

/* BF16 inference path of a network of the default shape (64^2 grid, width
 * 32, 16 x 32 modes, 3 input features): enabled (the default), each spectral
 * layer runs as one fused forward-transform pass and one fused inverse-
 * transform pass (MFMA, one workgroup per sample, intermediates in LDS) around
 * the per-mode mixing product; disabled, as the five batched products of the
 * training path.  Both round the same intermediates to BF16; the fused passes
 * evaluate GELU by a sigmoid form within 2.6e-5 of the exact erf form.  No
 * counterpart in the reference. */
int hzl_surrogate_set_fused(hzl_surrogate_t h, int enabled);

/* The error estimate (negative until evaluated or set). */
int hzl_surrogate_error_estimate(hzl_surrogate_t h, double *rms);
int hzl_surrogate_set_error_estimate(hzl_surrogate_t h, double rms);

/* Grid RMS deviation of an FP32 ensemble's members from their mean. */
int hzl_surrogate_spread(hzl_surrogate_t h, hzl_ensemble_t ensemble, double *rms);

/* Surrogate forecast with plain-solver fallback: after the network forecast,
 * when the error estimate exceeds threshold x the spread of the forecast
 * ensemble, the states are restored and the engine advances the ensemble
 * `steps` RK4 steps with its own forcing instead (used_solver = 1; identical
 * to hzl_ensemble_step).  spread receives the forecast ensemble's spread. */
int hzl_surrogate_forecast_or_solve(hzl_surrogate_t h, hzl_ensemble_t ensemble, double dt,
                                    size_t steps, int precision, double threshold,
                                    int *used_solver, double *spread);

/* Live judge of the network's forecast error, with the plain solver as the
 * judge: the first members(judge) members of `ensemble` are copied (states,
 * nu, alpha, Kolmogorov amplitude) into `judge`, an FP32 ensemble on the same
 * grid run without shell noise; there the same states are forecast over
 * steps * dt by the network and by `steps` plain RK4 steps.  The grid RMS
 * difference per member (vorticity units) becomes the error estimate that
 * hzl_surrogate_forecast_or_solve compares with the ensemble spread, and is
 * returned in *rms.  `ensemble` is not changed. */
int hzl_surrogate_judge(hzl_surrogate_t h, hzl_ensemble_t ensemble, hzl_ensemble_t judge,
                        double dt, size_t steps, int precision, double *rms);

/* The same judge, enqueued without waiting: judge's stream waits for the
 * work already queued on ensemble's stream and copies the judged states;
 * ensemble's stream waits only for that copy.  The network forecast of the
 * copies is ordered with the ensemble's forecasts; the solver run and the
 * comparison overlap the ensemble's later work.  Parameters are taken from
 * the ensemble's host-side values.  One judge is in flight per surrogate: a
 * start (or hzl_surrogate_judge) first finishes the previous one.  No
 * counterpart in the reference: lets a surrogate filter keep judging without
 * stalling its cycles. */
int hzl_surrogate_judge_start(hzl_surrogate_t h, hzl_ensemble_t ensemble, hzl_ensemble_t judge,
                              double dt, size_t steps, int precision);

/* Fold a judge in flight into the error estimate.  wait != 0 blocks until it
 * has finished; wait == 0 returns at once.  *finished = 1 when a result was
 * folded in by this call (0: none in flight, or not finished yet); *rms
 * receives the current error estimate (-1 before any). */
int hzl_surrogate_judge_poll(hzl_surrogate_t h, int wait, int *finished, double *rms);

/* Advance every member of an FP32 ensemble by steps * dt with the network
 * (steps and dt must equal the trained horizon).  The ensemble's step counter
 * advances by `steps`; filter calls may follow with steps = 0.  The ensemble
 * runs the white forcing process (HZL_EINVAL for OU; also for
 * hzl_surrogate_forecast_or_solve).  Forecasts of one surrogate share its
 * scratch buffers: a forecast of another ensemble (another stream) waits on
 * the device for the previous forecast to finish.  Up to four forecast launch
 * sequences (per ensemble, horizon and precision) are captured and replayed
 * when graphs are on. */
int hzl_surrogate_forecast(hzl_surrogate_t h, hzl_ensemble_t ensemble, double dt, size_t steps,
                           int precision);

#ifdef __cplusplus
}
#endif
#endif
