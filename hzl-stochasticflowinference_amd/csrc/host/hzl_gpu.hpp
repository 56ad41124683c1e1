// hzl_gpu.hpp -- C++ drop-in for the solver / filter path of
// pvasbob/HZL-stochasticFlowInference, implemented over the C-ABI in
// include/hzl_amd.h (libhzl_amd.so, gfx950 kernels).
//
// The class and function signatures are those of the reference headers, so a
// caller of
//   hzl::CudaSolver             (include/hzl/cuda_solver.hpp:7-22)
//   hzl::Parameters, Diagnostics, stream_seed        (include/hzl/solver.hpp:15-40)
//   hzl::Fourier2D, wave_number (include/hzl/fft.hpp:10-25)
//   hzl::Observation, Dataset   (include/hzl/inference.hpp:33-49)
//   hzl::ensemble               (include/hzl/inference.hpp:84-85)
//   hzl::bootstrap_filter, ensemble_kalman_filter, particle_log_likelihood,
//   averaged_particle_log_likelihood, particle_metropolis, stochastic_log_prior
//   and their records           (include/hzl/filter.hpp:5-54)
// compiles and links against libhzl_gpu.so instead of the reference library.
// The reference-signature solver functions (bootstrap_filter,
// ensemble_kalman_filter, ensemble and the particle-likelihood family) run the
// reference's own forcing process: the Ornstein-Uhlenbeck shell forcing with
// one mt19937_64 NormalStream per member seeded as the reference seeds its
// Solvers (the EnKF analyses it as a latent state).  Differences of behaviour:
//   * the batch (MFMA) EnKF analysis of an hzl::Ensemble updates the vorticity
//     only and needs the white-in-time forcing;
//   * hzl::Ensemble defaults to the white-in-time forcing (variance
//     sigma^2 / (2 dt) per component held through the step); select the OU
//     process with set_forcing_process;
//   * CudaSolver's CFL check is read when the step's results are read
//     (synchronize / coefficients) instead of after every step.
// hzl::Ensemble (GPU batches of members) has no reference counterpart: it
// replaces the reference's std::vector<Solver> loops.
#pragma once
#include "hzl_amd.h"

#include <complex>
#include <cstddef>
#include <cstdint>
#include <filesystem>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace hzl {

using Complex = std::complex<double>;
using Field = std::vector<Complex>;

// Host 2D FFT with the reference normalisation: physical(x) = sum_k
// coefficient(k) exp(ikx); forward divides by n^2, inverse does not.
class Fourier2D {
  public:
    explicit Fourier2D(std::size_t n);
    [[nodiscard]] std::size_t size() const noexcept {
        return n_;
    }
    void forward(Field &data);
    void inverse(Field &data);

  private:
    void lines(Field &data, bool columns, bool inverse);
    std::size_t n_;
    std::vector<std::size_t> bitrev_;
    Field twiddle_, line_;
};
[[nodiscard]] int wave_number(std::size_t index, std::size_t n);

enum class TimeIntegrator { rk4, ssprk3 };

struct Parameters {
    double viscosity = 0.01;
    double friction = 0.1;
    double forcing_sigma = 0.2;
    double forcing_tau = 0.5;
    int forcing_min = 2;
    int forcing_max = 3;
    void validate(std::size_t n) const;
};
struct Diagnostics {
    double time{}, energy{}, enstrophy{}, dissipation{}, injection{}, cfl_rate{};
};
std::uint64_t stream_seed(std::uint64_t base, std::uint64_t index);

// Named initial conditions of Solver::initialize (taylor-green, multimode,
// zero) as masked spectral coefficients, and Solver::initialize_physical.
Field initial_condition(std::size_t n, const std::string &name);
Field physical_to_spectral(std::size_t n, const Field &vorticity);

class CudaSolver {
  public:
    CudaSolver(std::size_t n, Parameters parameters, const Field &initial);
    ~CudaSolver();
    CudaSolver(const CudaSolver &) = delete;
    CudaSolver &operator=(const CudaSolver &) = delete;
    void step(double dt);
    void step(double dt, const Field &midpoint_forcing);
    void synchronize();
    [[nodiscard]] Field coefficients() const;

  private:
    struct Implementation;
    std::unique_ptr<Implementation> implementation_;
};

enum class Precision { fp32, fp64 };
enum class ForcingProcess { white, ornstein_uhlenbeck };

// A batch of members on one grid in one engine handle.  Members start at
// zero; member m's noise stream is keyed by stream_seed(seed, m).
class Ensemble {
  public:
    Ensemble(std::size_t n, std::size_t members, Parameters parameters,
             Precision precision = Precision::fp64, std::uint64_t seed = 0,
             double kolmogorov_amplitude = 0.0, int kolmogorov_wavenumber = 4);
    ~Ensemble();
    Ensemble(const Ensemble &) = delete;
    Ensemble &operator=(const Ensemble &) = delete;
    void set_member_parameters(std::size_t member, double viscosity, double friction,
                               double forcing_sigma, double kolmogorov_amplitude);
    void set_state(std::size_t member, const Field &coefficients);
    void set_all(const Field &coefficients); // every member
    [[nodiscard]] Field state(std::size_t member) const;
    void step(double dt, std::size_t steps = 1);
    // Solver::step(dt, integrator) (src/solver.cpp:217-254) for every member
    void step(double dt, std::size_t steps, TimeIntegrator integrator);
    void step(double dt, const Field &shared_midpoint_forcing);
    // Solver::drift / linearized_drift / adjoint_drift (src/solver.cpp:255-323)
    // of every member's state; one Field per member in and out
    [[nodiscard]] std::vector<Field> drift();
    [[nodiscard]] std::vector<Field> linearized_drift(const std::vector<Field> &directions);
    [[nodiscard]] std::vector<Field> adjoint_drift(const std::vector<Field> &cotangents);
    void synchronize();
    [[nodiscard]] std::vector<Diagnostics> diagnostics();
    // Solver::energy_spectrum (src/solver.cpp:426-433) of every member: n bins each
    [[nodiscard]] std::vector<std::vector<double>> energy_spectra();
    // Shell forcing process: white in time (default) or the reference's OU
    // process (ShellForcing, src/solver.cpp:84-119); selecting OU draws every
    // member's stationary start from its stream.
    void set_forcing_process(ForcingProcess process);
    // Solver::forcing() of one member (OU process only)
    [[nodiscard]] Field forcing(std::size_t member) const;
    // Solver::reseed_future (include/hzl/solver.hpp:73-75)
    void reseed_future(std::size_t member, std::uint64_t seed);
    // A member's OU stream (312 words and the position, Checkpoint::random_engine)
    // and forcing (OU process only): with a Checkpoint's, the member resumes
    // the checkpointed Solver's stochastic run draw for draw.
    void set_forcing_state(std::size_t member, const std::vector<std::uint64_t> &engine,
                           const Field &forcing);
    [[nodiscard]] std::pair<std::vector<std::uint64_t>, Field>
    forcing_state(std::size_t member) const;
    [[nodiscard]] std::size_t size() const noexcept {
        return n_;
    }
    [[nodiscard]] std::size_t members() const noexcept {
        return members_;
    }
    [[nodiscard]] hzl_ensemble_t handle() const noexcept {
        return handle_;
    }

  private:
    std::size_t n_, members_;
    hzl_ensemble_t handle_ = nullptr;
};

// HZL_CHECKPOINT_V1 text files, the format of Solver::save / Solver::load
// (src/solver.cpp:434-491): grid, time, the first six Parameters fields, the
// solver's NormalStream (mt19937_64 state as std::mt19937_64's text form: 312
// words and the position, then has_spare and spare; src/solver.cpp:54-60),
// and per spectral index y * n + x the vorticity and the OU forcing.  load()
// applies the reference's checks (finite values, dealiasing mask, forcing only
// on the shell pairs, conjugate symmetry) and errors; save() writes the same
// bytes the reference would.  The random stream is carried verbatim: the GPU
// engine's forcing noise is counter-based and does not consume it.
struct Checkpoint {
    std::size_t n = 0;
    double time = 0;
    Parameters parameters;
    std::vector<std::uint64_t> random_engine; // 312 state words + position
    bool random_has_spare = false;
    double random_spare = 0;
    Field vorticity, forcing; // spectral, n * n each
    void save(const std::filesystem::path &path) const;
    static Checkpoint load(const std::filesystem::path &path);
};

struct Observation {
    std::size_t step{}, x{}, y{};
    double value{}, standard_deviation{};
};
// Time-ordered point observations; HZL_OBSERVATIONS_V2 CSV files
// (src/inference.cpp:197-255; V1 files load too).
struct Dataset {
    std::size_t n = 16;
    double dt = 0.01;
    bool stochastic = false;
    int forcing_min = 2, forcing_max = 3;
    std::vector<Observation> observations;
    void validate() const;
    void save(const std::filesystem::path &path) const;
    static Dataset load(const std::filesystem::path &path);
};

// Observations of a truth run with the reference's OU forcing on the GPU
// (synthetic_observations, src/inference.cpp:256-288): Solver(n, truth,
// stream_seed(seed, 0)) from "multimode", eight fixed sensors every `every`
// steps with noise from NormalStream(stream_seed(seed, 1)); truth_fields
// (optional) receives the physical fields at those times.
Dataset synthetic_observations(std::size_t n, double dt, std::size_t steps, std::size_t every,
                               Parameters truth, double noise, std::uint64_t seed,
                               std::vector<Field> *truth_fields = nullptr);

struct FilterRecord {
    std::size_t step{};
    double log_likelihood_increment{}, effective_particles{}, maximum_weight{};
    Field mean_vorticity; // physical, as the reference records it
};
struct FilterResult {
    double log_likelihood = 0;
    std::size_t resampling_events = 0;
    std::vector<FilterRecord> history;
};
// Bootstrap particle filter with systematic resampling below N/2 ESS;
// particles start from initialize("multimode") (src/filter.cpp:38-126).
FilterResult bootstrap_filter(const Dataset &data, Parameters p, std::size_t particles,
                              std::uint64_t seed, bool retain_states = true);
// The same filter over an initialised GPU ensemble (members are particles).
FilterResult bootstrap_filter(Ensemble &particles, const Dataset &data, std::uint64_t seed,
                              bool retain_states = true);

struct EnsembleFilterRecord {
    std::size_t step{};
    double forecast_sensor_rmse{}, posterior_spread{};
    Field mean_vorticity; // physical
    bool is_forecast = false;
};
struct EnKFOptions {
    double inflation = 1.0;
    double localization_radius = 0.0; // Gaspari-Cohn half support, grid cells (serial)
    std::size_t forecast_steps = 0;
    std::size_t forecast_every = 1;
};
// Serial perturbed-observation EnKF; members start from initialize("multimode")
// with OU forcing, analysed together with the vorticity (src/filter.cpp:204-356).
std::vector<EnsembleFilterRecord> ensemble_kalman_filter(const Dataset &data, Parameters p,
                                                         std::size_t members, std::uint64_t seed,
                                                         EnKFOptions options = {});
// The same filter over an initialised GPU ensemble; batch = true assimilates
// each observation group at once (MFMA analysis, FP32 ensembles).
std::vector<EnsembleFilterRecord> ensemble_kalman_filter(Ensemble &members, const Dataset &data,
                                                         std::uint64_t seed,
                                                         EnKFOptions options = {},
                                                         bool batch = false);

using Vector = std::vector<double>;

// Particle-marginal inference over q = (log nu, log alpha, log sigma, log tau)
// (src/filter.cpp:127-203): Gaussian log prior with bounds, the bootstrap
// filter's log-likelihood (OU forcing, multimode particles, FP64 on the GPU),
// its log-mean-exp over replicates, and the pseudo-marginal random-walk
// Metropolis chain with the reference's streams.
struct ParticleSample {
    Vector parameters;
    double log_likelihood{}, log_posterior{};
    std::uint64_t likelihood_seed{};
    bool accepted{};
};
double stochastic_log_prior(const Vector &log_parameters);
double particle_log_likelihood(const Dataset &data, const Vector &log_parameters,
                               std::size_t particles, std::uint64_t seed);
double averaged_particle_log_likelihood(const Dataset &data, const Vector &log_parameters,
                                        std::size_t particles, std::size_t replicates,
                                        std::uint64_t seed);
std::vector<ParticleSample> particle_metropolis(const Dataset &data, Vector initial,
                                                double proposal_scale, std::size_t count,
                                                std::size_t particles, std::uint64_t seed);

struct ValueGradient {
    double value{};
    Vector gradient;
};
struct AdjointEvaluation {
    ValueGradient objective;
    std::size_t checkpoint_fields = 0, peak_block_fields = 0, replayed_steps = 0;
};
// Deterministic negative log-likelihood and its gradient with respect to
// (log viscosity, log friction) by the discrete adjoint of RK4
// (src/adjoint.cpp:43-161), from initialize("multimode"), on the GPU in FP64.
AdjointEvaluation likelihood_adjoint(const Dataset &data, const Vector &log_parameters,
                                     std::size_t checkpoint_stride = 20);
// The same for many parameter vectors at once (one GPU member each).
std::vector<AdjointEvaluation> likelihood_adjoint(const Dataset &data,
                                                  const std::vector<Vector> &log_parameters,
                                                  std::size_t checkpoint_stride = 20);
// The same objective with its gradient by forward sensitivities
// (Solver::step_tangent; src/inference.cpp:340-382), on the GPU in FP64.
ValueGradient likelihood_gradient(const Dataset &data, const Vector &log_parameters);

// Independent members from the zero state, final diagnostics in member order
// (src/inference.cpp:555-590); `workers` is accepted and unused.
std::vector<Diagnostics> ensemble(std::size_t n, Parameters p, double dt, std::size_t steps,
                                  std::size_t count, std::size_t workers, std::uint64_t seed);

} // namespace hzl
