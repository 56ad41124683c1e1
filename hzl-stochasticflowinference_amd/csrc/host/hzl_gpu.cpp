// hzl_gpu.cpp -- C++ drop-in layer over the C-ABI (see hzl_gpu.hpp).
//
// Everything that computes runs in libhzl_amd.so: the solver, filters and
// samplers on the GPU, and the reference's parameter rules, stream seeding and
// prior behind the C-ABI.  This file converts between the reference's types
// (Field, Parameters, Dataset, records) and the ABI's flat arrays, and holds
// the two text formats (HZL_OBSERVATIONS_V2, HZL_CHECKPOINT_V1) through the
// shared encodings of text_io.hpp.
#include "hzl_gpu.hpp"
#include "text_io.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace hzl {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kNegativeInfinity = -std::numeric_limits<double>::infinity();

// C-ABI status -> the reference's exception types (src/cuda_solver.cu:11-19).
void check(int rc) {
    if (rc == HZL_OK)
        return;
    const std::string msg = hzl_last_error();
    if (rc == HZL_EINVAL)
        throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

const double *raw(const Field &f) {
    return reinterpret_cast<const double *>(f.data());
}
double *raw(Field &f) {
    return reinterpret_cast<double *>(f.data());
}

hzl_parameters abi_parameters(const Parameters &p, double amp, int kf) {
    return hzl_parameters{p.viscosity,   p.friction,    p.forcing_sigma, p.forcing_tau,
                          p.forcing_min, p.forcing_max, amp,             kf};
}

Field physical(std::size_t n, Field spectral) {
    Fourier2D(n).inverse(spectral);
    return spectral;
}

// Index geometry of an n x n spectral field in the reference layout y n + x:
// signed wave numbers, the conjugate partner, the 2/3-retained set (without
// the mean) and the forcing shell of a parameter set.
struct Grid {
    std::size_t n;
    int kx(std::size_t i) const { return wave_number(i % n, n); }
    int ky(std::size_t i) const { return wave_number(i / n, n); }
    std::size_t partner(std::size_t i) const {
        return ((n - i / n) % n) * n + (n - i % n) % n;
    }
    bool retained(std::size_t i) const {
        const int third = (int)n;
        return i != 0 && 3 * std::abs(kx(i)) < third && 3 * std::abs(ky(i)) < third;
    }
    bool in_shell(std::size_t i, const Parameters &p) const {
        const int k2 = kx(i) * kx(i) + ky(i) * ky(i);
        return retained(i) && k2 >= p.forcing_min * p.forcing_min &&
               k2 <= p.forcing_max * p.forcing_max;
    }
    void mask(Field &f) const {
        for (std::size_t i = 0; i < f.size(); ++i)
            if (!retained(i))
                f[i] = Complex{};
    }
};

} // namespace

// ------------------------------------------------------------ Fourier2D --
Fourier2D::Fourier2D(std::size_t n) : n_(n), bitrev_(n), twiddle_(n / 2), line_(n) {
    if (n < 8 || n > 4096 || (n & (n - 1)) != 0)
        throw std::invalid_argument("grid must be a power of two between 8 and 4096");
    int bits = 0;
    while ((std::size_t{1} << bits) < n)
        ++bits;
    for (std::size_t i = 0; i < n; ++i) {
        std::size_t r = 0;
        for (int b = 0; b < bits; ++b)
            if (i & (std::size_t{1} << b))
                r |= std::size_t{1} << (bits - 1 - b);
        bitrev_[i] = r;
    }
    for (std::size_t k = 0; k < n / 2; ++k)
        twiddle_[k] = std::polar(1.0, -2 * kPi * (double)k / (double)n);
}

// In-place radix-2 decimation-in-time transform of every row (or column).
void Fourier2D::lines(Field &data, bool columns, bool inverse) {
    const std::size_t n = n_;
    for (std::size_t l = 0; l < n; ++l) {
        for (std::size_t i = 0; i < n; ++i)
            line_[bitrev_[i]] = columns ? data[i * n + l] : data[l * n + i];
        for (std::size_t len = 2; len <= n; len <<= 1) {
            const std::size_t half = len / 2, stride = n / len;
            for (std::size_t s = 0; s < n; s += len)
                for (std::size_t k = 0; k < half; ++k) {
                    const Complex w = inverse ? std::conj(twiddle_[k * stride]) : twiddle_[k * stride];
                    const Complex a = line_[s + k], b = line_[s + k + half] * w;
                    line_[s + k] = a + b;
                    line_[s + k + half] = a - b;
                }
        }
        for (std::size_t i = 0; i < n; ++i)
            (columns ? data[i * n + l] : data[l * n + i]) = line_[i];
    }
}

void Fourier2D::forward(Field &data) {
    if (data.size() != n_ * n_)
        throw std::invalid_argument("FFT field size mismatch");
    lines(data, false, false);
    lines(data, true, false);
    const double scale = 1.0 / (double)(n_ * n_);
    for (auto &v : data)
        v *= scale;
}

void Fourier2D::inverse(Field &data) {
    if (data.size() != n_ * n_)
        throw std::invalid_argument("FFT field size mismatch");
    lines(data, false, true);
    lines(data, true, true);
}

int wave_number(std::size_t index, std::size_t n) {
    return index < n / 2 ? (int)index : (int)index - (int)n;
}

// ----------------------------------------------------------- parameters --
void Parameters::validate(std::size_t n) const {
    const hzl_parameters p = abi_parameters(*this, 0.0, 0);
    check(hzl_parameters_validate(n, &p));
}

std::uint64_t stream_seed(std::uint64_t base, std::uint64_t index) {
    return hzl_stream_seed(base, index);
}

Field initial_condition(std::size_t n, const std::string &name) {
    // Solver::initialize's named fields (src/solver.cpp:140-161): Taylor-Green
    // 2 cos x cos y, plus three modes for multimode, or zero
    const bool zero = name == "zero", multimode = name == "multimode";
    if (!zero && !multimode && name != "taylor-green")
        throw std::invalid_argument("unknown initial condition");
    Field w(n * n);
    const double h = 2 * kPi / (double)n;
    for (std::size_t i = 0; i < n * n; ++i) {
        const double x = h * (double)(i % n), y = h * (double)(i / n);
        double v = zero ? 0.0 : 2 * std::cos(x) * std::cos(y);
        if (multimode)
            v += 0.6 * std::sin(2 * x + y) + 0.4 * std::cos(x - 3 * y) +
                 0.3 * std::sin(3 * x + 2 * y);
        w[i] = v;
    }
    Fourier2D(n).forward(w);
    Grid{n}.mask(w);
    return w;
}

Field physical_to_spectral(std::size_t n, const Field &vorticity) {
    if (vorticity.size() != n * n)
        throw std::invalid_argument("physical initial field size mismatch");
    const bool real_and_finite = std::all_of(vorticity.begin(), vorticity.end(), [](Complex v) {
        return std::isfinite(v.real()) && std::isfinite(v.imag()) && v.imag() == 0;
    });
    if (!real_and_finite)
        throw std::invalid_argument("physical initial vorticity must be finite and real");
    Field w = vorticity;
    Fourier2D(n).forward(w);
    Grid{n}.mask(w);
    return w;
}

// ----------------------------------------------------------- CudaSolver --
struct CudaSolver::Implementation {
    std::size_t n;
    Parameters p;
    hzl_ensemble_t h = nullptr;
    ~Implementation() {
        if (h)
            hzl_ensemble_destroy(h);
    }
};

CudaSolver::CudaSolver(std::size_t n, Parameters parameters, const Field &initial)
    : implementation_(std::make_unique<Implementation>()) {
    parameters.validate(n);
    if (initial.size() != n * n)
        throw std::invalid_argument("CUDA initial coefficient size mismatch");
    auto &s = *implementation_;
    s.n = n;
    s.p = parameters;
    const hzl_parameters ap = abi_parameters(parameters, 0.0, 4);
    check(hzl_ensemble_create(&s.h, n, 1, HZL_FP64, &ap, 0));
    check(hzl_ensemble_set_state(s.h, 0, raw(initial)));
}

CudaSolver::~CudaSolver() = default;

void CudaSolver::step(double dt) {
    // the reference GPU solver has no forcing generator (src/cuda_solver.cu:172-183)
    if (implementation_->p.forcing_sigma != 0)
        throw std::invalid_argument("stochastic GPU step requires midpoint forcing coefficients");
    // sigma = 0 and no Kolmogorov term: the engine step adds no forcing
    check(hzl_ensemble_step(implementation_->h, dt, 1));
}

void CudaSolver::step(double dt, const Field &midpoint_forcing) {
    if (midpoint_forcing.size() != implementation_->n * implementation_->n)
        throw std::invalid_argument("GPU forcing size mismatch");
    check(hzl_ensemble_step_forced(implementation_->h, dt, raw(midpoint_forcing), 1));
}

void CudaSolver::synchronize() {
    check(hzl_ensemble_synchronize(implementation_->h));
}

Field CudaSolver::coefficients() const {
    const std::size_t n = implementation_->n;
    check(hzl_ensemble_synchronize(implementation_->h));
    Field out(n * n);
    check(hzl_ensemble_get_state(implementation_->h, 0, raw(out)));
    return out;
}

// ------------------------------------------------------------- Ensemble --
Ensemble::Ensemble(std::size_t n, std::size_t members, Parameters parameters,
                   Precision precision, std::uint64_t seed, double kolmogorov_amplitude,
                   int kolmogorov_wavenumber)
    : n_(n), members_(members) {
    parameters.validate(n);
    const hzl_parameters ap = abi_parameters(parameters, kolmogorov_amplitude,
                                             kolmogorov_wavenumber);
    check(hzl_ensemble_create(&handle_, n, members,
                              precision == Precision::fp32 ? HZL_FP32 : HZL_FP64, &ap, seed));
}

Ensemble::~Ensemble() {
    if (handle_)
        hzl_ensemble_destroy(handle_);
}

void Ensemble::set_member_parameters(std::size_t member, double viscosity, double friction,
                                     double forcing_sigma, double kolmogorov_amplitude) {
    check(hzl_ensemble_set_member_parameters(handle_, member, viscosity, friction,
                                             forcing_sigma, kolmogorov_amplitude));
}

void Ensemble::set_state(std::size_t member, const Field &coefficients) {
    if (coefficients.size() != n_ * n_)
        throw std::invalid_argument("state size mismatch");
    check(hzl_ensemble_set_state(handle_, member, raw(coefficients)));
}

void Ensemble::set_all(const Field &coefficients) {
    if (coefficients.size() != n_ * n_)
        throw std::invalid_argument("state size mismatch");
    for (std::size_t m = 0; m < members_; ++m)
        check(hzl_ensemble_set_state(handle_, m, raw(coefficients)));
}

Field Ensemble::state(std::size_t member) const {
    Field out(n_ * n_);
    check(hzl_ensemble_get_state(handle_, member, raw(out)));
    return out;
}

void Ensemble::set_forcing_process(ForcingProcess process) {
    check(hzl_ensemble_set_forcing_process(
        handle_, process == ForcingProcess::white ? HZL_FORCING_WHITE : HZL_FORCING_OU));
}

Field Ensemble::forcing(std::size_t member) const {
    Field out(n_ * n_);
    check(hzl_ensemble_get_forcing(handle_, member, raw(out)));
    return out;
}

void Ensemble::reseed_future(std::size_t member, std::uint64_t seed) {
    check(hzl_ensemble_reseed_future(handle_, member, seed));
}

void Ensemble::set_forcing_state(std::size_t member, const std::vector<std::uint64_t> &engine,
                                 const Field &forcing) {
    if (engine.size() != 313 || forcing.size() != n_ * n_)
        throw std::invalid_argument("forcing state size mismatch");
    check(hzl_ensemble_set_forcing_state(handle_, member, engine.data(), raw(forcing)));
}

std::pair<std::vector<std::uint64_t>, Field> Ensemble::forcing_state(std::size_t member) const {
    std::vector<std::uint64_t> engine(313);
    Field forcing(n_ * n_);
    check(hzl_ensemble_get_forcing_state(handle_, member, engine.data(), raw(forcing)));
    return {std::move(engine), std::move(forcing)};
}

void Ensemble::step(double dt, std::size_t steps) {
    check(hzl_ensemble_step(handle_, dt, steps));
}

void Ensemble::step(double dt, std::size_t steps, TimeIntegrator integrator) {
    check(hzl_ensemble_step_integrator(handle_, dt, steps,
                                       integrator == TimeIntegrator::rk4 ? HZL_RK4 : HZL_SSPRK3));
}

namespace {
std::vector<double> pack_fields(const std::vector<Field> &fields, std::size_t count,
                                std::size_t n) {
    if (fields.size() != count)
        throw std::invalid_argument("one field per member expected");
    std::vector<double> flat;
    flat.reserve(2 * n * n * count);
    for (const auto &f : fields) {
        if (f.size() != n * n)
            throw std::invalid_argument("field size mismatch");
        for (const auto &c : f) {
            flat.push_back(c.real());
            flat.push_back(c.imag());
        }
    }
    return flat;
}

std::vector<Field> unpack_fields(const std::vector<double> &flat, std::size_t count,
                                 std::size_t n) {
    std::vector<Field> out(count, Field(n * n));
    for (std::size_t m = 0; m < count; ++m)
        for (std::size_t i = 0; i < n * n; ++i)
            out[m][i] = Complex(flat[2 * (m * n * n + i)], flat[2 * (m * n * n + i) + 1]);
    return out;
}
} // namespace

std::vector<Field> Ensemble::drift() {
    std::vector<double> out(2 * n_ * n_ * members_);
    check(hzl_ensemble_drift(handle_, out.data()));
    return unpack_fields(out, members_, n_);
}

std::vector<Field> Ensemble::linearized_drift(const std::vector<Field> &directions) {
    const auto in = pack_fields(directions, members_, n_);
    std::vector<double> out(in.size());
    check(hzl_ensemble_linearized_drift(handle_, in.data(), out.data()));
    return unpack_fields(out, members_, n_);
}

std::vector<Field> Ensemble::adjoint_drift(const std::vector<Field> &cotangents) {
    const auto in = pack_fields(cotangents, members_, n_);
    std::vector<double> out(in.size());
    check(hzl_ensemble_adjoint_drift(handle_, in.data(), out.data()));
    return unpack_fields(out, members_, n_);
}

void Ensemble::step(double dt, const Field &shared_midpoint_forcing) {
    if (shared_midpoint_forcing.size() != n_ * n_)
        throw std::invalid_argument("GPU forcing size mismatch");
    check(hzl_ensemble_step_forced(handle_, dt, raw(shared_midpoint_forcing), 1));
}

void Ensemble::synchronize() {
    check(hzl_ensemble_synchronize(handle_));
}

std::vector<std::vector<double>> Ensemble::energy_spectra() {
    std::vector<double> bins(members_ * n_);
    check(hzl_ensemble_energy_spectrum(handle_, bins.data()));
    std::vector<std::vector<double>> out(members_);
    for (std::size_t m = 0; m < members_; ++m)
        out[m].assign(bins.begin() + (long)(m * n_), bins.begin() + (long)((m + 1) * n_));
    return out;
}

std::vector<Diagnostics> Ensemble::diagnostics() {
    std::vector<double> d(6 * members_);
    check(hzl_ensemble_diagnostics(handle_, d.data()));
    std::vector<Diagnostics> out(members_);
    for (std::size_t m = 0; m < members_; ++m)
        out[m] = Diagnostics{0.0, d[6 * m], d[6 * m + 1], d[6 * m + 2], d[6 * m + 3],
                             d[6 * m + 4]};
    return out;
}

// -------------------------------------------------- likelihood_adjoint --
std::vector<AdjointEvaluation> likelihood_adjoint(const Dataset &data,
                                                  const std::vector<Vector> &log_parameters,
                                                  std::size_t checkpoint_stride) {
    data.validate();
    if (data.stochastic || checkpoint_stride == 0 || log_parameters.empty())
        throw std::invalid_argument(
            "adjoint requires deterministic data, two parameters, and positive checkpoint stride");
    for (const auto &q : log_parameters)
        if (q.size() != 2)
            throw std::invalid_argument(
                "adjoint requires deterministic data, two parameters, and positive checkpoint "
                "stride");
    const std::size_t B = log_parameters.size(), nobs = data.observations.size();
    Parameters p;
    p.forcing_sigma = 0;
    p.forcing_min = data.forcing_min;
    p.forcing_max = data.forcing_max;
    Ensemble e(data.n, B, p, Precision::fp64, 0);
    std::vector<std::size_t> steps(nobs);
    std::vector<int> x(nobs), y(nobs);
    std::vector<double> value(nobs), sd(nobs), q(2 * B), obj(B), grad(2 * B);
    for (std::size_t o = 0; o < nobs; ++o) {
        const auto &ob = data.observations[o];
        steps[o] = ob.step;
        x[o] = (int)ob.x;
        y[o] = (int)ob.y;
        value[o] = ob.value;
        sd[o] = ob.standard_deviation;
    }
    for (std::size_t m = 0; m < B; ++m) {
        q[2 * m] = log_parameters[m][0];
        q[2 * m + 1] = log_parameters[m][1];
    }
    const Field init = initial_condition(data.n, "multimode");
    std::size_t counts[3] = {0, 0, 0};
    check(hzl_likelihood_adjoint(e.handle(), raw(init), data.dt, nobs, steps.data(), x.data(),
                                 y.data(), value.data(), sd.data(), q.data(), checkpoint_stride,
                                 obj.data(), grad.data(), counts));
    std::vector<AdjointEvaluation> out(B);
    for (std::size_t m = 0; m < B; ++m) {
        out[m].objective.value = obj[m];
        out[m].objective.gradient = {grad[2 * m], grad[2 * m + 1]};
        if (std::isfinite(obj[m])) {
            out[m].checkpoint_fields = counts[0];
            out[m].peak_block_fields = counts[1];
            out[m].replayed_steps = counts[2];
        }
    }
    return out;
}

ValueGradient likelihood_gradient(const Dataset &data, const Vector &log_parameters) {
    data.validate();
    if (data.stochastic || log_parameters.size() != 2)
        throw std::invalid_argument(
            "tangent likelihood requires deterministic data and two parameters");
    const std::size_t nobs = data.observations.size();
    Parameters p;
    p.forcing_sigma = 0;
    p.forcing_min = data.forcing_min;
    p.forcing_max = data.forcing_max;
    Ensemble e(data.n, 1, p, Precision::fp64, 0);
    std::vector<std::size_t> steps(nobs);
    std::vector<int> x(nobs), y(nobs);
    std::vector<double> value(nobs), sd(nobs);
    for (std::size_t o = 0; o < nobs; ++o) {
        const auto &ob = data.observations[o];
        steps[o] = ob.step;
        x[o] = (int)ob.x;
        y[o] = (int)ob.y;
        value[o] = ob.value;
        sd[o] = ob.standard_deviation;
    }
    const Field init = initial_condition(data.n, "multimode");
    double obj = 0, grad[2] = {0, 0};
    check(hzl_likelihood_gradient(e.handle(), raw(init), data.dt, nobs, steps.data(), x.data(),
                                  y.data(), value.data(), sd.data(), log_parameters.data(), &obj,
                                  grad));
    return ValueGradient{obj, {grad[0], grad[1]}};
}

AdjointEvaluation likelihood_adjoint(const Dataset &data, const Vector &log_parameters,
                                     std::size_t checkpoint_stride) {
    return likelihood_adjoint(data, std::vector<Vector>{log_parameters}, checkpoint_stride)[0];
}

// -------------------------------------------------------------- Dataset --
// Dataset::validate (src/inference.cpp:197-213): the grid and the forcing
// shell, then the dataset's own rules.
void Dataset::validate() const {
    (void)Fourier2D(n);
    Parameters shell;
    shell.forcing_min = forcing_min;
    shell.forcing_max = forcing_max;
    shell.validate(n);
    if (n < 16 || !(std::isfinite(dt) && dt > 0) || observations.empty())
        throw std::invalid_argument("invalid observation dataset");
    std::size_t last = 0;
    const auto admissible = [&](const Observation &o) {
        const bool ok = o.step >= last && o.x < n && o.y < n && std::isfinite(o.value) &&
                        std::isfinite(o.standard_deviation) && o.standard_deviation > 0;
        last = o.step;
        return ok;
    };
    if (!std::all_of(observations.begin(), observations.end(), admissible))
        throw std::invalid_argument("observations must be time ordered, finite, in bounds, and "
                                    "have positive uncertainty");
}

namespace {
// HZL_OBSERVATIONS_V2 (src/inference.cpp:214-255): a header line
// "magic,n,dt,stochastic,forcing_min,forcing_max" (V1: "magic,n,dt"), the
// column line, then one observation per line; numbers with 17 digits.
constexpr const char *kObservationsV1 = "HZL_OBSERVATIONS_V1";
constexpr const char *kObservationsV2 = "HZL_OBSERVATIONS_V2";
constexpr const char *kObservationColumns = "step,x,y,vorticity,standard_deviation";
} // namespace

void Dataset::save(const std::filesystem::path &path) const {
    validate();
    text::LineWriter out(path, ',', "observation write failed");
    out.line(kObservationsV2, n, dt, stochastic, forcing_min, forcing_max);
    out.line(kObservationColumns);
    for (const Observation &o : observations)
        out.line(o.step, o.x, o.y, o.value, o.standard_deviation);
    out.close();
}

Dataset Dataset::load(const std::filesystem::path &path) {
    const std::string text = text::Scanner::slurp(path);
    const auto lines = text::lines_of(text);
    if (lines.empty())
        throw std::runtime_error("cannot read observation dataset");
    Dataset data;
    text::Scanner header(std::string(lines[0]), true);
    std::string magic;
    if (!header.read(magic) || !header.read(data.n) || !header.read(data.dt) ||
        (magic != kObservationsV1 && magic != kObservationsV2))
        throw std::runtime_error("invalid dataset header");
    if (magic == kObservationsV2 && !(header.read(data.stochastic) &&
                                      header.read(data.forcing_min) &&
                                      header.read(data.forcing_max)))
        throw std::runtime_error("invalid stochastic dataset metadata");
    if (lines.size() < 2 || lines[1] != kObservationColumns)
        throw std::runtime_error("invalid dataset columns");
    for (std::size_t l = 2; l < lines.size(); ++l) {
        if (lines[l].empty())
            continue;
        text::Scanner row(std::string(lines[l]), true);
        Observation o;
        if (!(row.read(o.step) && row.read(o.x) && row.read(o.y) && row.read(o.value) &&
              row.read(o.standard_deviation)) ||
            !row.at_end())
            throw std::runtime_error("invalid observation row");
        data.observations.push_back(o);
    }
    data.validate();
    return data;
}

// Observations of a truth run (synthetic_observations, src/inference.cpp:256-288):
// one GPU member started from "multimode" with the reference's OU forcing on
// stream stream_seed(seed, 0); every `every` steps its physical field at the
// eight sensors x_j = (j n / 8 + 1) mod n, y_j = (j^2 + 3 j + 2) mod n, plus
// noise from NormalStream(stream_seed(seed, 1)) drawn in sensor order.
Dataset synthetic_observations(std::size_t n, double dt, std::size_t steps, std::size_t every,
                               Parameters truth, double noise, std::uint64_t seed,
                               std::vector<Field> *truth_fields) {
    if (every == 0 || steps < every || !std::isfinite(noise) || noise <= 0)
        throw std::invalid_argument(
            "observation experiment requires steps >= every and positive noise");
    (void)Fourier2D(n);
    truth.validate(n);
    constexpr std::size_t kSensors = 8;
    const std::size_t times = steps / every;
    std::vector<double> draws(kSensors * times);
    check(hzl_normal_draws(stream_seed(seed, 1), draws.size(), draws.data()));

    // member 0 of an ensemble created with `seed` runs the OU forcing of
    // Solver(n, truth, stream_seed(seed, 0))
    Ensemble run(n, 1, truth, Precision::fp64, seed);
    run.set_all(initial_condition(n, "multimode"));
    run.set_forcing_process(ForcingProcess::ornstein_uhlenbeck);

    Dataset data;
    data.n = n;
    data.dt = dt;
    data.stochastic = truth.forcing_sigma > 0;
    data.forcing_min = truth.forcing_min;
    data.forcing_max = truth.forcing_max;
    if (truth_fields)
        truth_fields->clear();
    for (std::size_t t = 0; t < times; ++t) {
        run.step(dt, every);
        run.synchronize();
        const Field field = physical(n, run.state(0));
        for (std::size_t j = 0; j < kSensors; ++j) {
            const std::size_t x = (j * n / kSensors + 1) % n, y = (j * j + 3 * j + 2) % n;
            data.observations.push_back({(t + 1) * every, x, y,
                                         field[y * n + x].real() + noise * draws[kSensors * t + j],
                                         noise});
        }
        if (truth_fields)
            truth_fields->push_back(field);
    }
    data.validate();
    return data;
}

// ----------------------------------------------------------- Checkpoint --
// HZL_CHECKPOINT_V1 (src/solver.cpp:434-491), whitespace separated:
//   magic / n time / nu alpha sigma tau kmin kmax / 312 engine words and the
//   position / has_spare spare / per index y n + x: Re w, Im w, Re F, Im F.
namespace {
constexpr const char *kCheckpointMagic = "HZL_CHECKPOINT_V1";
constexpr std::size_t kEngineWords = 312;
} // namespace

void Checkpoint::save(const std::filesystem::path &path) const {
    if (vorticity.size() != n * n || forcing.size() != n * n ||
        random_engine.size() != kEngineWords + 1)
        throw std::invalid_argument("checkpoint fields do not match the grid");
    text::LineWriter out(path, ' ', "checkpoint write failed", "cannot open checkpoint for writing");
    const Parameters &p = parameters;
    out.line(kCheckpointMagic);
    out.line(n, time);
    out.line(p.viscosity, p.friction, p.forcing_sigma, p.forcing_tau, p.forcing_min,
             p.forcing_max);
    for (std::uint64_t word : random_engine)
        out.add(word);
    out.end();
    out.line(random_has_spare, random_spare);
    for (std::size_t i = 0; i < n * n; ++i)
        out.line(vorticity[i].real(), vorticity[i].imag(), forcing[i].real(), forcing[i].imag());
    out.close();
}

Checkpoint Checkpoint::load(const std::filesystem::path &path) {
    text::Scanner in(text::Scanner::slurp(path), false);
    Checkpoint c;
    Parameters &p = c.parameters;
    std::string magic;
    const bool header = in.read(magic) && in.read(c.n) && in.read(c.time) &&
                        in.read(p.viscosity) && in.read(p.friction) &&
                        in.read(p.forcing_sigma) && in.read(p.forcing_tau) &&
                        in.read(p.forcing_min) && in.read(p.forcing_max);
    if (!header || magic != kCheckpointMagic || !std::isfinite(c.time) || c.time < 0)
        throw std::runtime_error("invalid checkpoint header");
    // the reference rebuilds Solver(n, p, 0): grid, parameters, non-empty shell
    (void)Fourier2D(c.n);
    p.validate(c.n);
    const Grid grid{c.n};
    const std::size_t nn = c.n * c.n;
    std::vector<bool> shell(nn);
    for (std::size_t i = 0; i < nn; ++i)
        shell[i] = grid.in_shell(i, p);
    if (std::none_of(shell.begin(), shell.end(), [](bool b) { return b; }))
        throw std::invalid_argument("empty forcing shell");

    c.random_engine.resize(kEngineWords + 1);
    bool stream = std::all_of(c.random_engine.begin(), c.random_engine.end(),
                              [&](std::uint64_t &w) { return in.read(w); });
    stream = stream && c.random_engine[kEngineWords] <= kEngineWords &&
             in.read(c.random_has_spare) && in.read(c.random_spare) &&
             std::isfinite(c.random_spare);
    if (!stream)
        throw std::runtime_error("invalid checkpoint random stream");

    c.vorticity.assign(nn, Complex{});
    c.forcing.assign(nn, Complex{});
    for (std::size_t i = 0; i < nn; ++i) {
        double v[4];
        for (double &x : v)
            if (!in.read(x) || !std::isfinite(x))
                throw std::runtime_error("invalid checkpoint field");
        c.vorticity[i] = {v[0], v[1]};
        c.forcing[i] = {v[2], v[3]};
    }
    // masked vorticity, forcing only on the shell, conjugate symmetry
    // (vorticity to 1e-10 relative, forcing exactly)
    for (std::size_t i = 0; i < nn; ++i) {
        const std::size_t j = grid.partner(i);
        const Complex w = c.vorticity[i], F = c.forcing[i];
        const bool fits = (grid.retained(i) || w == Complex{}) && (shell[i] || F == Complex{}) &&
                          std::abs(w - std::conj(c.vorticity[j])) <= 1e-10 * (1 + std::abs(w)) &&
                          F == std::conj(c.forcing[j]);
        if (!fits)
            throw std::runtime_error("checkpoint violates spectral constraints");
    }
    if (!in.at_end())
        throw std::runtime_error("unexpected checkpoint data");
    return c;
}

// -------------------------------------------------------------- filters --
namespace {

struct Group {
    std::size_t step;
    std::vector<int> x, y;
    std::vector<double> value, sd;
};

std::vector<Group> groups_of(const Dataset &data) {
    std::vector<Group> out;
    for (const auto &o : data.observations) {
        if (out.empty() || out.back().step != o.step)
            out.push_back(Group{o.step, {}, {}, {}, {}});
        out.back().x.push_back((int)o.x);
        out.back().y.push_back((int)o.y);
        out.back().value.push_back(o.value);
        out.back().sd.push_back(o.standard_deviation);
    }
    return out;
}

Parameters with_data_shell(Parameters p, const Dataset &data) {
    p.forcing_min = data.forcing_min;
    p.forcing_max = data.forcing_max;
    return p;
}

} // namespace

FilterResult bootstrap_filter(Ensemble &particles, const Dataset &data, std::uint64_t seed,
                              bool retain_states) {
    data.validate();
    if (particles.size() != data.n)
        throw std::invalid_argument("particle grid does not match the dataset");
    const auto groups = groups_of(data);
    check(hzl_pf_begin(particles.handle(), seed, groups.size() + 1));
    FilterResult result;
    std::size_t current = 0;
    const std::size_t n = data.n;
    for (std::size_t g = 0; g < groups.size(); ++g) {
        const Group &gr = groups[g];
        double rec[4];
        Field mean(retain_states ? n * n : 0);
        check(hzl_pf_assimilate(particles.handle(), data.dt, gr.step - current, gr.x.size(),
                                gr.x.data(), gr.y.data(), gr.value.data(), gr.sd.data(),
                                g + 1 == groups.size(), rec, retain_states ? raw(mean) : nullptr));
        // every weight -inf: the reference returns before recording the group
        // (src/filter.cpp:76-80)
        if (!std::isfinite(rec[0]))
            break;
        FilterRecord r{gr.step, rec[0], rec[1], rec[2], {}};
        if (retain_states)
            r.mean_vorticity = physical(n, std::move(mean));
        result.history.push_back(std::move(r));
        current = gr.step;
    }
    check(hzl_pf_result(particles.handle(), &result.log_likelihood, &result.resampling_events));
    return result;
}

FilterResult bootstrap_filter(const Dataset &data, Parameters p, std::size_t particles,
                              std::uint64_t seed, bool retain_states) {
    data.validate();
    p = with_data_shell(p, data);
    p.validate(data.n);
    if (particles < 2)
        throw std::invalid_argument("particle filter needs at least two particles");
    // particle i is Solver(n, p, stream_seed(seed, i)) from multimode with its
    // own OU forcing (src/filter.cpp:46-51)
    Ensemble e(data.n, particles, p, Precision::fp64, seed);
    e.set_all(initial_condition(data.n, "multimode"));
    e.set_forcing_process(ForcingProcess::ornstein_uhlenbeck);
    return bootstrap_filter(e, data, seed, retain_states);
}

// ------------------------------------------------- particle likelihoods --
// The particle-marginal family over q = (log nu, log alpha, log sigma,
// log tau) (src/filter.cpp:127-203) runs on the engine's batched drivers: the
// prior and the sampler's streams live behind the C-ABI, every likelihood is a
// GPU bootstrap filter with the OU forcing (FP64), and replicates or chain
// proposals are segments of one ensemble.
namespace {

// Observation rows as the batched drivers take them.
struct Rows {
    std::vector<int> step, x, y;
    std::vector<double> value, sd;
    explicit Rows(const Dataset &data) {
        for (const Observation &o : data.observations) {
            step.push_back((int)o.step);
            x.push_back((int)o.x);
            y.push_back((int)o.y);
            value.push_back(o.value);
            sd.push_back(o.standard_deviation);
        }
    }
};

// The data's grid shell on default parameters: the batched drivers set every
// member's nu, alpha, sigma and tau from its filter's q.
Parameters shell_of(const Dataset &data) {
    Parameters p;
    p.forcing_min = data.forcing_min;
    p.forcing_max = data.forcing_max;
    return p;
}

hzl_pmcmc_options stochastic_options() {
    hzl_pmcmc_options o{};
    o.prior = HZL_PRIOR_STOCHASTIC;
    return o;
}

// particle_log_likelihood at one q for every seed (one filter per seed).
Vector likelihoods(const Dataset &data, const Vector &q, std::size_t particles,
                   const std::vector<std::uint64_t> &seeds) {
    double prior = 0;
    check(hzl_stochastic_log_prior(q.data(), q.size(), &prior));
    if (!std::isfinite(prior))
        return Vector(seeds.size(), kNegativeInfinity);
    data.validate();
    if (particles < 2)
        throw std::invalid_argument("particle filter needs at least two particles");
    Ensemble e(data.n, seeds.size() * particles, shell_of(data), Precision::fp64, 0);
    e.set_forcing_process(ForcingProcess::ornstein_uhlenbeck);
    const Rows rows(data);
    const Field start = initial_condition(data.n, "multimode");
    const hzl_pmcmc_options o = stochastic_options();
    Vector points, ll(seeds.size());
    for (std::size_t f = 0; f < seeds.size(); ++f)
        points.insert(points.end(), q.begin(), q.end());
    check(hzl_pf_log_likelihoods(e.handle(), &o, data.dt, rows.step.size(), rows.step.data(),
                                 rows.x.data(), rows.y.data(), rows.value.data(), rows.sd.data(),
                                 raw(start), seeds.size(), points.data(), seeds.data(),
                                 ll.data()));
    return ll;
}

} // namespace

double stochastic_log_prior(const Vector &q) {
    double out = 0;
    check(hzl_stochastic_log_prior(q.data(), q.size(), &out));
    return out;
}

double particle_log_likelihood(const Dataset &data, const Vector &q, std::size_t particles,
                               std::uint64_t seed) {
    return likelihoods(data, q, particles, {seed})[0];
}

// The replicates run as one batch of filters; their log-mean-exp is formed
// around the largest estimate.
double averaged_particle_log_likelihood(const Dataset &data, const Vector &q,
                                        std::size_t particles, std::size_t replicates,
                                        std::uint64_t seed) {
    if (replicates == 0)
        throw std::invalid_argument("averaged particle likelihood needs at least one replicate");
    std::vector<std::uint64_t> seeds(replicates);
    for (std::size_t i = 0; i < replicates; ++i)
        seeds[i] = stream_seed(seed, i);
    const Vector ll = likelihoods(data, q, particles, seeds);
    const double top = *std::max_element(ll.begin(), ll.end());
    if (!std::isfinite(top))
        return kNegativeInfinity;
    double mean = 0;
    for (double v : ll)
        mean += std::exp(v - top);
    return top + std::log(mean / (double)replicates);
}

// One chain of the batched sampler with the reference's own streams
// (chain seed = seed) and prior.
std::vector<ParticleSample> particle_metropolis(const Dataset &data, Vector initial,
                                                double scale, std::size_t count,
                                                std::size_t particles, std::uint64_t seed) {
    if (!data.stochastic)
        throw std::invalid_argument("particle sampler expects a stochastic observation dataset");
    if (count == 0 || !std::isfinite(scale) || scale <= 0)
        throw std::invalid_argument("invalid particle sampler configuration");
    (void)stochastic_log_prior(initial); // four components
    data.validate();
    if (particles < 2)
        throw std::invalid_argument("particle filter needs at least two particles");
    Ensemble e(data.n, particles, shell_of(data), Precision::fp64, 0);
    e.set_forcing_process(ForcingProcess::ornstein_uhlenbeck);
    const Rows rows(data);
    const Field start = initial_condition(data.n, "multimode");
    hzl_pmcmc_options o = stochastic_options();
    o.chains = 1;
    o.samples = count;
    o.proposal_scale = scale;
    o.chain_seeds = &seed;
    constexpr std::size_t D = 4;
    Vector q(D * count), ll(count), density(count);
    std::vector<int> accepted(count);
    std::vector<std::uint64_t> seeds(count);
    check(hzl_pmcmc_run(e.handle(), &o, data.dt, rows.step.size(), rows.step.data(),
                        rows.x.data(), rows.y.data(), rows.value.data(), rows.sd.data(),
                        raw(start), initial.data(), q.data(), ll.data(), density.data(),
                        accepted.data(), seeds.data()));
    std::vector<ParticleSample> chain(count);
    for (std::size_t s = 0; s < count; ++s)
        chain[s] = ParticleSample{Vector(q.begin() + (long)(D * s), q.begin() + (long)(D * s + D)),
                                  ll[s], density[s], seeds[s], accepted[s] != 0};
    return chain;
}

std::vector<EnsembleFilterRecord> ensemble_kalman_filter(Ensemble &members, const Dataset &data,
                                                         std::uint64_t seed, EnKFOptions options,
                                                         bool batch) {
    data.validate();
    if (members.members() < 2)
        throw std::invalid_argument("EnKF requires at least two ensemble members");
    if (!std::isfinite(options.inflation) || options.inflation < 1 ||
        !std::isfinite(options.localization_radius) || options.localization_radius < 0)
        throw std::invalid_argument(
            "EnKF inflation must be >= 1 and localization radius nonnegative");
    if (options.forecast_every == 0)
        throw std::invalid_argument("EnKF forecast output interval must be positive");
    const std::size_t n = data.n;
    check(hzl_enkf_begin(members.handle(), seed, batch ? HZL_ENKF_BATCH : HZL_ENKF_SERIAL));
    std::vector<EnsembleFilterRecord> history;
    std::size_t current = 0;
    for (const Group &gr : groups_of(data)) {
        double rec[2];
        Field mean(n * n);
        check(hzl_enkf_assimilate(members.handle(), data.dt, gr.step - current, gr.x.size(),
                                  gr.x.data(), gr.y.data(), gr.value.data(), gr.sd.data(),
                                  options.inflation, options.localization_radius, rec,
                                  raw(mean)));
        history.push_back({gr.step, rec[0], rec[1], physical(n, std::move(mean)), false});
        current = gr.step;
    }
    for (std::size_t horizon = 1; horizon <= options.forecast_steps; ++horizon) {
        members.step(data.dt, 1);
        if (horizon % options.forecast_every != 0 && horizon != options.forecast_steps)
            continue;
        double spread = 0;
        Field mean(n * n);
        check(hzl_ensemble_mean_spread(members.handle(), &spread, raw(mean)));
        history.push_back({current + horizon, std::numeric_limits<double>::quiet_NaN(), spread,
                           physical(n, std::move(mean)), true});
    }
    return history;
}

std::vector<EnsembleFilterRecord> ensemble_kalman_filter(const Dataset &data, Parameters p,
                                                         std::size_t count, std::uint64_t seed,
                                                         EnKFOptions options) {
    data.validate();
    if (count < 2)
        throw std::invalid_argument("EnKF requires at least two ensemble members");
    p = with_data_shell(p, data);
    p.validate(data.n);
    // member i is Solver(n, p, stream_seed(seed, i)) from multimode; its OU
    // forcing is part of the analysed state (src/filter.cpp:217-223,237-321)
    Ensemble e(data.n, count, p, Precision::fp64, seed);
    e.set_all(initial_condition(data.n, "multimode"));
    e.set_forcing_process(ForcingProcess::ornstein_uhlenbeck);
    return ensemble_kalman_filter(e, data, seed, options, false);
}

std::vector<Diagnostics> ensemble(std::size_t n, Parameters p, double dt, std::size_t steps,
                                  std::size_t count, std::size_t workers, std::uint64_t seed) {
    if (count == 0 || workers == 0)
        throw std::invalid_argument("ensemble count and workers must be positive");
    // member i is Solver(n, p, stream_seed(seed, i)) from zero vorticity with
    // its own OU forcing (src/inference.cpp:563-575)
    Ensemble e(n, count, p, Precision::fp64, seed);
    e.set_forcing_process(ForcingProcess::ornstein_uhlenbeck);
    e.step(dt, steps);
    e.synchronize();
    auto d = e.diagnostics();
    double time = 0; // Solver::time(): t += dt per step (src/solver.cpp:253)
    for (std::size_t s = 0; s < steps; ++s)
        time += dt;
    for (auto &x : d)
        x.time = time;
    return d;
}

} // namespace hzl
