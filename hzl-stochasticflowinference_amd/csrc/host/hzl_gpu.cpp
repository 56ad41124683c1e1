// hzl_gpu.cpp -- C++ drop-in layer over the C-ABI (see hzl_gpu.hpp).
#include "hzl_gpu.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <limits>
#include <numbers>
#include <random>
#include <sstream>
#include <stdexcept>

namespace hzl {

namespace {

constexpr double kPi = 3.14159265358979323846;

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
    if (!std::isfinite(viscosity) || viscosity < 0 || !std::isfinite(friction) || friction < 0 ||
        !std::isfinite(forcing_sigma) || forcing_sigma < 0 || !std::isfinite(forcing_tau) ||
        forcing_tau <= 0)
        throw std::invalid_argument(
            "nu, alpha, sigma must be finite and nonnegative; tau must be positive");
    if (forcing_min < 1 || forcing_max < forcing_min ||
        3 * static_cast<long long>(forcing_max) >= static_cast<long long>(n))
        throw std::invalid_argument("forcing shell must lie strictly inside the dealiased grid");
}

std::uint64_t stream_seed(std::uint64_t base, std::uint64_t index) {
    // SplitMix64 finaliser of base + golden-ratio increments (src/solver.cpp:61-66)
    std::uint64_t x = base + 0x9e3779b97f4a7c15ULL * (index + 1);
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

namespace {
void mask(std::size_t n, Field &f) {
    for (std::size_t y = 0; y < n; ++y)
        for (std::size_t x = 0; x < n; ++x) {
            const int a = wave_number(x, n), b = wave_number(y, n);
            const bool active = 3 * std::abs(a) < (int)n && 3 * std::abs(b) < (int)n &&
                                (x != 0 || y != 0);
            if (!active)
                f[y * n + x] = 0;
        }
}
} // namespace

Field initial_condition(std::size_t n, const std::string &name) {
    if (name != "taylor-green" && name != "multimode" && name != "zero")
        throw std::invalid_argument("unknown initial condition");
    Field w(n * n);
    for (std::size_t y = 0; y < n; ++y)
        for (std::size_t x = 0; x < n; ++x) {
            const double a = 2 * kPi * (double)x / (double)n, b = 2 * kPi * (double)y / (double)n;
            double v = name == "zero" ? 0.0 : 2 * std::cos(a) * std::cos(b);
            if (name == "multimode")
                v += 0.6 * std::sin(2 * a + b) + 0.4 * std::cos(a - 3 * b) +
                     0.3 * std::sin(3 * a + 2 * b);
            w[y * n + x] = v;
        }
    Fourier2D(n).forward(w);
    mask(n, w);
    return w;
}

Field physical_to_spectral(std::size_t n, const Field &vorticity) {
    if (vorticity.size() != n * n)
        throw std::invalid_argument("physical initial field size mismatch");
    for (const auto &v : vorticity)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()) || v.imag() != 0)
            throw std::invalid_argument("physical initial vorticity must be finite and real");
    Field w = vorticity;
    Fourier2D(n).forward(w);
    mask(n, w);
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
void Dataset::validate() const {
    Parameters p;
    p.forcing_min = forcing_min;
    p.forcing_max = forcing_max;
    (void)Fourier2D(n);
    p.validate(n);
    if (n < 16 || !std::isfinite(dt) || dt <= 0 || observations.empty())
        throw std::invalid_argument("invalid observation dataset");
    std::size_t previous = 0;
    for (const auto &o : observations) {
        if (o.step < previous || o.x >= n || o.y >= n || !std::isfinite(o.value) ||
            !std::isfinite(o.standard_deviation) || o.standard_deviation <= 0)
            throw std::invalid_argument("observations must be time ordered, finite, in bounds, and "
                                        "have positive uncertainty");
        previous = o.step;
    }
}

void Dataset::save(const std::filesystem::path &path) const {
    validate();
    std::ofstream out(path);
    out << std::setprecision(17);
    out << "HZL_OBSERVATIONS_V2," << n << ',' << dt << ',' << stochastic << ',' << forcing_min
        << ',' << forcing_max << "\nstep,x,y,vorticity,standard_deviation\n";
    for (const auto &o : observations)
        out << o.step << ',' << o.x << ',' << o.y << ',' << o.value << ',' << o.standard_deviation
            << '\n';
    if (!out)
        throw std::runtime_error("observation write failed");
}

namespace {
// Retained strictly-inside-2/3 mask without the mean (src/solver.cpp:126-135).
bool retained(std::size_t n, std::size_t i) {
    const int a = wave_number(i % n, n), b = wave_number(i / n, n);
    return 3 * std::abs(a) < (int)n && 3 * std::abs(b) < (int)n && i != 0;
}
// Indices of the forcing shell pairs, both members (src/solver.cpp:67-83).
std::vector<bool> shell_indices(std::size_t n, const Parameters &p) {
    std::vector<bool> forced(n * n, false);
    bool any = false;
    for (std::size_t i = 0; i < n * n; ++i) {
        const int a = wave_number(i % n, n), b = wave_number(i / n, n);
        const int k2 = a * a + b * b;
        if (retained(n, i) && k2 >= p.forcing_min * p.forcing_min &&
            k2 <= p.forcing_max * p.forcing_max) {
            forced[i] = true;
            any = true;
        }
    }
    if (!any)
        throw std::invalid_argument("empty forcing shell");
    return forced;
}
constexpr std::size_t kEngineWords = 312;
} // namespace

void Checkpoint::save(const std::filesystem::path &path) const {
    if (vorticity.size() != n * n || forcing.size() != n * n ||
        random_engine.size() != kEngineWords + 1)
        throw std::invalid_argument("checkpoint fields do not match the grid");
    std::ofstream out(path);
    if (!out)
        throw std::runtime_error("cannot open checkpoint for writing");
    out << std::setprecision(17);
    const auto &p = parameters;
    out << "HZL_CHECKPOINT_V1\n"
        << n << ' ' << time << '\n'
        << p.viscosity << ' ' << p.friction << ' ' << p.forcing_sigma << ' ' << p.forcing_tau
        << ' ' << p.forcing_min << ' ' << p.forcing_max << '\n';
    for (std::size_t k = 0; k < kEngineWords; ++k)
        out << random_engine[k] << ' ';
    out << random_engine[kEngineWords] << '\n' << (random_has_spare ? 1 : 0) << ' ' << random_spare
        << '\n';
    for (std::size_t i = 0; i < n * n; ++i)
        out << vorticity[i].real() << ' ' << vorticity[i].imag() << ' ' << forcing[i].real() << ' '
            << forcing[i].imag() << '\n';
    if (!out)
        throw std::runtime_error("checkpoint write failed");
}

Checkpoint Checkpoint::load(const std::filesystem::path &path) {
    std::ifstream in(path);
    std::string magic;
    Checkpoint c;
    Parameters &p = c.parameters;
    if (!(in >> magic >> c.n >> c.time >> p.viscosity >> p.friction >> p.forcing_sigma >>
          p.forcing_tau >> p.forcing_min >> p.forcing_max) ||
        magic != "HZL_CHECKPOINT_V1" || !std::isfinite(c.time) || c.time < 0)
        throw std::runtime_error("invalid checkpoint header");
    Fourier2D grid_check(c.n); // power-of-two grid, as the reference's Solver(n, p, 0)
    p.validate(c.n);
    const auto forced = shell_indices(c.n, p);
    c.random_engine.resize(kEngineWords + 1);
    for (auto &w : c.random_engine)
        if (!(in >> w))
            throw std::runtime_error("invalid checkpoint random stream");
    if (c.random_engine[kEngineWords] > kEngineWords ||
        !(in >> c.random_has_spare >> c.random_spare) || !std::isfinite(c.random_spare))
        throw std::runtime_error("invalid checkpoint random stream");
    const std::size_t n = c.n;
    c.vorticity.assign(n * n, Complex{});
    c.forcing.assign(n * n, Complex{});
    for (std::size_t i = 0; i < n * n; ++i) {
        double a = 0, b = 0, f = 0, g = 0;
        if (!(in >> a >> b >> f >> g) || !std::isfinite(a) || !std::isfinite(b) ||
            !std::isfinite(f) || !std::isfinite(g))
            throw std::runtime_error("invalid checkpoint field");
        c.vorticity[i] = {a, b};
        c.forcing[i] = {f, g};
    }
    for (std::size_t i = 0; i < n * n; ++i) {
        const std::size_t y = i / n, x = i % n, partner = ((n - y) % n) * n + (n - x) % n;
        const Complex w = c.vorticity[i], F = c.forcing[i];
        if ((!retained(n, i) && w != Complex{}) || (!forced[i] && F != Complex{}) ||
            std::abs(w - std::conj(c.vorticity[partner])) > 1e-10 * (1 + std::abs(w)) ||
            F != std::conj(c.forcing[partner]))
            throw std::runtime_error("checkpoint violates spectral constraints");
    }
    std::string trailing;
    if (in >> trailing)
        throw std::runtime_error("unexpected checkpoint data");
    return c;
}

Dataset Dataset::load(const std::filesystem::path &path) {
    std::ifstream in(path);
    std::string line, magic;
    Dataset data;
    if (!std::getline(in, line))
        throw std::runtime_error("cannot read observation dataset");
    std::replace(line.begin(), line.end(), ',', ' ');
    std::istringstream header(line);
    if (!(header >> magic >> data.n >> data.dt) ||
        (magic != "HZL_OBSERVATIONS_V1" && magic != "HZL_OBSERVATIONS_V2"))
        throw std::runtime_error("invalid dataset header");
    if (magic == "HZL_OBSERVATIONS_V2" &&
        !(header >> data.stochastic >> data.forcing_min >> data.forcing_max))
        throw std::runtime_error("invalid stochastic dataset metadata");
    if (!std::getline(in, line) || line != "step,x,y,vorticity,standard_deviation")
        throw std::runtime_error("invalid dataset columns");
    while (std::getline(in, line)) {
        if (line.empty())
            continue;
        std::replace(line.begin(), line.end(), ',', ' ');
        std::istringstream row(line);
        Observation o;
        std::string extra;
        if (!(row >> o.step >> o.x >> o.y >> o.value >> o.standard_deviation) || (row >> extra))
            throw std::runtime_error("invalid observation row");
        data.observations.push_back(o);
    }
    data.validate();
    return data;
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
        FilterRecord r{gr.step, rec[0], rec[1], rec[2], {}};
        if (retain_states)
            r.mean_vorticity = physical(n, std::move(mean));
        result.history.push_back(std::move(r));
        current = gr.step;
        if (!std::isfinite(rec[0]))
            break;
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

namespace {

constexpr double negative_infinity = -std::numeric_limits<double>::infinity();

// NormalStream (src/solver.cpp:41-54): mt19937_64, open-interval uniforms,
// Box-Muller returning r cos and caching r sin.
class NormalStream {
  public:
    explicit NormalStream(std::uint64_t seed) : engine_(seed) {}
    double operator()() {
        if (has_spare_) {
            has_spare_ = false;
            return spare_;
        }
        const auto open = [this] { return ((double)(engine_() >> 11) + 0.5) * 0x1.0p-53; };
        const double r = std::sqrt(-2 * std::log(open())), angle = 2 * std::numbers::pi * open();
        spare_ = r * std::sin(angle);
        has_spare_ = true;
        return r * std::cos(angle);
    }

  private:
    std::mt19937_64 engine_;
    bool has_spare_ = false;
    double spare_ = 0;
};

// decode (src/filter.cpp:15-24)
Parameters decode(const Dataset &data, const Vector &q) {
    Parameters p;
    p.viscosity = std::exp(q[0]);
    p.friction = std::exp(q[1]);
    p.forcing_sigma = std::exp(q[2]);
    p.forcing_tau = std::exp(q[3]);
    p.forcing_min = data.forcing_min;
    p.forcing_max = data.forcing_max;
    return p;
}

} // namespace

double stochastic_log_prior(const Vector &q) {
    if (q.size() != 4)
        throw std::invalid_argument(
            "stochastic inference expects log(nu), log(alpha), log(sigma), log(tau)");
    const double centers[] = {std::log(0.03), std::log(0.1), std::log(0.2), std::log(0.5)};
    double result = 0;
    for (std::size_t i = 0; i < q.size(); ++i) {
        if (!std::isfinite(q[i]) || q[i] < -10 || q[i] > (i < 2 ? 0 : 2))
            return negative_infinity;
        const double z = (q[i] - centers[i]) / 1.5;
        result -= 0.5 * z * z;
    }
    return result;
}

double particle_log_likelihood(const Dataset &data, const Vector &q, std::size_t particles,
                               std::uint64_t seed) {
    if (!std::isfinite(stochastic_log_prior(q)))
        return negative_infinity;
    const Parameters p = decode(data, q);
    const double largest = (double)((data.n - 1) / 3);
    if (data.dt * (2 * p.viscosity * largest * largest + p.friction) > 2.5)
        return negative_infinity;
    return bootstrap_filter(data, p, particles, seed, false).log_likelihood;
}

double averaged_particle_log_likelihood(const Dataset &data, const Vector &q,
                                        std::size_t particles, std::size_t replicates,
                                        std::uint64_t seed) {
    if (replicates == 0)
        throw std::invalid_argument("averaged particle likelihood needs at least one replicate");
    Vector values(replicates);
    for (std::size_t i = 0; i < replicates; ++i)
        values[i] = particle_log_likelihood(data, q, particles, stream_seed(seed, i));
    const double maximum = *std::max_element(values.begin(), values.end());
    if (!std::isfinite(maximum))
        return negative_infinity;
    double total = 0;
    for (double v : values)
        total += std::exp(v - maximum);
    return maximum + std::log(total / (double)replicates);
}

// Pseudo-marginal random-walk Metropolis (src/filter.cpp:166-203): proposals
// from NormalStream(seed), likelihood seeds and acceptance uniforms from
// mt19937_64(stream_seed(seed, 1)); a rejected proposal keeps the current
// likelihood estimate and its seed.  Each likelihood is one GPU filter.
std::vector<ParticleSample> particle_metropolis(const Dataset &data, Vector current,
                                                double scale, std::size_t count,
                                                std::size_t particles, std::uint64_t seed) {
    if (!data.stochastic)
        throw std::invalid_argument("particle sampler expects a stochastic observation dataset");
    if (count == 0 || !std::isfinite(scale) || scale <= 0)
        throw std::invalid_argument("invalid particle sampler configuration");
    NormalStream normal(seed);
    std::mt19937_64 random(stream_seed(seed, 1));
    const auto uniform = [&random] { return (double)(random() >> 11) * 0x1.0p-53; };
    auto likelihood_seed = random();
    double likelihood = particle_log_likelihood(data, current, particles, likelihood_seed);
    double density = likelihood + stochastic_log_prior(current);
    if (!std::isfinite(density))
        throw std::invalid_argument("initial particle posterior estimate is not finite");
    std::vector<ParticleSample> chain;
    chain.reserve(count);
    for (std::size_t i = 0; i < count; ++i) {
        auto proposal = current;
        for (auto &q : proposal)
            q += scale * normal();
        const auto proposed_seed = random();
        const double proposed_likelihood =
            particle_log_likelihood(data, proposal, particles, proposed_seed);
        const double proposed_density = proposed_likelihood + stochastic_log_prior(proposal);
        const double draw = uniform();
        const bool accepted = std::isfinite(proposed_density) &&
                              std::log(std::max(draw, std::numeric_limits<double>::min())) <
                                  std::min(0.0, proposed_density - density);
        if (accepted) {
            current = std::move(proposal);
            likelihood = proposed_likelihood;
            density = proposed_density;
            likelihood_seed = proposed_seed;
        }
        chain.push_back({current, likelihood, density, likelihood_seed, accepted});
    }
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
