// hzl_cli.cpp -- the reference CLI's `ensemble`, `filter` and `particle-sample`
// commands (src/main.cpp:384-526) over the GPU drop-in layer (libhzl_gpu.so).
//
// Same command names, options, defaults, output directory rules, files and
// their columns as the reference's flows, so scripts that drive those commands
// switch to the GPU engine by changing the executable's name.  Options are
// unique `--name value` pairs; an option the command does not read is an
// error; the output directory must be new or empty and receives
// invocation.json with every resolved option.  Errors print `error: <what>` on
// stderr and exit 1.  The other reference commands (calibration, ROM, Monte
// Carlo experiments) are outside this engine and are rejected as unknown.
#include "hzl_gpu.hpp"

#include <charconv>
#include <chrono>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>

namespace {

// --name value pairs; get() records the value in force (given or default),
// finish() rejects options no one read (src/main.cpp:23-95).
class Options {
  public:
    Options(int argc, char **argv) : command_(argv[1]) {
        for (int i = 2; i < argc; i += 2) {
            const std::string key = argv[i];
            if (key.rfind("--", 0) != 0 || i + 1 >= argc || given_.count(key))
                throw std::invalid_argument("options require unique --name value pairs");
            given_[key] = argv[i + 1];
        }
    }
    std::string get(const std::string &key, const std::string &fallback) {
        read_.insert(key);
        const auto it = given_.find(key);
        const std::string value = it == given_.end() ? fallback : it->second;
        resolved_[key.substr(2)] = value;
        return value;
    }
    template <class T> T number(const std::string &key, T fallback) {
        const std::string text = get(key, std::to_string(fallback));
        T value{};
        const auto [end, error] = std::from_chars(text.data(), text.data() + text.size(), value);
        if (error != std::errc{} || end != text.data() + text.size())
            throw std::invalid_argument("invalid value for " + key);
        if constexpr (std::is_floating_point_v<T>)
            if (!std::isfinite(value))
                throw std::invalid_argument("nonfinite value for " + key);
        return value;
    }
    void finish() const {
        for (const auto &kv : given_)
            if (!read_.count(kv.first))
                throw std::invalid_argument("unknown option " + kv.first);
    }
    // invocation.json: {"schema":1,"command":...,"options":{resolved, by name}}
    void record(const std::filesystem::path &path) const {
        std::ofstream out(path);
        const auto quoted = [&out](const std::string &v) {
            out << '"';
            for (unsigned char c : v) {
                if (c == '"' || c == '\\')
                    out << '\\' << (char)c;
                else if (c < 32)
                    out << "\\u" << std::hex << std::setw(4) << std::setfill('0') << (unsigned)c
                        << std::dec;
                else
                    out << (char)c;
            }
            out << '"';
        };
        out << "{\"schema\":1,\"command\":";
        quoted(command_);
        out << ",\"options\":{";
        bool first = true;
        for (const auto &kv : resolved_) {
            if (!first)
                out << ',';
            first = false;
            quoted(kv.first);
            out << ':';
            quoted(kv.second);
        }
        out << "}}\n";
        if (!out)
            throw std::runtime_error("cannot record invocation");
    }

  private:
    std::string command_;
    std::map<std::string, std::string> given_, resolved_;
    std::set<std::string> read_;
};

std::ofstream output(const std::filesystem::path &path) {
    std::ofstream out(path);
    out.exceptions(std::ios::badbit | std::ios::failbit);
    out << std::setprecision(17);
    return out;
}

// --output (default results/<name>) must be new or empty; every option has been
// read by now, so unknown ones fail before anything is written.
std::filesystem::path directory(Options &args, const std::string &name) {
    const std::filesystem::path path = args.get("--output", "results/" + name);
    if (std::filesystem::exists(path) && !std::filesystem::is_empty(path))
        throw std::invalid_argument(
            "output directory is not empty; choose a new --output directory");
    args.finish();
    std::filesystem::create_directories(path);
    args.record(path / "invocation.json");
    return path;
}

hzl::Parameters parameters(Options &args) {
    hzl::Parameters p;
    p.viscosity = args.number("--nu", p.viscosity);
    p.friction = args.number("--alpha", p.friction);
    p.forcing_sigma = args.number("--sigma", p.forcing_sigma);
    p.forcing_tau = args.number("--tau", p.forcing_tau);
    p.forcing_min = args.number("--kmin", p.forcing_min);
    p.forcing_max = args.number("--kmax", p.forcing_max);
    return p;
}

double seconds_since(std::chrono::steady_clock::time_point start) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
}

// Welford running mean / variance (src/inference.cpp:540-554).
struct Moments {
    std::size_t count = 0;
    double mean = 0, m2 = 0;
    void add(double x) {
        if (!std::isfinite(x))
            throw std::invalid_argument("nonfinite statistical sample");
        ++count;
        const double delta = x - mean;
        mean += delta / (double)count;
        m2 += delta * (x - mean);
    }
    double standard_error() const {
        const double var = count > 1 ? m2 / (double)(count - 1)
                                     : std::numeric_limits<double>::quiet_NaN();
        return std::sqrt(var / (double)count);
    }
};

// `ensemble` (src/main.cpp:384-416): independent members from zero vorticity.
void ensemble_command(Options &args) {
    const auto n = args.number<std::size_t>("--n", 32), steps = args.number<std::size_t>("--steps", 200),
               count = args.number<std::size_t>("--count", 16),
               workers = args.number<std::size_t>("--workers", 4);
    const double dt = args.number("--dt", 0.005);
    const auto seed = args.number<std::uint64_t>("--seed", 42);
    const hzl::Parameters p = parameters(args);
    const auto path = directory(args, "ensemble");
    args.finish();
    const auto start = std::chrono::steady_clock::now();
    const auto results = hzl::ensemble(n, p, dt, steps, count, workers, seed);
    const double seconds = seconds_since(start);
    auto members = output(path / "members.csv");
    members << "member,seed,time,energy,enstrophy,dissipation\n";
    Moments energy;
    for (std::size_t i = 0; i < results.size(); ++i) {
        const auto &d = results[i];
        energy.add(d.energy);
        members << i << ',' << hzl::stream_seed(seed, i) << ',' << d.time << ',' << d.energy
                << ',' << d.enstrophy << ',' << d.dissipation << '\n';
    }
    p.validate(n);
    auto meta = output(path / "metadata.json");
    meta << "{\n  \"schema\": 1,\n  \"version\": \"0.1.0\",\n  \"backend\": "
            "\"mi355x-gfx950-hip-fp64\",\n  \"compiler\": \""
         << __VERSION__ << "\",\n"
         << "  \"n\": " << n << ",\n  \"dt\": " << dt << ",\n  \"integrator\": \"rk4\""
         << ",\n  \"steps_this_run\": " << steps << ",\n  \"seed\": " << seed << ",\n"
         << "  \"initial\": \"zero\",\n  \"nu\": " << p.viscosity << ",\n  \"alpha\": "
         << p.friction << ",\n"
         << "  \"sigma\": " << p.forcing_sigma << ",\n  \"tau\": " << p.forcing_tau
         << ",\n  \"kmin\": " << p.forcing_min << ",\n  \"kmax\": " << p.forcing_max << ",\n"
         << "  \"final_time\": " << dt * (double)steps
         << ",\n  \"wall_seconds_including_diagnostics\": " << seconds << "\n}\n";
    auto summary = output(path / "summary.json");
    summary << "{\"count\":" << count << ",\"workers\":" << workers
            << ",\"mean_energy\":" << energy.mean << ",\"energy_standard_error\":"
            << (count > 1 ? std::to_string(energy.standard_error()) : "null")
            << ",\"wall_seconds\":" << seconds << "}\n";
    std::cout << "ensemble complete: " << path << "; energy mean=" << energy.mean
              << ", SE=" << energy.standard_error() << ", seconds=" << seconds << '\n';
}

// `filter` (src/main.cpp:417-485): bootstrap particle filter or serial EnKF.
void filter_command(Options &args) {
    const auto input = args.get("--data", "results/observations/observations.csv");
    const auto count = args.number<std::size_t>("--particles", 64);
    const auto method = args.get("--method", "particle");
    const auto seed = args.number<std::uint64_t>("--seed", 2026);
    hzl::EnKFOptions options;
    options.inflation = args.number("--inflation", 1.0);
    options.localization_radius = args.number("--localization-radius", 0.0);
    options.forecast_steps = args.number<std::size_t>("--forecast-steps", 0);
    options.forecast_every = args.number<std::size_t>("--forecast-every", 1);
    const hzl::Parameters p = parameters(args);
    const auto path = directory(args, "filter");
    args.finish();
    const auto data = hzl::Dataset::load(input);
    data.save(path / "observations.csv");
    const auto start = std::chrono::steady_clock::now();
    if (method == "enkf") {
        const auto history = hzl::ensemble_kalman_filter(data, p, count, seed, options);
        auto diagnostics = output(path / "filter.csv"),
             fields = output(path / "reconstruction.csv");
        diagnostics << "step,phase,forecast_sensor_rmse,posterior_spread\n";
        fields << "step,phase,ix,iy,vorticity\n";
        for (const auto &r : history) {
            const char *phase = r.is_forecast ? "forecast" : "analysis";
            diagnostics << r.step << ',' << phase << ',';
            if (!r.is_forecast)
                diagnostics << r.forecast_sensor_rmse;
            diagnostics << ',' << r.posterior_spread << '\n';
            for (std::size_t y = 0; y < data.n; ++y)
                for (std::size_t x = 0; x < data.n; ++x)
                    fields << r.step << ',' << phase << ',' << x << ',' << y << ','
                           << r.mean_vorticity[y * data.n + x].real() << '\n';
        }
        auto summary = output(path / "summary.json");
        summary << "{\"method\":\"serial perturbed-observation EnKF\",\"members\":" << count
                << ",\"seed\":" << seed << ",\"inflation\":" << options.inflation
                << ",\"localization_radius_grid_cells\":" << options.localization_radius
                << ",\"forecast_steps\":" << options.forecast_steps
                << ",\"forecast_every\":" << options.forecast_every
                << ",\"wall_seconds\":" << seconds_since(start) << "}\n";
        std::cout << "ensemble Kalman reconstruction complete: " << path << '\n';
        return;
    }
    if (method != "particle")
        throw std::invalid_argument("filter method must be particle or enkf");
    const auto result = hzl::bootstrap_filter(data, p, count, seed);
    if (!std::isfinite(result.log_likelihood))
        throw std::runtime_error("particle likelihood collapsed numerically");
    auto diagnostics = output(path / "filter.csv"), fields = output(path / "reconstruction.csv");
    diagnostics << "step,log_likelihood_increment,effective_particles,maximum_weight\n";
    fields << "step,ix,iy,vorticity\n";
    for (const auto &r : result.history) {
        diagnostics << r.step << ',' << r.log_likelihood_increment << ','
                    << r.effective_particles << ',' << r.maximum_weight << '\n';
        for (std::size_t y = 0; y < data.n; ++y)
            for (std::size_t x = 0; x < data.n; ++x)
                fields << r.step << ',' << x << ',' << y << ','
                       << r.mean_vorticity[y * data.n + x].real() << '\n';
    }
    auto summary = output(path / "summary.json");
    summary << "{\"particles\":" << count << ",\"seed\":" << seed
            << ",\"log_likelihood_estimate\":" << result.log_likelihood
            << ",\"resampling_events\":" << result.resampling_events
            << ",\"wall_seconds\":" << seconds_since(start) << "}\n";
    std::cout << "particle filter complete: " << path
              << "; log likelihood estimate=" << result.log_likelihood
              << ", resampling events=" << result.resampling_events << '\n';
}

// `particle-sample` (src/main.cpp:486-526): pseudo-marginal Metropolis chain.
void particle_sample_command(Options &args) {
    const auto input = args.get("--data", "results/observations/training.csv");
    const auto count = args.number<std::size_t>("--samples", 500),
               particles = args.number<std::size_t>("--particles", 32);
    const auto seed = args.number<std::uint64_t>("--seed", 123);
    const double scale = args.number("--proposal", 0.05);
    const double nu = args.number("--initial-nu", 0.03),
                 alpha = args.number("--initial-alpha", 0.1),
                 sigma = args.number("--initial-sigma", 0.2),
                 tau = args.number("--initial-tau", 0.5);
    if (nu <= 0 || alpha <= 0 || sigma <= 0 || tau <= 0)
        throw std::invalid_argument("initial stochastic parameters must be positive");
    const hzl::Vector initial{std::log(nu), std::log(alpha), std::log(sigma), std::log(tau)};
    const auto path = directory(args, "particle-posterior");
    args.finish();
    const auto data = hzl::Dataset::load(input);
    data.save(path / "training.csv");
    const auto chain = hzl::particle_metropolis(data, initial, scale, count, particles, seed);
    auto out = output(path / "chain.csv");
    out << "iteration,nu,alpha,sigma,tau,log_likelihood,log_posterior,likelihood_seed,accepted\n";
    std::size_t accepted = 0;
    for (std::size_t i = 0; i < chain.size(); ++i) {
        const auto &s = chain[i];
        accepted += s.accepted;
        out << i;
        for (double q : s.parameters)
            out << ',' << std::exp(q);
        out << ',' << s.log_likelihood << ',' << s.log_posterior << ',' << s.likelihood_seed
            << ',' << s.accepted << '\n';
    }
    const double rate = (double)accepted / (double)count;
    auto summary = output(path / "summary.json");
    summary << "{\"samples\":" << count << ",\"particles\":" << particles
            << ",\"acceptance_rate\":" << rate
            << ",\"method\":\"pseudo-marginal random-walk Metropolis\"}\n";
    std::cout << "particle posterior chain written: " << path << "; acceptance=" << rate
              << '\n';
}

void usage() {
    std::cout << "HZL stochastic turbulence engine (MI355X) 0.1.0\n"
              << "Commands: ensemble, filter, particle-sample\n"
              << "Parameters: --nu 0.01 --alpha 0.1 --sigma 0.2 --tau 0.5 --kmin 2 --kmax 3\n"
              << "ensemble: --n 32 --steps 200 --dt 0.005 --count 16 --workers 4 --seed 42 "
                 "(zero initial vorticity)\n"
              << "filter: --data observations.csv --particles 64 --method particle|enkf "
                 "--seed 2026 --inflation 1 --localization-radius 0 --forecast-steps 0 "
                 "--forecast-every 1 and physical parameters\n"
              << "particle-sample: --data training.csv --particles 32 --samples 500 "
                 "--proposal 0.05 --seed 123 --initial-nu 0.03 --initial-alpha 0.1 "
                 "--initial-sigma 0.2 --initial-tau 0.5\n"
              << "All commands: --output NEW_DIRECTORY\n";
}

} // namespace

int main(int argc, char **argv) {
    try {
        if (argc < 2 || std::string_view(argv[1]) == "--help") {
            usage();
            return 0;
        }
        Options args(argc, argv);
        const std::string command = argv[1];
        if (command == "ensemble")
            ensemble_command(args);
        else if (command == "filter")
            filter_command(args);
        else if (command == "particle-sample")
            particle_sample_command(args);
        else
            throw std::invalid_argument("unknown command: " + command);
        return 0;
    } catch (const std::exception &error) {
        std::cerr << "error: " << error.what() << '\n';
        return 1;
    }
}
