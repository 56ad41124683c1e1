// hzl_cli.cpp -- command-line front end of the GPU drop-in (libhzl_gpu.so) for
// three flows of the reference executable: `ensemble`, `filter` and
// `particle-sample` (reference src/main.cpp:384-526 define their contract).
//
// Every command is a row of the table in commands(): its options (name, type,
// default) in the order their values are checked, an optional check that runs
// before anything is written, the default output directory, and the action.
// The contract kept from the reference, so scripts can switch executables:
//   * arguments after the command are unique `--name value` pairs;
//   * an option the command does not declare is an error, reported after the
//     values of the declared ones have been checked;
//   * the output directory must be new or empty; it receives invocation.json
//     ({"schema":1,"command":..,"options":{name: value in force}}, values as
//     given or as the default is spelled by std::to_string);
//   * the same output files, columns, summary keys and stdout lines;
//   * failures print `error: <message>` on stderr and exit with status 1.
// The reference's other commands (experiments, ROM, calibration) are not part
// of this engine and are reported as unknown commands.
#include "hzl_gpu.hpp"
#include "text_io.hpp"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <functional>
#include <iostream>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <string_view>
#include <variant>
#include <vector>

namespace {

namespace fs = std::filesystem;
using hzl::text::fixed6;
using hzl::text::JsonObject;
using hzl::text::LineWriter;

enum class Kind { text, count, seed, integer, real };
// seeds are unsigned long long, counts std::size_t (distinct types on LP64)
using Value = std::variant<std::string, std::size_t, unsigned long long, int, double>;

struct OptionSpec {
    std::string name; // without the leading "--"
    Kind kind;
    Value fallback;
};

// The spelling of a default in invocation.json (std::to_string's).
std::string spell(const Value &v) {
    return std::visit(
        [](const auto &x) -> std::string {
            using T = std::decay_t<decltype(x)>;
            if constexpr (std::is_same_v<T, std::string>)
                return x;
            else if constexpr (std::is_same_v<T, double>)
                return fixed6(x);
            else
                return hzl::text::number(x);
        },
        v);
}

Value parse(const OptionSpec &o, const std::string &text) {
    const std::string flag = "--" + o.name;
    const auto typed = [&](auto zero) -> Value {
        using T = decltype(zero);
        T v{};
        // from_chars over the whole token, as the reference parses option values
        const auto r = std::from_chars(text.data(), text.data() + text.size(), v);
        if (r.ec != std::errc{} || r.ptr != text.data() + text.size())
            throw std::invalid_argument("invalid value for " + flag);
        if constexpr (std::is_floating_point_v<T>)
            if (!std::isfinite(v))
                throw std::invalid_argument("nonfinite value for " + flag);
        return v;
    };
    switch (o.kind) {
    case Kind::text:
        return text;
    case Kind::count:
        return typed(std::size_t{});
    case Kind::seed:
        return typed((unsigned long long)0);
    case Kind::integer:
        return typed(int{});
    case Kind::real:
        return typed(double{});
    }
    throw std::logic_error("option kind");
}

// Resolved values of one invocation, by option name.
class Values {
  public:
    template <class T> const T &get(const std::string &name) const {
        return std::get<T>(values_.at(name));
    }
    double real(const std::string &name) const { return get<double>(name); }
    std::size_t count(const std::string &name) const { return get<std::size_t>(name); }
    std::uint64_t seed(const std::string &name) const { return get<unsigned long long>(name); }
    const std::string &text(const std::string &name) const { return get<std::string>(name); }
    void set(const std::string &name, Value v, std::string spelled) {
        values_[name] = std::move(v);
        shown_[name] = std::move(spelled);
    }
    const std::map<std::string, std::string> &shown() const { return shown_; }

    hzl::Parameters physics() const {
        hzl::Parameters p;
        p.viscosity = real("nu");
        p.friction = real("alpha");
        p.forcing_sigma = real("sigma");
        p.forcing_tau = real("tau");
        p.forcing_min = get<int>("kmin");
        p.forcing_max = get<int>("kmax");
        return p;
    }

  private:
    std::map<std::string, Value> values_;
    std::map<std::string, std::string> shown_;
};

struct Command {
    std::string name;
    std::string default_output;
    std::vector<OptionSpec> options;
    std::function<void(const Values &)> check; // before anything is written
    std::function<void(const Values &, const fs::path &)> run;
};

// The six model parameters every command that builds solvers accepts.
std::vector<OptionSpec> physics_options() {
    const hzl::Parameters d;
    return {{"nu", Kind::real, d.viscosity},        {"alpha", Kind::real, d.friction},
            {"sigma", Kind::real, d.forcing_sigma}, {"tau", Kind::real, d.forcing_tau},
            {"kmin", Kind::integer, d.forcing_min}, {"kmax", Kind::integer, d.forcing_max}};
}

std::vector<OptionSpec> concat(std::vector<OptionSpec> a, const std::vector<OptionSpec> &b) {
    a.insert(a.end(), b.begin(), b.end());
    return a;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

LineWriter csv(const fs::path &path) {
    return LineWriter(path, ',', "cannot write " + path.string());
}

void emit(const fs::path &path, const std::string &text) {
    hzl::text::write_file(path, text, "cannot write " + path.string());
}

// Physical field of a record: rows of (ix, iy, value) in y-major order.
template <class Prefix>
void field_rows(LineWriter &out, std::size_t n, const hzl::Field &f, Prefix prefix) {
    for (std::size_t y = 0; y < n; ++y)
        for (std::size_t x = 0; x < n; ++x) {
            prefix(out);
            out.line(x, y, f[y * n + x].real());
        }
}

// Sample mean and the standard error of the mean (sample variance / count).
struct MeanError {
    double mean = 0, standard_error = 0;
};
MeanError mean_error(const std::vector<double> &v) {
    for (double x : v)
        if (!std::isfinite(x))
            throw std::invalid_argument("nonfinite statistical sample");
    MeanError r;
    double sum = 0;
    for (double x : v)
        sum += x;
    r.mean = v.empty() ? 0 : sum / (double)v.size();
    double ss = 0;
    for (double x : v)
        ss += (x - r.mean) * (x - r.mean);
    const double variance =
        v.size() > 1 ? ss / (double)(v.size() - 1) : std::numeric_limits<double>::quiet_NaN();
    r.standard_error = std::sqrt(variance / (double)v.size());
    return r;
}

// ------------------------------------------------------------- ensemble --
void run_ensemble(const Values &a, const fs::path &dir) {
    const std::size_t n = a.count("n"), steps = a.count("steps"), count = a.count("count"),
                      workers = a.count("workers");
    const double dt = a.real("dt");
    const std::uint64_t seed = a.seed("seed");
    const hzl::Parameters p = a.physics();
    const auto t0 = std::chrono::steady_clock::now();
    const auto members = hzl::ensemble(n, p, dt, steps, count, workers, seed);
    const double seconds = seconds_since(t0);

    auto table = csv(dir / "members.csv");
    table.line("member", "seed", "time", "energy", "enstrophy", "dissipation");
    std::vector<double> energies;
    for (std::size_t i = 0; i < members.size(); ++i) {
        const hzl::Diagnostics &d = members[i];
        table.line(i, hzl::stream_seed(seed, i), d.time, d.energy, d.enstrophy, d.dissipation);
        energies.push_back(d.energy);
    }
    table.close();
    const MeanError e = mean_error(energies);
    p.validate(n);

    JsonObject meta;
    meta.value("schema", 1)
        .value("version", "0.1.0")
        .value("backend", "mi355x-gfx950-hip-fp64")
        .value("compiler", __VERSION__)
        .value("n", n)
        .value("dt", dt)
        .value("integrator", "rk4")
        .value("steps_this_run", steps)
        .value("seed", seed)
        .value("initial", "zero")
        .value("nu", p.viscosity)
        .value("alpha", p.friction)
        .value("sigma", p.forcing_sigma)
        .value("tau", p.forcing_tau)
        .value("kmin", p.forcing_min)
        .value("kmax", p.forcing_max)
        .value("final_time", dt * (double)steps)
        .value("wall_seconds_including_diagnostics", seconds);
    emit(dir / "metadata.json", meta.pretty());

    JsonObject summary;
    summary.value("count", count)
        .value("workers", workers)
        .value("mean_energy", e.mean)
        .raw("energy_standard_error", count > 1 ? fixed6(e.standard_error) : "null")
        .value("wall_seconds", seconds);
    emit(dir / "summary.json", summary.compact());
    std::cout << "ensemble complete: " << dir << "; energy mean=" << e.mean
              << ", SE=" << e.standard_error << ", seconds=" << seconds << '\n';
}

// --------------------------------------------------------------- filter --
void run_filter(const Values &a, const fs::path &dir) {
    const std::size_t count = a.count("particles");
    const std::uint64_t seed = a.seed("seed");
    const std::string &method = a.text("method");
    const hzl::Parameters p = a.physics();
    const hzl::Dataset data = hzl::Dataset::load(a.text("data"));
    data.save(dir / "observations.csv");
    const std::size_t n = data.n;
    const auto t0 = std::chrono::steady_clock::now();

    if (method == "enkf") {
        hzl::EnKFOptions o;
        o.inflation = a.real("inflation");
        o.localization_radius = a.real("localization-radius");
        o.forecast_steps = a.count("forecast-steps");
        o.forecast_every = a.count("forecast-every");
        const auto history = hzl::ensemble_kalman_filter(data, p, count, seed, o);
        auto records = csv(dir / "filter.csv");
        auto fields = csv(dir / "reconstruction.csv");
        records.line("step", "phase", "forecast_sensor_rmse", "posterior_spread");
        fields.line("step", "phase", "ix", "iy", "vorticity");
        for (const auto &r : history) {
            const char *phase = r.is_forecast ? "forecast" : "analysis";
            records.add(r.step).add(phase);
            if (r.is_forecast)
                records.skip(); // free forecasts have no sensor RMSE
            else
                records.add(r.forecast_sensor_rmse);
            records.add(r.posterior_spread).end();
            field_rows(fields, n, r.mean_vorticity,
                       [&](LineWriter &w) { w.add(r.step).add(phase); });
        }
        records.close();
        fields.close();
        JsonObject summary;
        summary.value("method", "serial perturbed-observation EnKF")
            .value("members", count)
            .value("seed", seed)
            .value("inflation", o.inflation)
            .value("localization_radius_grid_cells", o.localization_radius)
            .value("forecast_steps", o.forecast_steps)
            .value("forecast_every", o.forecast_every)
            .value("wall_seconds", seconds_since(t0));
        emit(dir / "summary.json", summary.compact());
        std::cout << "ensemble Kalman reconstruction complete: " << dir << '\n';
        return;
    }
    if (method != "particle")
        throw std::invalid_argument("filter method must be particle or enkf");

    const hzl::FilterResult result = hzl::bootstrap_filter(data, p, count, seed);
    if (!std::isfinite(result.log_likelihood))
        throw std::runtime_error("particle likelihood collapsed numerically");
    auto records = csv(dir / "filter.csv");
    auto fields = csv(dir / "reconstruction.csv");
    records.line("step", "log_likelihood_increment", "effective_particles", "maximum_weight");
    fields.line("step", "ix", "iy", "vorticity");
    for (const auto &r : result.history) {
        records.line(r.step, r.log_likelihood_increment, r.effective_particles, r.maximum_weight);
        field_rows(fields, n, r.mean_vorticity, [&](LineWriter &w) { w.add(r.step); });
    }
    records.close();
    fields.close();
    JsonObject summary;
    summary.value("particles", count)
        .value("seed", seed)
        .value("log_likelihood_estimate", result.log_likelihood)
        .value("resampling_events", result.resampling_events)
        .value("wall_seconds", seconds_since(t0));
    emit(dir / "summary.json", summary.compact());
    std::cout << "particle filter complete: " << dir
              << "; log likelihood estimate=" << result.log_likelihood
              << ", resampling events=" << result.resampling_events << '\n';
}

// ------------------------------------------------------ particle-sample --
const char *const kInitial[4] = {"initial-nu", "initial-alpha", "initial-sigma", "initial-tau"};

void check_particle_sample(const Values &a) {
    for (const char *name : kInitial)
        if (!(a.real(name) > 0))
            throw std::invalid_argument("initial stochastic parameters must be positive");
}

void run_particle_sample(const Values &a, const fs::path &dir) {
    const std::size_t samples = a.count("samples"), particles = a.count("particles");
    hzl::Vector q0;
    for (const char *name : kInitial)
        q0.push_back(std::log(a.real(name)));
    const hzl::Dataset data = hzl::Dataset::load(a.text("data"));
    data.save(dir / "training.csv");
    const auto chain = hzl::particle_metropolis(data, q0, a.real("proposal"), samples, particles,
                                                a.seed("seed"));
    auto out = csv(dir / "chain.csv");
    out.line("iteration", "nu", "alpha", "sigma", "tau", "log_likelihood", "log_posterior",
             "likelihood_seed", "accepted");
    std::size_t accepted = 0;
    for (std::size_t i = 0; i < chain.size(); ++i) {
        const hzl::ParticleSample &s = chain[i];
        out.add(i);
        for (double q : s.parameters)
            out.add(std::exp(q));
        out.line(s.log_likelihood, s.log_posterior, s.likelihood_seed, s.accepted);
        accepted += s.accepted ? 1 : 0;
    }
    out.close();
    const double rate = (double)accepted / (double)samples;
    JsonObject summary;
    summary.value("samples", samples)
        .value("particles", particles)
        .value("acceptance_rate", rate)
        .value("method", "pseudo-marginal random-walk Metropolis");
    emit(dir / "summary.json", summary.compact());
    std::cout << "particle posterior chain written: " << dir << "; acceptance=" << rate << '\n';
}

const std::vector<Command> &commands() {
    static const std::vector<Command> table = {
        {"ensemble",
         "results/ensemble",
         concat({{"n", Kind::count, std::size_t{32}},
                 {"steps", Kind::count, std::size_t{200}},
                 {"count", Kind::count, std::size_t{16}},
                 {"workers", Kind::count, std::size_t{4}},
                 {"dt", Kind::real, 0.005},
                 {"seed", Kind::seed, 42ULL}},
                physics_options()),
         nullptr,
         run_ensemble},
        {"filter",
         "results/filter",
         concat({{"data", Kind::text, std::string("results/observations/observations.csv")},
                 {"particles", Kind::count, std::size_t{64}},
                 {"method", Kind::text, std::string("particle")},
                 {"seed", Kind::seed, 2026ULL},
                 {"inflation", Kind::real, 1.0},
                 {"localization-radius", Kind::real, 0.0},
                 {"forecast-steps", Kind::count, std::size_t{0}},
                 {"forecast-every", Kind::count, std::size_t{1}}},
                physics_options()),
         nullptr,
         run_filter},
        {"particle-sample",
         "results/particle-posterior",
         {{"data", Kind::text, std::string("results/observations/training.csv")},
          {"samples", Kind::count, std::size_t{500}},
          {"particles", Kind::count, std::size_t{32}},
          {"seed", Kind::seed, 123ULL},
          {"proposal", Kind::real, 0.05},
          {"initial-nu", Kind::real, 0.03},
          {"initial-alpha", Kind::real, 0.1},
          {"initial-sigma", Kind::real, 0.2},
          {"initial-tau", Kind::real, 0.5}},
         check_particle_sample,
         run_particle_sample},
    };
    return table;
}

void usage() {
    std::cout << "HZL stochastic turbulence engine (MI355X) 0.1.0\nCommands:\n";
    for (const Command &c : commands()) {
        std::cout << "  " << c.name;
        for (const OptionSpec &o : c.options)
            std::cout << " --" << o.name << ' ' << spell(o.fallback);
        std::cout << '\n';
    }
    std::cout << "All commands: --output NEW_DIRECTORY\n";
}

int dispatch(int argc, char **argv) {
    if (argc < 2 || std::string_view(argv[1]) == "--help") {
        usage();
        return 0;
    }
    // unique --name value pairs
    std::map<std::string, std::string> given;
    for (int i = 2; i < argc; i += 2) {
        const std::string key = argv[i];
        if (key.size() < 2 || key.compare(0, 2, "--") != 0 || i + 1 >= argc ||
            !given.emplace(key, argv[i + 1]).second)
            throw std::invalid_argument("options require unique --name value pairs");
    }
    const Command *command = nullptr;
    for (const Command &c : commands())
        if (c.name == argv[1])
            command = &c;
    if (!command)
        throw std::invalid_argument("unknown command: " + std::string(argv[1]));

    Values values;
    std::map<std::string, bool> declared;
    for (const OptionSpec &o : command->options) {
        const auto it = given.find("--" + o.name);
        const std::string shown = it == given.end() ? spell(o.fallback) : it->second;
        values.set(o.name, parse(o, shown), shown);
        declared["--" + o.name] = true;
    }
    if (command->check)
        command->check(values);
    const auto out = given.find("--output");
    const std::string dir = out == given.end() ? command->default_output : out->second;
    values.set("output", dir, dir);
    declared["--output"] = true;
    if (fs::exists(dir) && !fs::is_empty(dir))
        throw std::invalid_argument(
            "output directory is not empty; choose a new --output directory");
    for (const auto &kv : given)
        if (!declared.count(kv.first))
            throw std::invalid_argument("unknown option " + kv.first);

    fs::create_directories(dir);
    JsonObject options;
    for (const auto &kv : values.shown())
        options.value(kv.first, kv.second);
    std::string invocation = options.compact();
    invocation.pop_back(); // the object is nested: drop its newline
    emit(fs::path(dir) / "invocation.json",
         "{\"schema\":1,\"command\":" + hzl::text::json_string(command->name) +
             ",\"options\":" + invocation + "}\n");
    command->run(values, dir);
    return 0;
}

} // namespace

int main(int argc, char **argv) {
    try {
        return dispatch(argc, argv);
    } catch (const std::exception &error) {
        std::cerr << "error: " << error.what() << '\n';
        return 1;
    }
}
