// pmcmc.hip -- batched particle-marginal Metropolis-Hastings and batched
// particle likelihoods (host drivers over the engine's segmented filters).
//
// Restates reference particle_metropolis (src/filter.cpp:166-203) for C chains
// at once: every chain is one particle filter over its own contiguous segment
// of the ensemble (pf_begin_filters), so one pass of the engine's batched RK4
// and filter kernels evaluates all C proposed likelihoods.  Chain c's stream
// seed s_c is chain_seeds[c] when given, else stream_seed(seed, c); from it,
// exactly the reference's streams:
//   proposals  q' = q + scale * NormalStream(s_c)() per component,
//   likelihood seeds and acceptance draws from mt19937_64(stream_seed(s_c, 1)),
//   accept iff the proposed density is finite and
//   log(max(u, DBL_MIN)) < min(0, density' - density); on rejection the
//   current likelihood estimate and its seed are kept (pseudo-marginal).
// The parameter vector is described by a Model: the viscosity and forcing
// amplitude posterior's q = (log nu, log Kolmogorov amplitude) with a
// log-uniform prior, or the reference's
// q = (log nu, log alpha, log sigma, log tau) with its bounded Gaussian prior
// (stochastic_log_prior, src/filter.cpp:127-140; decode, :15-24).  The
// explicit-step guard dt (2 nu K^2 + alpha) > 2.5 gives a -inf likelihood
// (src/filter.cpp:145-148).  Each likelihood evaluation restarts every particle
// from the same initial state at step 0.
#include "engine.hpp"

#include <array>
#include <cmath>
#include <limits>
#include <random>
#include <stdexcept>
#include <vector>

namespace hzl_amd {

namespace {

constexpr double NEG_INF = -std::numeric_limits<double>::infinity();
constexpr int MAX_DIMS = 4;
using Point = std::array<double, MAX_DIMS>;

// The reference NormalStream (src/solver.cpp:41-54): mt19937_64, open-interval
// uniforms, Box-Muller returning r cos and keeping r sin for the next call.
struct Normals {
    std::mt19937_64 engine;
    bool has_spare = false;
    double spare = 0;
    explicit Normals(uint64_t seed) : engine(seed) {}
    double operator()() {
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        const auto u = [this]() { return ((double)(engine() >> 11) + 0.5) * 0x1.0p-53; };
        const double r = std::sqrt(-2 * std::log(u())), angle = 2 * PI * u();
        spare = r * std::sin(angle);
        has_spare = true;
        return r * std::cos(angle);
    }
};

// One axis of the reference's stochastic prior: a Gaussian of width 1.5 in
// the log parameter around `center`, zero density outside [lower, upper].
struct Axis {
    double center, lower, upper;
};
const std::array<Axis, 4> &stochastic_axes() {
    static const std::array<Axis, 4> axes = {Axis{std::log(0.03), -10, 0},
                                             Axis{std::log(0.1), -10, 0},
                                             Axis{std::log(0.2), -10, 2},
                                             Axis{std::log(0.5), -10, 2}};
    return axes;
}
constexpr double STOCHASTIC_WIDTH = 1.5;

double stochastic_prior(const double *q) {
    double sum = 0;
    for (int i = 0; i < 4; ++i) {
        const Axis &a = stochastic_axes()[i];
        if (!(std::isfinite(q[i]) && q[i] >= a.lower && q[i] <= a.upper))
            return NEG_INF;
        const double z = (q[i] - a.center) / STOCHASTIC_WIDTH;
        sum -= 0.5 * z * z;
    }
    return sum;
}

// What a parameter vector means for the members of a filter.
struct Model {
    int prior;
    int dims;
    double lower[2], upper[2]; // log-uniform bounds
    double friction, sigma;    // fixed under the log-uniform prior
    double dt;
    int K;
    hzl_surrogate_t surrogate; // forecast model of every interval (null: the solver)
    int surrogate_precision;

    explicit Model(const hzl_pmcmc_options &o, double step, int k)
        : prior(o.prior), dims(o.prior == HZL_PRIOR_STOCHASTIC ? 4 : 2),
          lower{o.log_lower[0], o.log_lower[1]}, upper{o.log_upper[0], o.log_upper[1]},
          friction(o.friction), sigma(o.forcing_sigma), dt(step), K(k), surrogate(o.surrogate),
          surrogate_precision(o.surrogate_precision) {
        if (prior != HZL_PRIOR_LOG_UNIFORM && prior != HZL_PRIOR_STOCHASTIC)
            throw std::invalid_argument("prior must be HZL_PRIOR_LOG_UNIFORM or HZL_PRIOR_STOCHASTIC");
        if (prior == HZL_PRIOR_LOG_UNIFORM) {
            for (int i = 0; i < 2; ++i)
                if (!std::isfinite(lower[i]) || !std::isfinite(upper[i]) || lower[i] >= upper[i])
                    throw std::invalid_argument("prior bounds must be finite and ordered");
            if (!std::isfinite(friction) || friction < 0 || !std::isfinite(sigma) || sigma < 0)
                throw std::invalid_argument(
                    "friction and forcing sigma must be finite and nonnegative");
        }
    }
    double log_prior(const Point &q) const {
        if (prior == HZL_PRIOR_STOCHASTIC)
            return stochastic_prior(q.data());
        for (int i = 0; i < 2; ++i)
            if (!(std::isfinite(q[i]) && q[i] >= lower[i] && q[i] <= upper[i]))
                return NEG_INF;
        return 0.0;
    }
    double viscosity(const Point &q) const { return std::exp(q[0]); }
    double alpha(const Point &q) const {
        return prior == HZL_PRIOR_STOCHASTIC ? std::exp(q[1]) : friction;
    }
    // a likelihood exists: inside the prior and within the explicit-step limit
    bool admissible(const Point &q) const {
        if (!(log_prior(q) > NEG_INF))
            return false;
        return dt * (2 * viscosity(q) * (double)K * (double)K + alpha(q)) <= 2.5;
    }
    void apply(EngineBase &eng, size_t member, const Point &q) const {
        if (prior == HZL_PRIOR_STOCHASTIC) {
            eng.set_member_parameters(member, std::exp(q[0]), std::exp(q[1]), std::exp(q[2]), 0.0);
            eng.set_member_tau(member, std::exp(q[3]));
        } else {
            eng.set_member_parameters(member, std::exp(q[0]), friction, sigma, std::exp(q[1]));
        }
    }
};

// Observation rows split into groups of one step each.
struct Window {
    std::vector<size_t> step, begin; // begin has one entry more than step
    const int *x, *y;
    const double *value, *sd;

    Window(size_t nobs, const int *s, const int *x_, const int *y_, const double *v,
           const double *d)
        : x(x_), y(y_), value(v), sd(d) {
        if (nobs == 0)
            throw std::invalid_argument("observation window is empty");
        for (size_t i = 0; i < nobs; ++i) {
            if (s[i] < 0 || (i && s[i] < s[i - 1]))
                throw std::invalid_argument("observations must be time ordered from step 0");
            if (i == 0 || s[i] != s[i - 1]) {
                step.push_back((size_t)s[i]);
                begin.push_back(i);
            }
        }
        begin.push_back(nobs);
    }
};

// Log-likelihood estimates of F filters of P particles: filter f at point
// q[f] with filter seed seeds[f], every particle restarting from `initial`.
// A filter whose point is not admissible runs on `stand_in` (any admissible
// point, so the engine's time-step checks pass) and reports -inf.
void evaluate(EngineBase &eng, const Model &model, const Window &w, const double *initial,
              size_t F, size_t P, const std::vector<Point> &q, const std::vector<uint64_t> &seeds,
              const std::vector<char> &ok, const std::vector<Point> &stand_in,
              std::vector<double> &ll) {
    eng.fill_state(initial);
    eng.set_step_index(0);
    for (size_t f = 0; f < F; ++f)
        for (size_t i = 0; i < P; ++i)
            model.apply(eng, f * P + i, ok[f] ? q[f] : stand_in[f]);
    eng.pf_begin_filters(seeds.data(), F, w.step.size() + 1, true);
    size_t now = 0;
    for (size_t g = 0; g < w.step.size(); ++g) {
        const size_t b = w.begin[g], e = w.begin[g + 1];
        size_t steps = w.step[g] - now;
        if (model.surrogate && steps) {
            surrogate_forecast(model.surrogate, eng, model.dt, steps, model.surrogate_precision);
            steps = 0; // the network advanced the particles
        }
        eng.pf_assimilate(model.dt, steps, e - b, w.x + b, w.y + b, w.value + b, w.sd + b,
                          g + 1 == w.step.size(), nullptr, nullptr);
        now = w.step[g];
    }
    ll.assign(F, NEG_INF);
    eng.pf_result(ll.data(), nullptr);
    for (size_t f = 0; f < F; ++f)
        if (!ok[f])
            ll[f] = NEG_INF;
}

struct Chain {
    Normals normal;
    std::mt19937_64 random;
    Point q{};
    double likelihood = NEG_INF, density = NEG_INF;
    uint64_t seed = 0;
    explicit Chain(uint64_t s) : normal(s), random(stream_seed(s, 1)) {}
    double uniform() { return (double)(random() >> 11) * 0x1.0p-53; }
};

void check_window_args(const void *step, const void *x, const void *y, const void *value,
                       const void *sd, const void *initial) {
    if (!step || !x || !y || !value || !sd || !initial)
        throw std::invalid_argument("null argument");
}

} // namespace

} // namespace hzl_amd

using hzl_amd::engine_of;
using hzl_amd::guarded;

extern "C" int hzl_normal_draws(uint64_t seed, size_t count, double *out) {
    return guarded([&] {
        if (count && !out)
            throw std::invalid_argument("null argument");
        hzl_amd::Normals normal(seed);
        for (size_t i = 0; i < count; ++i)
            out[i] = normal();
    });
}

extern "C" int hzl_stochastic_log_prior(const double *q, size_t dims, double *out) {
    return guarded([&] {
        if (!out || (dims && !q))
            throw std::invalid_argument("null argument");
        if (dims != 4)
            throw std::invalid_argument(
                "stochastic inference expects log(nu), log(alpha), log(sigma), log(tau)");
        *out = hzl_amd::stochastic_prior(q);
    });
}

extern "C" int hzl_pf_log_likelihoods(hzl_ensemble_t h, const hzl_pmcmc_options *o, double dt,
                                      size_t nobs, const int *step, const int *x, const int *y,
                                      const double *value, const double *sd,
                                      const double *initial_field, size_t filters,
                                      const double *q, const uint64_t *seeds,
                                      double *log_likelihood) {
    return guarded([&] {
        using namespace hzl_amd;
        if (!o || !q || !seeds || !log_likelihood)
            throw std::invalid_argument("null argument");
        check_window_args(step, x, y, value, sd, initial_field);
        EngineBase &eng = engine_of(h);
        const EngineViews v = eng.views();
        if (filters == 0 || (size_t)v.B % filters != 0 || (size_t)v.B / filters < 2)
            throw std::invalid_argument("members must split into filters of at least two particles");
        const Model model(*o, dt, v.K);
        const Window w(nobs, step, x, y, value, sd);
        std::vector<Point> points(filters), stand_in(filters);
        std::vector<uint64_t> s(seeds, seeds + filters);
        std::vector<char> ok(filters);
        const Point *any = nullptr;
        for (size_t f = 0; f < filters; ++f) {
            for (int i = 0; i < model.dims; ++i)
                points[f][i] = q[f * model.dims + i];
            ok[f] = model.admissible(points[f]);
            if (ok[f] && !any)
                any = &points[f];
        }
        std::vector<double> ll(filters, NEG_INF);
        if (any) {
            for (auto &p : stand_in)
                p = *any;
            evaluate(eng, model, w, initial_field, filters, (size_t)v.B / filters, points, s, ok,
                     stand_in, ll);
        }
        for (size_t f = 0; f < filters; ++f)
            log_likelihood[f] = ll[f];
    });
}

extern "C" int hzl_pmcmc_run(hzl_ensemble_t h, const hzl_pmcmc_options *o, double dt,
                             size_t nobs, const int *step, const int *x, const int *y,
                             const double *value, const double *sd, const double *initial_field,
                             const double *initial_q, double *samples_q, double *samples_ll,
                             double *samples_density, int *samples_accepted,
                             uint64_t *samples_seed) {
    return guarded([&] {
        using namespace hzl_amd;
        if (!o || !initial_q || !samples_q || !samples_ll)
            throw std::invalid_argument("null argument");
        check_window_args(step, x, y, value, sd, initial_field);
        EngineBase &eng = engine_of(h);
        const EngineViews v = eng.views();
        if (o->chains == 0 || (size_t)v.B % o->chains != 0 || (size_t)v.B / o->chains < 2)
            throw std::invalid_argument("members must split into chains of at least two particles");
        if (o->samples == 0 || !std::isfinite(o->proposal_scale) || o->proposal_scale <= 0)
            throw std::invalid_argument("invalid particle sampler configuration");
        const Model model(*o, dt, v.K);
        const Window w(nobs, step, x, y, value, sd);
        const size_t C = o->chains, P = (size_t)v.B / C, S = o->samples;
        const int D = model.dims;

        std::vector<Chain> chains;
        chains.reserve(C);
        std::vector<Point> current(C), proposed(C);
        std::vector<uint64_t> seeds(C);
        std::vector<char> ok(C);
        std::vector<double> ll;
        for (size_t c = 0; c < C; ++c) {
            chains.emplace_back(o->chain_seeds ? o->chain_seeds[c] : stream_seed(o->seed, c));
            for (int i = 0; i < D; ++i)
                chains[c].q[i] = initial_q[c * D + i];
            chains[c].seed = chains[c].random();
            current[c] = chains[c].q;
            seeds[c] = chains[c].seed;
            ok[c] = model.admissible(current[c]);
            if (!ok[c])
                throw std::invalid_argument("initial particle posterior estimate is not finite");
        }
        evaluate(eng, model, w, initial_field, C, P, current, seeds, ok, current, ll);
        for (size_t c = 0; c < C; ++c) {
            chains[c].likelihood = ll[c];
            chains[c].density = ll[c] + model.log_prior(chains[c].q);
            if (!std::isfinite(chains[c].density))
                throw std::invalid_argument("initial particle posterior estimate is not finite");
        }
        for (size_t s = 0; s < S; ++s) {
            for (size_t c = 0; c < C; ++c) {
                Chain &ch = chains[c];
                proposed[c] = ch.q;
                for (int i = 0; i < D; ++i)
                    proposed[c][i] += o->proposal_scale * ch.normal();
                seeds[c] = ch.random();
                ok[c] = model.admissible(proposed[c]);
                current[c] = ch.q;
            }
            evaluate(eng, model, w, initial_field, C, P, proposed, seeds, ok, current, ll);
            for (size_t c = 0; c < C; ++c) {
                Chain &ch = chains[c];
                const double density = ll[c] + model.log_prior(proposed[c]);
                const double draw = ch.uniform();
                const bool accepted =
                    std::isfinite(density) &&
                    std::log(std::max(draw, std::numeric_limits<double>::min())) <
                        std::min(0.0, density - ch.density);
                if (accepted) {
                    ch.q = proposed[c];
                    ch.likelihood = ll[c];
                    ch.density = density;
                    ch.seed = seeds[c];
                }
                const size_t k = c * S + s;
                for (int i = 0; i < D; ++i)
                    samples_q[k * D + i] = ch.q[i];
                samples_ll[k] = ch.likelihood;
                if (samples_density)
                    samples_density[k] = ch.density;
                if (samples_accepted)
                    samples_accepted[k] = accepted ? 1 : 0;
                if (samples_seed)
                    samples_seed[k] = ch.seed;
            }
        }
    });
}
