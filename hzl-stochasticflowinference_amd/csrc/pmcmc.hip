// pmcmc.hip -- batched particle-marginal Metropolis-Hastings (host driver).
//
// Restates reference particle_metropolis (src/filter.cpp:166-203) for C chains
// at once: every chain is one particle filter over its own contiguous segment
// of the ensemble (hzl_pf_begin_filters), so one pass of the engine's batched
// RK4 and filter kernels evaluates all C proposed likelihoods.  Per chain c,
// with chain seed s_c = stream_seed(seed, c), exactly the reference's streams:
//   proposals  q' = q + scale * NormalStream(s_c)() per component,
//   likelihood seeds and acceptance draws from mt19937_64(stream_seed(s_c, 1)),
//   accept iff the proposed density is finite and
//   log(max(u, DBL_MIN)) < min(0, density' - density); on rejection the
//   current likelihood estimate and its seed are kept (pseudo-marginal).
// Parameters are q = (log nu, log kolmogorov amplitude) with a log-uniform
// prior (uniform in q inside [lower, upper], -inf outside); the reference's
// explicit-step guard dt (2 nu K^2 + alpha) > 2.5 gives -inf likelihood
// (src/filter.cpp:145-148).  Each likelihood evaluation restarts every
// particle from the same initial state at step 0.
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

// The reference NormalStream (src/solver.cpp:41-54).
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

struct Chain {
    Normals normal;
    std::mt19937_64 random;
    std::array<double, 2> q;
    double likelihood = NEG_INF, density = NEG_INF;
    uint64_t seed = 0;
    explicit Chain(uint64_t s) : normal(s), random(stream_seed(s, 1)) {}
    double uniform() {
        return (double)(random() >> 11) * 0x1.0p-53;
    }
};

struct Driver {
    EngineBase &eng;
    const hzl_pmcmc_options &o;
    double dt;
    std::vector<size_t> gstep, gbegin; // observation groups
    const int *x, *y;
    const double *value, *sd, *initial;
    size_t C, P;
    int K;

    double log_prior(const std::array<double, 2> &q) const {
        for (int i = 0; i < 2; ++i)
            if (!std::isfinite(q[i]) || q[i] < o.log_lower[i] || q[i] > o.log_upper[i])
                return NEG_INF;
        return 0.0;
    }
    bool stable(const std::array<double, 2> &q) const {
        const double nu = std::exp(q[0]);
        return dt * (2 * nu * (double)K * (double)K + o.friction) <= 2.5;
    }

    // Log-likelihood estimates of every chain at (q[c], seed[c]); chains with
    // active[c] false (outside the prior or unstable) run on `fallback[c]`
    // and report -inf.
    void evaluate(const std::vector<std::array<double, 2>> &q, const std::vector<uint64_t> &seeds,
                  const std::vector<char> &active,
                  const std::vector<std::array<double, 2>> &fallback, std::vector<double> &ll) {
        eng.fill_state(initial);
        eng.set_step_index(0);
        for (size_t c = 0; c < C; ++c) {
            const auto &qc = active[c] ? q[c] : fallback[c];
            for (size_t i = 0; i < P; ++i)
                eng.set_member_parameters(c * P + i, std::exp(qc[0]), o.friction, o.forcing_sigma,
                                          std::exp(qc[1]));
        }
        eng.pf_begin_filters(seeds.data(), C, gstep.size() + 1, true);
        size_t current = 0;
        std::vector<double> rec(4 * C);
        for (size_t g = 0; g < gstep.size(); ++g) {
            const size_t b = gbegin[g], e = gbegin[g + 1];
            eng.pf_assimilate(dt, gstep[g] - current, e - b, x + b, y + b, value + b, sd + b,
                              g + 1 == gstep.size(), rec.data(), nullptr);
            current = gstep[g];
        }
        ll.assign(C, NEG_INF);
        eng.pf_result(ll.data(), nullptr);
        for (size_t c = 0; c < C; ++c)
            if (!active[c])
                ll[c] = NEG_INF;
    }
};

} // namespace

} // namespace hzl_amd

using hzl_amd::engine_of;
using hzl_amd::guarded;

extern "C" int hzl_pmcmc_run(hzl_ensemble_t h, const hzl_pmcmc_options *o, double dt,
                             size_t nobs, const int *step, const int *x, const int *y,
                             const double *value, const double *sd, const double *initial_field,
                             const double *initial_q, double *samples_q, double *samples_ll,
                             double *samples_density, int *samples_accepted) {
    return guarded([&] {
        using namespace hzl_amd;
        if (!o || !step || !x || !y || !value || !sd || !initial_field || !initial_q ||
            !samples_q || !samples_ll)
            throw std::invalid_argument("null argument");
        EngineBase &eng = engine_of(h);
        const EngineViews v = eng.views();
        if (o->chains == 0 || (size_t)v.B % o->chains != 0 || (size_t)v.B / o->chains < 2)
            throw std::invalid_argument("members must split into chains of at least two particles");
        if (o->samples == 0 || !std::isfinite(o->proposal_scale) || o->proposal_scale <= 0)
            throw std::invalid_argument("invalid particle sampler configuration");
        for (int i = 0; i < 2; ++i)
            if (!std::isfinite(o->log_lower[i]) || !std::isfinite(o->log_upper[i]) ||
                o->log_lower[i] >= o->log_upper[i])
                throw std::invalid_argument("prior bounds must be finite and ordered");
        if (!std::isfinite(o->friction) || o->friction < 0 || !std::isfinite(o->forcing_sigma) ||
            o->forcing_sigma < 0)
            throw std::invalid_argument("friction and forcing sigma must be finite and nonnegative");
        if (nobs == 0)
            throw std::invalid_argument("observation window is empty");
        Driver d{eng, *o, dt, {}, {}, x, y, value, sd, initial_field, o->chains,
                 (size_t)v.B / o->chains, v.K};
        for (size_t i = 0; i < nobs; ++i) {
            if (step[i] <= 0 || (i && step[i] < step[i - 1]))
                throw std::invalid_argument("observations must be time ordered after step 0");
            if (i == 0 || step[i] != step[i - 1]) {
                d.gstep.push_back((size_t)step[i]);
                d.gbegin.push_back(i);
            }
        }
        d.gbegin.push_back(nobs);
        const size_t C = o->chains;
        std::vector<Chain> chains;
        chains.reserve(C);
        std::vector<std::array<double, 2>> q(C), prop(C);
        std::vector<uint64_t> seeds(C);
        std::vector<char> active(C);
        std::vector<double> ll;
        for (size_t c = 0; c < C; ++c) {
            chains.emplace_back(stream_seed(o->seed, c));
            chains[c].q = {initial_q[2 * c], initial_q[2 * c + 1]};
            chains[c].seed = chains[c].random();
            q[c] = chains[c].q;
            seeds[c] = chains[c].seed;
            active[c] = d.log_prior(q[c]) > NEG_INF && d.stable(q[c]);
            if (!active[c])
                throw std::invalid_argument("initial chain parameters lie outside the prior");
        }
        d.evaluate(q, seeds, active, q, ll);
        for (size_t c = 0; c < C; ++c) {
            chains[c].likelihood = ll[c];
            chains[c].density = ll[c] + d.log_prior(chains[c].q);
            if (!std::isfinite(chains[c].density))
                throw std::invalid_argument("initial particle posterior estimate is not finite");
        }
        for (size_t s = 0; s < o->samples; ++s) {
            for (size_t c = 0; c < C; ++c) {
                Chain &ch = chains[c];
                prop[c] = ch.q;
                for (auto &v2 : prop[c])
                    v2 += o->proposal_scale * ch.normal();
                seeds[c] = ch.random();
                active[c] = d.log_prior(prop[c]) > NEG_INF && d.stable(prop[c]);
                q[c] = ch.q;
            }
            d.evaluate(prop, seeds, active, q, ll);
            for (size_t c = 0; c < C; ++c) {
                Chain &ch = chains[c];
                const double density = ll[c] + d.log_prior(prop[c]);
                const double draw = ch.uniform();
                const bool accepted =
                    std::isfinite(density) &&
                    std::log(std::max(draw, std::numeric_limits<double>::min())) <
                        std::min(0.0, density - ch.density);
                if (accepted) {
                    ch.q = prop[c];
                    ch.likelihood = ll[c];
                    ch.density = density;
                    ch.seed = seeds[c];
                }
                const size_t k = c * o->samples + s;
                samples_q[2 * k] = ch.q[0];
                samples_q[2 * k + 1] = ch.q[1];
                samples_ll[k] = ch.likelihood;
                if (samples_density)
                    samples_density[k] = ch.density;
                if (samples_accepted)
                    samples_accepted[k] = accepted ? 1 : 0;
            }
        }
    });
}
