// ref_shim.cpp -- C entry points over the UNMODIFIED reference CPU library.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) together
// with the reference's own sources where they lie under /root/reference/src
// (fft.cpp, solver.cpp, inference.cpp, adjoint.cpp, filter.cpp) into
// oracle/_ref/libhzlref.so.  Nothing here is copied from the reference: this
// file only calls its public C++ API (include/hzl/*.hpp) so that the Python
// tests and the golden-vector script can drive it through ctypes.
//
// The engine's physics (Kolmogorov forcing, shell white noise, Gaussian random
// field) is injected through the reference's public API: a frozen OU process
// (sigma = 0, tau = 1e300, so transition_ou leaves the forcing unchanged) and
// Solver::update_latent_state(coefficients(), F_step) before every step
// (include/hzl/solver.hpp:84, src/solver.cpp:403-425).
#include "hzl/filter.hpp"
#include "hzl/inference.hpp"
#include "hzl/solver.hpp"
#include <array>
#include <cstring>
#include <exception>
#include <string>

namespace {
thread_local std::string last_error;

template <class F> int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        last_error = e.what();
        return 1;
    } catch (...) {
        last_error = "unknown exception";
        return 1;
    }
}

hzl::Field to_field(const double *d, std::size_t count) {
    hzl::Field f(count);
    for (std::size_t i = 0; i < count; ++i)
        f[i] = {d[2 * i], d[2 * i + 1]};
    return f;
}

void from_field(const hzl::Field &f, double *d) {
    for (std::size_t i = 0; i < f.size(); ++i) {
        d[2 * i] = f[i].real();
        d[2 * i + 1] = f[i].imag();
    }
}

hzl::Parameters frozen(double nu, double alpha, int kmin, int kmax) {
    hzl::Parameters p;
    p.viscosity = nu;
    p.friction = alpha;
    p.forcing_sigma = 0;
    p.forcing_tau = 1e300;
    p.forcing_min = kmin;
    p.forcing_max = kmax;
    return p;
}
} // namespace

extern "C" {

const char *ref_last_error() {
    return last_error.c_str();
}

int ref_fft2d(std::size_t n, double *data, int inverse) {
    return guarded([&] {
        hzl::Fourier2D fft(n);
        auto f = to_field(data, n * n);
        if (inverse)
            fft.inverse(f);
        else
            fft.forward(f);
        from_field(f, data);
    });
}

int ref_wave_number(std::size_t index, std::size_t n) {
    return hzl::wave_number(index, n);
}

std::uint64_t ref_stream_seed(std::uint64_t base, std::uint64_t index) {
    return hzl::stream_seed(base, index);
}

// `count` normals of NormalStream(seed) (src/solver.cpp:41-54).
void ref_normals(std::uint64_t seed, std::size_t count, double *out) {
    hzl::NormalStream s(seed);
    for (std::size_t i = 0; i < count; ++i)
        out[i] = s();
}

// Run `steps` reference RK4 steps from a spectral state.  The forcing of step k
// is forcing + k * forcing_stride complex entries (stride 0 holds one field).
// Forcing is applied only on shell pairs kmin..kmax, as update_latent_state does.
int ref_run(std::size_t n, double nu, double alpha, double dt, std::size_t steps, int kmin,
            int kmax, const double *initial, const double *forcing, std::size_t forcing_stride,
            double *out) {
    return guarded([&] {
        hzl::Solver s(n, frozen(nu, alpha, kmin, kmax), 0);
        const hzl::Field zero(n * n);
        s.update_latent_state(to_field(initial, n * n),
                              forcing ? to_field(forcing, n * n) : zero);
        for (std::size_t k = 0; k < steps; ++k) {
            if (forcing && (k == 0 || forcing_stride != 0))
                s.update_latent_state(s.coefficients(),
                                      to_field(forcing + 2 * forcing_stride * k, n * n));
            s.step(dt);
        }
        from_field(s.coefficients(), out);
    });
}

// Solver(n, p, seed) with OU shell forcing (sigma, tau on kmin..kmax): the
// given spectral state (the constructor's forcing kept), `steps` RK4 steps;
// final state and forcing out.
int ref_run_ou(std::size_t n, double nu, double alpha, double sigma, double tau, int kmin,
               int kmax, std::uint64_t seed, double dt, std::size_t steps, const double *initial,
               double *out_w, double *out_F) {
    return guarded([&] {
        hzl::Parameters p;
        p.viscosity = nu;
        p.friction = alpha;
        p.forcing_sigma = sigma;
        p.forcing_tau = tau;
        p.forcing_min = kmin;
        p.forcing_max = kmax;
        hzl::Solver s(n, p, seed);
        const hzl::Field start_forcing = s.forcing(); // copy: the call rewrites forcing_
        s.update_latent_state(to_field(initial, n * n), start_forcing);
        for (std::size_t k = 0; k < steps; ++k)
            s.step(dt);
        from_field(s.coefficients(), out_w);
        from_field(s.forcing(), out_F);
    });
}

// Solver(n, p, seed) with OU forcing from multimode: `before` steps, save a
// checkpoint, `after` more steps, save another (Solver::save, src/solver.cpp:434-454).
int ref_checkpoint_ou(std::size_t n, double nu, double alpha, double sigma, double tau,
                      int kmin, int kmax, std::uint64_t seed, double dt, std::size_t before,
                      std::size_t after, const char *path_before, const char *path_after) {
    return guarded([&] {
        hzl::Parameters p;
        p.viscosity = nu;
        p.friction = alpha;
        p.forcing_sigma = sigma;
        p.forcing_tau = tau;
        p.forcing_min = kmin;
        p.forcing_max = kmax;
        hzl::Solver s(n, p, seed);
        s.initialize("multimode");
        for (std::size_t k = 0; k < before; ++k)
            s.step(dt);
        s.save(path_before);
        for (std::size_t k = 0; k < after; ++k)
            s.step(dt);
        s.save(path_after);
    });
}

// `steps` reference SSPRK(3,3) steps (Solver::step(dt, TimeIntegrator::ssprk3))
// from a spectral state with one forcing field held on the shell pairs.
int ref_run_ssprk3(std::size_t n, double nu, double alpha, double dt, std::size_t steps,
                   int kmin, int kmax, const double *initial, const double *forcing,
                   double *out) {
    return guarded([&] {
        hzl::Solver s(n, frozen(nu, alpha, kmin, kmax), 0);
        s.update_latent_state(to_field(initial, n * n),
                              forcing ? to_field(forcing, n * n) : hzl::Field(n * n));
        for (std::size_t k = 0; k < steps; ++k)
            s.step(dt, hzl::TimeIntegrator::ssprk3);
        from_field(s.coefficients(), out);
    });
}

// Solver::linearized_drift and Solver::adjoint_drift of a spectral state.
int ref_linearized_drift(std::size_t n, double nu, double alpha, const double *state,
                         const double *direction, double *out) {
    return guarded([&] {
        hzl::Solver s(n, frozen(nu, alpha, 1, 1), 0);
        from_field(s.linearized_drift(to_field(state, n * n), to_field(direction, n * n)), out);
    });
}

int ref_adjoint_drift(std::size_t n, double nu, double alpha, const double *state,
                      const double *cotangent, double *out) {
    return guarded([&] {
        hzl::Solver s(n, frozen(nu, alpha, 1, 1), 0);
        from_field(s.adjoint_drift(to_field(state, n * n), to_field(cotangent, n * n)), out);
    });
}

// Physical-space initial condition through Solver::initialize_physical (one
// forward FFT and projection, src/solver.cpp:162-173); returns the spectrum.
int ref_initialize_physical(std::size_t n, const double *physical_real, double *out) {
    return guarded([&] {
        hzl::Solver s(n, frozen(0, 0, 1, 1), 0);
        hzl::Field f(n * n);
        for (std::size_t i = 0; i < n * n; ++i)
            f[i] = physical_real[i];
        s.initialize_physical(f);
        from_field(s.coefficients(), out);
    });
}

// Named initial condition ("taylor-green", "multimode", "zero").
int ref_initialize(std::size_t n, const char *name, double *out) {
    return guarded([&] {
        hzl::Solver s(n, frozen(0, 0, 1, 1), 0);
        s.initialize(name);
        from_field(s.coefficients(), out);
    });
}

// Drift (RHS without forcing) of a spectral state (src/solver.cpp:255-263).
int ref_drift(std::size_t n, double nu, double alpha, const double *state, double *out) {
    return guarded([&] {
        hzl::Solver s(n, frozen(nu, alpha, 1, 1), 0);
        from_field(s.drift(to_field(state, n * n)), out);
    });
}

// out = energy, enstrophy, dissipation, injection, cfl_rate.
int ref_diagnostics(std::size_t n, double nu, double alpha, int kmin, int kmax,
                    const double *state, const double *forcing, double *out) {
    return guarded([&] {
        hzl::Solver s(n, frozen(nu, alpha, kmin, kmax), 0);
        s.update_latent_state(to_field(state, n * n),
                              forcing ? to_field(forcing, n * n) : hzl::Field(n * n));
        const auto d = s.diagnostics();
        out[0] = d.energy;
        out[1] = d.enstrophy;
        out[2] = d.dissipation;
        out[3] = d.injection;
        out[4] = d.cfl_rate;
    });
}

// A reference Solver with OU shell forcing (seed `seed`) started from a named
// initial condition, advanced `steps` RK4 steps of dt, written by Solver::save.
int ref_checkpoint_write(std::size_t n, double nu, double alpha, double sigma, double tau,
                         int kmin, int kmax, std::uint64_t seed, const char *initial, double dt,
                         std::size_t steps, const char *path) {
    return guarded([&] {
        hzl::Parameters p;
        p.viscosity = nu;
        p.friction = alpha;
        p.forcing_sigma = sigma;
        p.forcing_tau = tau;
        p.forcing_min = kmin;
        p.forcing_max = kmax;
        hzl::Solver s(n, p, seed);
        s.initialize(initial);
        for (std::size_t k = 0; k < steps; ++k)
            s.step(dt);
        s.save(path);
    });
}

// Solver::load of a checkpoint, then its time and diagnostics:
// out = time, energy, enstrophy, dissipation, injection, cfl_rate.
int ref_checkpoint_diagnostics(const char *path, double *out) {
    return guarded([&] {
        hzl::Solver s = hzl::Solver::load(path);
        const auto d = s.diagnostics();
        out[0] = d.time;
        out[1] = d.energy;
        out[2] = d.enstrophy;
        out[3] = d.dissipation;
        out[4] = d.injection;
        out[5] = d.cfl_rate;
    });
}

// Solver::load of a checkpoint: its vorticity and forcing coefficients
// (n * n each, interleaved complex).
int ref_checkpoint_fields(const char *path, double *vorticity, double *forcing) {
    return guarded([&] {
        const hzl::Solver s = hzl::Solver::load(path);
        from_field(s.coefficients(), vorticity);
        from_field(s.forcing(), forcing);
    });
}

int ref_energy_spectrum(std::size_t n, const double *state, double *bins) {
    return guarded([&] {
        hzl::Solver s(n, frozen(0, 0, 1, 1), 0);
        s.update_latent_state(to_field(state, n * n), hzl::Field(n * n));
        const auto b = s.energy_spectrum();
        std::memcpy(bins, b.data(), sizeof(double) * b.size());
    });
}

// The reference's own member-parallel CPU ensemble (src/inference.cpp:555-590):
// zero initial condition, OU shell forcing, one std::jthread per worker.
int ref_ensemble(std::size_t n, double nu, double alpha, double sigma, double tau, int kmin,
                 int kmax, double dt, std::size_t steps, std::size_t count, std::size_t workers,
                 std::uint64_t seed, double *energies) {
    return guarded([&] {
        hzl::Parameters p;
        p.viscosity = nu;
        p.friction = alpha;
        p.forcing_sigma = sigma;
        p.forcing_tau = tau;
        p.forcing_min = kmin;
        p.forcing_max = kmax;
        const auto d = hzl::ensemble(n, p, dt, steps, count, workers, seed);
        for (std::size_t i = 0; i < count; ++i)
            energies[i] = d[i].energy;
    });
}

// ---------------------------------------------------------------- filters --
struct RefDataset {
    hzl::Dataset data;
};

// synthetic_observations (src/inference.cpp:256-288); returns an opaque dataset.
void *ref_synthetic(std::size_t n, double dt, std::size_t steps, std::size_t every, double nu,
                    double alpha, double sigma, double tau, int kmin, int kmax, double noise,
                    std::uint64_t seed) {
    RefDataset *d = nullptr;
    const int rc = guarded([&] {
        hzl::Parameters p;
        p.viscosity = nu;
        p.friction = alpha;
        p.forcing_sigma = sigma;
        p.forcing_tau = tau;
        p.forcing_min = kmin;
        p.forcing_max = kmax;
        d = new RefDataset{hzl::synthetic_observations(n, dt, steps, every, p, noise, seed)};
    });
    return rc ? nullptr : d;
}

void ref_dataset_free(void *d) {
    delete static_cast<RefDataset *>(d);
}

std::size_t ref_dataset_size(void *d) {
    return static_cast<RefDataset *>(d)->data.observations.size();
}

// Observation rows as (step, x, y, value, sd).
void ref_dataset_rows(void *d, double *rows) {
    const auto &obs = static_cast<RefDataset *>(d)->data.observations;
    for (std::size_t i = 0; i < obs.size(); ++i) {
        rows[5 * i + 0] = static_cast<double>(obs[i].step);
        rows[5 * i + 1] = static_cast<double>(obs[i].x);
        rows[5 * i + 2] = static_cast<double>(obs[i].y);
        rows[5 * i + 3] = obs[i].value;
        rows[5 * i + 4] = obs[i].standard_deviation;
    }
}

// bootstrap_filter (src/filter.cpp:38-126).  summary = log_likelihood,
// resampling_events, number of records; per-record (increment, ESS, max weight)
// into records (3 per record, capacity max_records); mean physical vorticity
// real parts into means (n*n per record) when non-null.
int ref_bootstrap_filter(void *d, double nu, double alpha, double sigma, double tau,
                         std::size_t particles, std::uint64_t seed, double *summary,
                         double *records, std::size_t max_records, double *means) {
    return guarded([&] {
        const auto &data = static_cast<RefDataset *>(d)->data;
        hzl::Parameters p;
        p.viscosity = nu;
        p.friction = alpha;
        p.forcing_sigma = sigma;
        p.forcing_tau = tau;
        const auto r = hzl::bootstrap_filter(data, p, particles, seed, means != nullptr);
        summary[0] = r.log_likelihood;
        summary[1] = static_cast<double>(r.resampling_events);
        summary[2] = static_cast<double>(r.history.size());
        for (std::size_t i = 0; i < r.history.size() && i < max_records; ++i) {
            records[3 * i + 0] = r.history[i].log_likelihood_increment;
            records[3 * i + 1] = r.history[i].effective_particles;
            records[3 * i + 2] = r.history[i].maximum_weight;
            if (means) {
                const auto &f = r.history[i].mean_vorticity;
                for (std::size_t j = 0; j < f.size(); ++j)
                    means[i * f.size() + j] = f[j].real();
            }
        }
    });
}

// particle_metropolis (src/filter.cpp:166-203) from q0 (4 log parameters):
// per sample the 4 parameters, log-likelihood, log-posterior, likelihood seed
// (as a double bit pattern) and the accepted flag, 8 doubles per sample.
int ref_particle_metropolis(void *d, const double *q0, double scale, std::size_t count,
                            std::size_t particles, std::uint64_t seed, double *out) {
    return guarded([&] {
        const auto &data = static_cast<RefDataset *>(d)->data;
        const auto chain = hzl::particle_metropolis(data, {q0[0], q0[1], q0[2], q0[3]}, scale,
                                                    count, particles, seed);
        for (std::size_t i = 0; i < chain.size(); ++i) {
            for (int j = 0; j < 4; ++j)
                out[8 * i + j] = chain[i].parameters[j];
            out[8 * i + 4] = chain[i].log_likelihood;
            out[8 * i + 5] = chain[i].log_posterior;
            std::memcpy(&out[8 * i + 6], &chain[i].likelihood_seed, sizeof(double));
            out[8 * i + 7] = chain[i].accepted ? 1.0 : 0.0;
        }
    });
}

// Dataset::save (src/inference.cpp:214-225): the reference's own CSV writer.
int ref_dataset_save(void *d, const char *path) {
    return guarded([&] { static_cast<RefDataset *>(d)->data.save(path); });
}

// Deterministic negative log-likelihood over (log nu, log alpha) (src/inference.cpp:321-339).
int ref_negative_log_likelihood(void *d, double log_nu, double log_alpha, double *out) {
    return guarded([&] {
        *out = hzl::negative_log_likelihood(static_cast<RefDataset *>(d)->data,
                                            {log_nu, log_alpha});
    });
}

// likelihood_gradient (src/inference.cpp:340-382): out = value, gradient[0..1].
int ref_likelihood_gradient(void *d, double log_nu, double log_alpha, double *out) {
    return guarded([&] {
        const auto r = hzl::likelihood_gradient(static_cast<RefDataset *>(d)->data,
                                                {log_nu, log_alpha});
        out[0] = r.value;
        out[1] = r.gradient[0];
        out[2] = r.gradient[1];
    });
}

// `steps` Solver::step_tangent steps from a spectral state and two
// sensitivity fields (no forcing); writes the state and the sensitivities.
int ref_step_tangent(std::size_t n, double nu, double alpha, double dt, std::size_t steps,
                     const double *state, const double *tangents, double *state_out,
                     double *tangents_out) {
    return guarded([&] {
        hzl::Solver s(n, frozen(nu, alpha, 1, 1), 0);
        s.update_latent_state(to_field(state, n * n), hzl::Field(n * n));
        std::array<hzl::Field, 2> t{to_field(tangents, n * n), to_field(tangents + 2 * n * n, n * n)};
        for (std::size_t k = 0; k < steps; ++k)
            s.step_tangent(dt, t);
        from_field(s.coefficients(), state_out);
        from_field(t[0], tangents_out);
        from_field(t[1], tangents_out + 2 * n * n);
    });
}

// likelihood_adjoint (src/adjoint.cpp:43-161): out = value, gradient[0..1],
// checkpoint_fields, peak_block_fields, replayed_steps.
int ref_likelihood_adjoint(void *d, double log_nu, double log_alpha, std::size_t stride,
                           double *out) {
    return guarded([&] {
        const auto r = hzl::likelihood_adjoint(static_cast<RefDataset *>(d)->data,
                                               {log_nu, log_alpha}, stride);
        out[0] = r.objective.value;
        out[1] = r.objective.gradient[0];
        out[2] = r.objective.gradient[1];
        out[3] = (double)r.checkpoint_fields;
        out[4] = (double)r.peak_block_fields;
        out[5] = (double)r.replayed_steps;
    });
}

// ensemble_kalman_filter (src/filter.cpp:204-356): per record (step,
// forecast_sensor_rmse, posterior_spread) and the mean physical field.
int ref_enkf(void *d, double nu, double alpha, double sigma, double tau, std::size_t members,
             std::uint64_t seed, double inflation, double radius, double *records,
             std::size_t max_records, double *means, std::size_t *count) {
    return guarded([&] {
        const auto &data = static_cast<RefDataset *>(d)->data;
        hzl::Parameters p;
        p.viscosity = nu;
        p.friction = alpha;
        p.forcing_sigma = sigma;
        p.forcing_tau = tau;
        hzl::EnKFOptions o;
        o.inflation = inflation;
        o.localization_radius = radius;
        const auto h = hzl::ensemble_kalman_filter(data, p, members, seed, o);
        *count = h.size();
        for (std::size_t i = 0; i < h.size() && i < max_records; ++i) {
            records[3 * i + 0] = static_cast<double>(h[i].step);
            records[3 * i + 1] = h[i].forecast_sensor_rmse;
            records[3 * i + 2] = h[i].posterior_spread;
            if (means)
                for (std::size_t j = 0; j < h[i].mean_vorticity.size(); ++j)
                    means[i * h[i].mean_vorticity.size() + j] = h[i].mean_vorticity[j].real();
        }
    });
}
}
