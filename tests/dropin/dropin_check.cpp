// dropin_check.cpp -- drives the C++ drop-in layer (libhzl_gpu.so) through the
// reference's own class and function signatures (TEST ONLY).  Each mode writes
// its results as raw little-endian doubles for tests/test_dropin.py.
//   dataset <in.csv> <out.csv>      Dataset::load then Dataset::save
//   fft <n> <out>                   Fourier2D forward of a fixed pattern, then inverse
//   initial <n> <name> <out>        initial_condition
//   solver <n> <steps> <out>        CudaSolver from multimode, deterministic steps
//   forced <n> <steps> <out>        CudaSolver::step(dt, F) with a fixed forcing
//   pf <csv> <out>                  bootstrap_filter(data, p, 8, 555)
//   enkf <csv> <out>                ensemble_kalman_filter(data, p, 8, 555)
//   pf_ou <csv> <out>               bootstrap_filter(data, p, 8, 555) with sigma 0.2,
//                                   tau 0.5 (the OU forcing process)
//   pmcmc <csv> <count> <out>       particle_metropolis(data, log(0.03, 0.1, 0.2, 0.5),
//                                   0.1, count, 8, 123): per sample q[4], log-likelihood,
//                                   log-posterior, likelihood seed (bits), accepted
//   ensemble_ou <out>               ensemble(16, p, 0.01, 20, 4, 2, 42) with sigma 0.2,
//                                   tau 0.5 on 2..3: energy, enstrophy, injection each
//   synthetic <sigma> <every> <steps> <noise> <out>
//                                   synthetic_observations(16, 0.01, steps, every, p, noise,
//                                   987) with nu 0.03, alpha 0.1, tau 0.5 on 2..3: rows
//                                   (step, x, y, value, sd), then the truth fields
//   spectrum <n> <out>              Ensemble (FP64) from multimode and taylor-green:
//                                   energy_spectra of both (n bins each)
//   likelihoods <csv> <out>         particle_log_likelihood at three seeds,
//                                   averaged_particle_log_likelihood over them (three
//                                   replicates) and over one, stochastic_log_prior
//   errors                          invalid arguments raise std::invalid_argument
//   adjoint <csv> <stride> <out>    likelihood_adjoint(data, q, stride) at three q:
//                                   value, gradient, checkpoint fields,
//                                   peak block fields, replayed steps each
//   resume_ou <in> <steps> <out>    Ensemble member from a checkpoint with OU forcing:
//                                   state, stream and forcing loaded, `steps` steps of
//                                   0.01, saved as a checkpoint
//   calibration <n> <out>           Ensemble (FP64, 2 members from multimode and
//                                   taylor-green): drift, linearized_drift along
//                                   each other's state, adjoint_drift, then
//                                   5 SSPRK3 steps of 1e-3
//   checkpoint <in> <out>           Checkpoint::load then Checkpoint::save
//   resume <in> <steps> <out>       CudaSolver from a checkpoint's vorticity, `steps`
//                                   steps of 1e-3 holding its forcing
//                                   (CudaSolver::step(dt, F)); writes the loaded
//                                   vorticity, forcing and the final state
#include "hzl_gpu.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

using namespace hzl;

static void write(const std::string &path, const std::vector<double> &v) {
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
}
static void append(std::vector<double> &v, const Field &f) {
    for (const auto &c : f) {
        v.push_back(c.real());
        v.push_back(c.imag());
    }
}

template <class F> static bool throws_invalid(F &&f) {
    try {
        f();
    } catch (const std::invalid_argument &) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

int main(int argc, char **argv) {
    if (argc < 2)
        return 2;
    const std::string mode = argv[1];
    try {
        if (mode == "dataset" && argc == 4) {
            Dataset::load(argv[2]).save(argv[3]);
        } else if (mode == "fft" && argc == 4) {
            const std::size_t n = std::stoul(argv[2]);
            Field f(n * n);
            for (std::size_t i = 0; i < f.size(); ++i)
                f[i] = Complex(std::sin(0.37 * (double)i) + 0.25, std::cos(1.3 * (double)i));
            std::vector<double> out;
            Fourier2D t(n);
            Field g = f;
            t.forward(g);
            append(out, g);
            t.inverse(g);
            append(out, g);
            write(argv[3], out);
        } else if (mode == "initial" && argc == 5) {
            std::vector<double> out;
            append(out, initial_condition(std::stoul(argv[2]), argv[3]));
            write(argv[4], out);
        } else if ((mode == "solver" || mode == "forced") && argc == 5) {
            const std::size_t n = std::stoul(argv[2]), steps = std::stoul(argv[3]);
            Parameters p;
            p.viscosity = 1e-2;
            p.friction = 0.1;
            p.forcing_sigma = 0;
            CudaSolver s(n, p, initial_condition(n, "multimode"));
            Field F(n * n);
            if (mode == "forced") {
                // 0.3 sin(2x) + 0.2 cos(3y) as coefficients
                F[2] = Complex(0, -0.15);
                F[n - 2] = Complex(0, 0.15);
                F[3 * n] = Complex(0.1, 0);
                F[(n - 3) * n] = Complex(0.1, 0);
            }
            for (std::size_t k = 0; k < steps; ++k) {
                if (mode == "forced")
                    s.step(1e-3, F);
                else
                    s.step(1e-3);
            }
            s.synchronize();
            std::vector<double> out;
            append(out, s.coefficients());
            write(argv[4], out);
        } else if (mode == "pf" && argc == 4) {
            const Dataset data = Dataset::load(argv[2]);
            Parameters p;
            p.viscosity = 0.03;
            p.friction = 0.1;
            p.forcing_sigma = 0;
            const FilterResult r = bootstrap_filter(data, p, 8, 555);
            std::vector<double> out{r.log_likelihood, (double)r.resampling_events};
            for (const auto &h : r.history) {
                out.insert(out.end(), {(double)h.step, h.log_likelihood_increment,
                                       h.effective_particles, h.maximum_weight});
                append(out, h.mean_vorticity);
            }
            write(argv[3], out);
        } else if (mode == "pf_ou" && argc == 4) {
            const Dataset data = Dataset::load(argv[2]);
            Parameters p;
            p.viscosity = 0.03;
            p.friction = 0.1;
            p.forcing_sigma = 0.2;
            p.forcing_tau = 0.5;
            const FilterResult r = bootstrap_filter(data, p, 8, 555);
            std::vector<double> out{r.log_likelihood, (double)r.resampling_events};
            for (const auto &h : r.history) {
                out.insert(out.end(), {(double)h.step, h.log_likelihood_increment,
                                       h.effective_particles, h.maximum_weight});
                append(out, h.mean_vorticity);
            }
            write(argv[3], out);
        } else if (mode == "pmcmc" && argc == 5) {
            const Dataset data = Dataset::load(argv[2]);
            const Vector q0{std::log(0.03), std::log(0.1), std::log(0.2), std::log(0.5)};
            const auto chain = particle_metropolis(data, q0, 0.1, std::stoul(argv[3]), 8, 123);
            std::vector<double> out;
            for (const auto &c : chain) {
                out.insert(out.end(), c.parameters.begin(), c.parameters.end());
                double bits;
                std::memcpy(&bits, &c.likelihood_seed, sizeof bits);
                out.insert(out.end(), {c.log_likelihood, c.log_posterior, bits,
                                       c.accepted ? 1.0 : 0.0});
            }
            write(argv[4], out);
        } else if (mode == "ensemble_ou" && argc == 3) {
            Parameters p;
            p.viscosity = 0.03;
            p.friction = 0.1;
            p.forcing_sigma = 0.2;
            p.forcing_tau = 0.5;
            std::vector<double> out;
            for (const auto &d : ensemble(16, p, 0.01, 20, 4, 2, 42))
                out.insert(out.end(), {d.energy, d.enstrophy, d.injection});
            write(argv[2], out);
        } else if (mode == "enkf" && argc == 4) {
            const Dataset data = Dataset::load(argv[2]);
            Parameters p;
            p.viscosity = 0.03;
            p.friction = 0.1;
            p.forcing_sigma = 0;
            const auto h = ensemble_kalman_filter(data, p, 8, 555);
            std::vector<double> out;
            for (const auto &r : h) {
                out.insert(out.end(),
                           {(double)r.step, r.forecast_sensor_rmse, r.posterior_spread});
                append(out, r.mean_vorticity);
            }
            write(argv[3], out);
        } else if (mode == "checkpoint" && argc == 4) {
            Checkpoint::load(argv[2]).save(argv[3]);
        } else if (mode == "resume" && argc == 5) {
            const Checkpoint c = Checkpoint::load(argv[2]);
            const std::size_t steps = std::stoul(argv[3]);
            CudaSolver s(c.n, c.parameters, c.vorticity);
            for (std::size_t k = 0; k < steps; ++k)
                s.step(1e-3, c.forcing);
            s.synchronize();
            std::vector<double> out;
            append(out, c.vorticity);
            append(out, c.forcing);
            append(out, s.coefficients());
            write(argv[4], out);
        } else if (mode == "resume_ou" && argc == 5) {
            Checkpoint c = Checkpoint::load(argv[2]);
            if (c.random_has_spare)
                throw std::runtime_error("stream holds a cached normal");
            const std::size_t steps = std::stoul(argv[3]);
            const double dt = 0.01;
            Ensemble e(c.n, 1, c.parameters, Precision::fp64, 0);
            e.set_state(0, c.vorticity);
            e.set_forcing_process(ForcingProcess::ornstein_uhlenbeck);
            e.set_forcing_state(0, c.random_engine, c.forcing);
            e.step(dt, steps);
            e.synchronize();
            c.vorticity = e.state(0);
            auto [engine, forcing] = e.forcing_state(0);
            c.random_engine = engine;
            c.forcing = forcing;
            for (std::size_t s = 0; s < steps; ++s)
                c.time += dt;
            c.save(argv[4]);
        } else if (mode == "calibration" && argc == 4) {
            const std::size_t n = std::stoul(argv[2]);
            Parameters p;
            p.viscosity = 1e-2;
            p.friction = 0.1;
            p.forcing_sigma = 0;
            Ensemble e(n, 2, p, Precision::fp64, 7);
            const Field a = initial_condition(n, "multimode"), b = initial_condition(n, "taylor-green");
            e.set_state(0, a);
            e.set_state(1, b);
            std::vector<double> out;
            for (const auto &f : e.drift())
                append(out, f);
            for (const auto &f : e.linearized_drift({b, a}))
                append(out, f);
            for (const auto &f : e.adjoint_drift({b, a}))
                append(out, f);
            e.step(1e-3, 5, TimeIntegrator::ssprk3);
            e.synchronize();
            append(out, e.state(0));
            append(out, e.state(1));
            write(argv[3], out);
        } else if (mode == "adjoint" && argc == 5) {
            const Dataset data = Dataset::load(argv[2]);
            const std::size_t stride = std::stoul(argv[3]);
            const std::vector<Vector> qs{{std::log(0.03), std::log(0.1)},
                                         {std::log(0.02), std::log(0.15)},
                                         {-13.0, -1.0}};
            std::vector<double> out;
            for (const auto &q : qs) {
                const AdjointEvaluation r = likelihood_adjoint(data, q, stride);
                out.insert(out.end(), {r.objective.value, r.objective.gradient[0],
                                       r.objective.gradient[1], (double)r.checkpoint_fields,
                                       (double)r.peak_block_fields, (double)r.replayed_steps});
            }
            write(argv[4], out);
        } else if (mode == "synthetic" && argc == 7) {
            Parameters p;
            p.viscosity = 0.03;
            p.friction = 0.1;
            p.forcing_sigma = std::stod(argv[2]);
            p.forcing_tau = 0.5;
            std::vector<Field> truth;
            const Dataset d = synthetic_observations(16, 0.01, std::stoul(argv[4]),
                                                     std::stoul(argv[3]), p, std::stod(argv[5]),
                                                     987, &truth);
            std::vector<double> out;
            for (const auto &o : d.observations)
                out.insert(out.end(), {(double)o.step, (double)o.x, (double)o.y, o.value,
                                       o.standard_deviation});
            for (const auto &f : truth)
                append(out, f);
            write(argv[6], out);
        } else if (mode == "spectrum" && argc == 4) {
            const std::size_t n = std::stoul(argv[2]);
            Parameters p;
            p.forcing_sigma = 0;
            Ensemble e(n, 2, p, Precision::fp64, 7);
            e.set_state(0, initial_condition(n, "multimode"));
            e.set_state(1, initial_condition(n, "taylor-green"));
            std::vector<double> out;
            for (const auto &bins : e.energy_spectra())
                out.insert(out.end(), bins.begin(), bins.end());
            write(argv[3], out);
        } else if (mode == "likelihoods" && argc == 4) {
            const Dataset data = Dataset::load(argv[2]);
            const Vector q{std::log(0.03), std::log(0.1), std::log(0.2), std::log(0.5)};
            std::vector<double> out;
            for (std::uint64_t i = 0; i < 3; ++i)
                out.push_back(particle_log_likelihood(data, q, 8, stream_seed(555, i)));
            out.push_back(averaged_particle_log_likelihood(data, q, 8, 3, 555));
            out.push_back(averaged_particle_log_likelihood(data, q, 8, 1, 555));
            out.push_back(stochastic_log_prior(q));
            out.push_back(particle_log_likelihood(data, {0.5, -1, -1, -1}, 8, 1)); // outside
            write(argv[3], out);
        } else if (mode == "errors") {
            Parameters p;
            const bool ok =
                throws_invalid([&] { CudaSolver(16, p, Field(15)); }) &&
                throws_invalid([&] { CudaSolver(16, p, initial_condition(16, "zero")).step(1e-3); }) &&
                throws_invalid([&] { Parameters q; q.forcing_max = 6; q.validate(16); }) &&
                throws_invalid([&] { (void)initial_condition(16, "vortex"); }) &&
                throws_invalid([&] {
                    Parameters q = p;
                    q.forcing_sigma = 0;
                    Field F(16 * 16);
                    F[1] = Complex(1, 0); // no conjugate partner
                    CudaSolver(16, q, initial_condition(16, "zero")).step(1e-3, F);
                });
            std::cout << (ok ? "ok" : "fail") << "\n";
            return ok ? 0 : 1;
        } else {
            return 2;
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
