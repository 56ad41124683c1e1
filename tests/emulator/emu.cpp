// emu.cpp -- CPU lane emulator for the engine's kernel bodies (TEST ONLY).
//
// Runs the unmodified __host__ __device__ kernel bodies of
// hzl-stochasticflowinference_amd/csrc/kernels.hpp on the host: one std::thread
// per lane of a workgroup, std::barrier for __syncthreads, a host vector for
// LDS, and the engine's launch sequence (kernels.hpp header comment) replayed
// over host arrays.  Used by the CPU test suite to check indexing, layouts and
// numerics against the oracle without a GPU.  Never linked into the product.
#include "tables.hpp"

#include <barrier>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

using namespace hzl_amd;

namespace {

struct CpuCtx {
    int lane;
    std::barrier<> *bar;
    int tid() const {
        return lane;
    }
    void sync() const {
        bar->arrive_and_wait();
    }
    // lanes are independent OS threads here, so wave-local order needs the
    // full barrier
    void wsync() const {
        bar->arrive_and_wait();
    }
};

// Launch `blocks` workgroups of `threads` lanes; body(ctx, block, lds).
template <class Body> void launch(long blocks, int threads, size_t lds_bytes, Body body) {
    std::barrier<> bar(threads);
    std::vector<unsigned char> lds(lds_bytes + 64);
    std::vector<std::thread> pool;
    pool.reserve((size_t)threads);
    for (int lane = 0; lane < threads; ++lane)
        pool.emplace_back([&, lane] {
            CpuCtx ctx{lane, &bar};
            for (long b = 0; b < blocks; ++b) {
                body(ctx, b, lds.data());
                bar.arrive_and_wait();
            }
        });
    for (auto &t : pool)
        t.join();
}

template <int N, typename R> struct Sim {
    using G = Geometry<N>;
    using FFT = GroupFFT<N, G::E, R>;
    int B;
    std::vector<cx<R>> w0, w1, acc, mix, nl, ext, tw, noise;
    int npairs = 0;
    std::vector<double> cfl;
    std::vector<double> nu, alpha, sigma, amp;
    std::vector<uint64_t> keys;
    std::vector<int> pairs;
    int kmin, kmax, kf;
    uint64_t step = 0;

    Sim(int b, double nu0, double alpha0, double sigma0, double amp0, int kf_, int kmin_,
        int kmax_, uint64_t seed)
        : B(b), w0((size_t)b * G::SPEC), w1((size_t)b * G::SPEC), acc((size_t)b * G::SPEC),
          mix((size_t)b * 4 * G::MIXF), nl((size_t)b * G::MIXF), ext((size_t)b * G::SPEC),
          tw(twiddles<R>(N)), cfl((size_t)b * (N / 2), 0.0), nu(b, nu0), alpha(b, alpha0),
          sigma(b, sigma0), amp(b, amp0), keys(b), pairs(pair_table(kmin_, kmax_)), kmin(kmin_),
          kmax(kmax_), kf(kf_) {
        for (int m = 0; m < b; ++m)
            keys[m] = stream_seed(seed, (uint64_t)m);
        pair_table(kmin_, kmax_, &npairs);
        noise.assign((size_t)b * npairs, cx<R>{0, 0});
    }

    ForcingModel<R> model() {
        return ForcingModel<R>{noise.data(), npairs, amp.data(), pairs.data(), kmin, kmax, kf};
    }

    void make_noise(double dt) {
        NoiseArgs<R> a{B, npairs, sigma.data(), keys.data(), dt, step, noise.data()};
        for (long i = 0; i < (long)B * npairs; ++i)
            noise_body<R>(a, i);
    }

    template <int MODE> void col_mode(const StageArgs<R> &a) {
        const long blocks = ((long)B * G::KP + G::COL_G - 1) / G::COL_G;
        launch(blocks, G::COL_THREADS, sizeof(cx<R>) * G::COL_G * FFT::LDS_ELEMS,
               [&](const CpuCtx &ctx, long b, unsigned char *lds) {
                   col_body<N, MODE, R>(ctx, a, reinterpret_cast<cx<R> *>(lds), b);
               });
    }

    void col(int mode, double dt, const cx<R> *e, long long es) {
        StageArgs<R> a;
        a.B = B;
        a.w0 = w0.data();
        a.w1 = w1.data();
        a.acc = acc.data();
        a.nl = nl.data();
        a.mix = mix.data();
        a.tw = tw.data();
        a.nu = nu.data();
        a.alpha = alpha.data();
        a.forcing = model();
        a.ext = e;
        a.ext_stride = es;
        a.h_half = (R)(dt / 2);
        a.h_full = (R)dt;
        a.h_sixth = (R)(dt / 6);
        switch (mode) {
        case 0:
            return col_mode<0>(a);
        case 1:
            return col_mode<1>(a);
        case 2:
            return col_mode<2>(a);
        case 3:
            return col_mode<3>(a);
        default:
            return col_mode<4>(a);
        }
    }

    void row(bool with_cfl) {
        RowArgs<R> a{B, mix.data(), nl.data(), tw.data(), with_cfl ? cfl.data() : nullptr};
        const long blocks = ((long)B * (N / 2) + G::ROW_PPW - 1) / G::ROW_PPW;
        launch(blocks, G::ROW_THREADS, sizeof(cx<R>) * RowShape<N>::template lds_elems<R>(),
               [&](const CpuCtx &ctx, long b, unsigned char *lds) {
                   if (with_cfl)
                       row_body<N, R, true>(ctx, a, reinterpret_cast<cx<R> *>(lds), b);
                   else
                       row_body<N, R, false>(ctx, a, reinterpret_cast<cx<R> *>(lds), b);
               });
    }

    void grf() {
        GrfArgs<R> a{G::K, w0.data(), keys.data()};
        launch(B, 64, sizeof(double) * 64, [&](const CpuCtx &ctx, long b, unsigned char *lds) {
            grf_body<R>(ctx, a, reinterpret_cast<double *>(lds), 64, b);
        });
    }

    void diag(double dt, double *out) {
        (void)dt;
        DiagArgs<R> a{G::K, w0.data(), nu.data(), model(), step > 0 ? 1 : 0, cfl.data(), N / 2,
                      N, out};
        launch(B, 64, sizeof(double) * 3 * 64, [&](const CpuCtx &ctx, long b, unsigned char *lds) {
            diag_body<R>(ctx, a, reinterpret_cast<double *>(lds), 64, b);
        });
    }

    void steps(double dt, int count, const cx<R> *e, long long es, bool &mix_valid) {
        for (int s = 0; s < count; ++s) {
            if (!mix_valid) {
                col(0, dt, e, es);
                mix_valid = true;
            }
            if (!e)
                make_noise(dt);
            for (int stage = 1; stage <= 4; ++stage) {
                row(stage == 1);
                col(stage, dt, e, es);
            }
            ++step;
        }
    }
};

template <int N, typename R>
int run(int B, double nu, double alpha, double sigma, double amp, int kf, int kmin, int kmax,
        uint64_t seed, double dt, int steps, int grf, const double *init, const double *forcing,
        int forcing_shared, double *out, double *diag) {
    Sim<N, R> s(B, nu, alpha, sigma, amp, kf, kmin, kmax, seed);
    const size_t nn = (size_t)N * N;
    if (grf)
        s.grf();
    else if (init)
        for (int m = 0; m < B; ++m)
            full_to_half<R>(N, init + 2 * nn * m, s.w0.data() + (size_t)m * Geometry<N>::SPEC);
    const cx<R> *e = nullptr;
    long long es = 0;
    if (forcing) {
        const int count = forcing_shared ? 1 : B;
        for (int m = 0; m < count; ++m)
            full_to_half<R>(N, forcing + 2 * nn * m, s.ext.data() + (size_t)m * Geometry<N>::SPEC);
        e = s.ext.data();
        es = forcing_shared ? 0 : (long long)Geometry<N>::SPEC;
    }
    bool valid = false;
    s.steps(dt, steps, e, es, valid);
    for (int m = 0; m < B; ++m)
        half_to_full<R>(N, s.w0.data() + (size_t)m * Geometry<N>::SPEC, out + 2 * nn * m);
    if (diag)
        s.diag(dt, diag);
    return 0;
}

template <typename R>
int run_any(int n, int B, double nu, double alpha, double sigma, double amp, int kf, int kmin,
            int kmax, uint64_t seed, double dt, int steps, int grf, const double *init,
            const double *forcing, int forcing_shared, double *out, double *diag) {
#define HZL_EMU_CASE(NN)                                                                       \
    case NN:                                                                                   \
        return run<NN, R>(B, nu, alpha, sigma, amp, kf, kmin, kmax, seed, dt, steps, grf, init, \
                          forcing, forcing_shared, out, diag);
    switch (n) {
        HZL_EMU_CASE(16)
        HZL_EMU_CASE(32)
        HZL_EMU_CASE(64)
        HZL_EMU_CASE(128)
        HZL_EMU_CASE(256)
        HZL_EMU_CASE(512)
    }
#undef HZL_EMU_CASE
    return -1;
}

// Batch of 1-D transforms through GroupFFT: lines[L][N] in natural order.
template <int N, typename R> int fft_lines(int lines, int inverse, double *data) {
    using F = GroupFFT<N, Shape<N>::E, R>;
    constexpr int E = Shape<N>::E, T = N / E;
    const auto tw = twiddles<R>(N);
    std::vector<cx<R>> buf((size_t)lines * N);
    for (size_t i = 0; i < buf.size(); ++i)
        buf[i] = {(R)data[2 * i], (R)data[2 * i + 1]};
    const int groups = 64 / T;
    const long blocks = (lines + groups - 1) / groups;
    launch(blocks, 64, sizeof(cx<R>) * groups * F::LDS_ELEMS,
           [&](const CpuCtx &ctx, long b, unsigned char *lds) {
               const int g = ctx.tid() / T, t = ctx.tid() % T;
               const long line = b * groups + g;
               const bool live = line < lines;
               cx<R> *L = reinterpret_cast<cx<R> *>(lds) + g * F::LDS_ELEMS;
               cx<R> v[E];
               cx<R> *x = buf.data() + (live ? line : 0) * N;
               if (!inverse) {
                   for (int j = 0; j < E; ++j)
                       v[j] = live ? x[t + T * j] : cx<R>{0, 0};
                   F::forward(ctx, v, L, t, tw.data());
                   ctx.sync();
                   if (live)
                       for (int i = 0; i < E; ++i)
                           x[F::index(t, i)] = v[i];
               } else {
                   for (int i = 0; i < E; ++i)
                       v[i] = live ? x[F::index(t, i)] : cx<R>{0, 0};
                   F::inverse(ctx, v, L, t, tw.data());
                   ctx.sync();
                   if (live)
                       for (int j = 0; j < E; ++j)
                           x[t + T * j] = v[j];
               }
           });
    for (size_t i = 0; i < buf.size(); ++i) {
        data[2 * i] = (double)buf[i].x;
        data[2 * i + 1] = (double)buf[i].y;
    }
    return 0;
}

} // namespace

extern "C" {

int emu_run(int n, int precision, int B, double nu, double alpha, double sigma, double amp,
            int kf, int kmin, int kmax, uint64_t seed, double dt, int steps, int grf,
            const double *init, const double *forcing, int forcing_shared, double *out,
            double *diag) {
    if (precision == 0)
        return run_any<float>(n, B, nu, alpha, sigma, amp, kf, kmin, kmax, seed, dt, steps, grf,
                              init, forcing, forcing_shared, out, diag);
    return run_any<double>(n, B, nu, alpha, sigma, amp, kf, kmin, kmax, seed, dt, steps, grf,
                           init, forcing, forcing_shared, out, diag);
}

int emu_fft_lines(int n, int precision, int lines, int inverse, double *data) {
#define HZL_EMU_FFT(NN)                                                                        \
    case NN:                                                                                   \
        return precision == 0 ? fft_lines<NN, float>(lines, inverse, data)                     \
                              : fft_lines<NN, double>(lines, inverse, data);
    switch (n) {
        HZL_EMU_FFT(16)
        HZL_EMU_FFT(32)
        HZL_EMU_FFT(64)
        HZL_EMU_FFT(128)
        HZL_EMU_FFT(256)
        HZL_EMU_FFT(512)
    }
#undef HZL_EMU_FFT
    return -1;
}
}
