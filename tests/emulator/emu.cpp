// emu.cpp -- CPU lane emulator for the engine's kernel bodies (TEST ONLY).
//
// Runs the unmodified __host__ __device__ kernel bodies of
// hzl-stochasticflowinference_amd/csrc/kernels.hpp on the host: one std::thread
// per lane of a workgroup, std::barrier for __syncthreads, a host vector for
// LDS, and the engine's launch sequence (kernels.hpp header comment) replayed
// over host arrays.  Used by the CPU test suite to check indexing, layouts and
// numerics against the oracle without a GPU.  Never linked into the product.
#include "fno_fused.hpp"
#include "gemm_mfma.hpp"
#include "surrogate_kernels.hpp"
#include "tables.hpp"
#include "variants.hpp"

#include <barrier>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

using namespace hzl_amd;

namespace {

// Operands of one emulated matrix instruction, exchanged between lanes.
struct MfmaScratch {
    float a[64][8], b[64][8];
};

struct CpuCtx {
    int lane;
    std::barrier<> *bar;
    MfmaScratch *mx = nullptr;
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
    // v_mfma_f32_32x32x16_bf16: lane l supplies A row l%32, k (l/32)*8 + j and
    // B column l%32, same k; accumulator i is D[8(i/4) + 4(l/32) + i%4][l%32].
    void mfma_bf16(const bf16x8_t &fa, const bf16x8_t &fb, f32x16_t &acc, int l) const {
        for (int j = 0; j < 8; ++j) {
            mx->a[l][j] = (float)fa[j];
            mx->b[l][j] = (float)fb[j];
        }
        bar->arrive_and_wait();
        for (int i = 0; i < 16; ++i) {
            const int row = 8 * (i / 4) + 4 * (l / 32) + i % 4, col = l % 32;
            float s = 0;
            for (int kb = 0; kb < 2; ++kb)
                for (int j = 0; j < 8; ++j)
                    s += mx->a[row + 32 * kb][j] * mx->b[col + 32 * kb][j];
            acc[i] += s;
        }
        bar->arrive_and_wait();
    }
    // value of v held by lane (l ^ mask)
    float shfl_xor(float v, int mask, int l) const {
        mx->a[l][0] = v;
        bar->arrive_and_wait();
        const float r = mx->a[l ^ mask][0];
        bar->arrive_and_wait();
        return r;
    }
    // v_mfma_f32_32x32x2_f32: lane l supplies A[l%32][l/32] and B[l/32][l%32].
    void mfma_f32(float fa, float fb, f32x16_t &acc, int l) const {
        mx->a[l][0] = fa;
        mx->b[l][0] = fb;
        bar->arrive_and_wait();
        for (int i = 0; i < 16; ++i) {
            const int row = 8 * (i / 4) + 4 * (l / 32) + i % 4, col = l % 32;
            acc[i] += mx->a[row][0] * mx->b[col][0] + mx->a[row + 32][0] * mx->b[col + 32][0];
        }
        bar->arrive_and_wait();
    }
};

// Launch `blocks` workgroups of `threads` lanes; body(ctx, block, lds).
// Each wavefront (64 lanes) exchanges its matrix-instruction operands through
// its own scratch; every lane of the workgroup reaches each barrier.
template <class Body> void launch(long blocks, int threads, size_t lds_bytes, Body body) {
    std::barrier<> bar(threads);
    std::vector<unsigned char> lds(lds_bytes + 64);
    std::vector<std::thread> pool;
    std::vector<MfmaScratch> mx((size_t)(threads + 63) / 64);
    pool.reserve((size_t)threads);
    for (int lane = 0; lane < threads; ++lane)
        pool.emplace_back([&, lane] {
            CpuCtx ctx{lane, &bar, &mx[(size_t)lane / 64]};
            for (long b = 0; b < blocks; ++b) {
                body(ctx, b, lds.data());
                bar.arrive_and_wait();
            }
        });
    for (auto &t : pool)
        t.join();
}

// GC / GR: launch shapes of the column and row kernels (variants.hpp);
// DIRECT: row spectra straight from the mixed fields (row_body_direct).
template <int N, typename R, class GC = DefaultGeometry<N, R>, class GR = DefaultGeometry<N, R>, bool DIRECT = false>
struct Sim {
    using G = Geometry<N>;
    using FFT = GroupFFT<N, GC::E, R>;
    int B;
    std::vector<cx<R>> w0, w1, acc, mix, nl, ext, tw, tw_row, noise;
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
          tw(twiddles<R>(N, GC::E)), tw_row(twiddles<R>(N, GR::E)), cfl((size_t)b * (N / 2), 0.0),
          nu(b, nu0), alpha(b, alpha0),
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
        const long blocks = ((long)B * G::KP + GC::COL_G - 1) / GC::COL_G;
        launch(blocks, GC::COL_THREADS, sizeof(cx<R>) * GC::COL_G * FFT::LDS_ELEMS,
               [&](const CpuCtx &ctx, long b, unsigned char *lds) {
                   col_body<N, MODE, R, GC>(ctx, a, reinterpret_cast<cx<R> *>(lds), b);
               });
    }

    void col(int mode, double dt, const cx<R> *e, long long es) {
        StageArgs<R> a = args(dt, e, es);
        col_any(mode, a);
    }

    void col_any(int mode, const StageArgs<R> &a) {
        switch (mode) {
        case 0:
            return col_mode<0>(a);
        case 1:
            return col_mode<1>(a);
        case 2:
            return col_mode<2>(a);
        case 3:
            return col_mode<3>(a);
        case 4:
            return col_mode<4>(a);
        case 5:
            return col_mode<5>(a);
        case 6:
            return col_mode<6>(a);
        case 7:
            return col_mode<7>(a);
        case 8:
            return col_mode<8>(a);
        case 9:
            return col_mode<9>(a);
        default:
            return col_mode<10>(a);
        }
    }

    StageArgs<R> args(double dt, const cx<R> *e, long long es) {
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
        return a;
    }

    // Row pass over an arbitrary mixed buffer into an arbitrary nonlinear field.
    void row_on(cx<R> *mx, cx<R> *out) {
        RowArgs<R> a{B, mx, out, tw_row.data(), nullptr};
        const long blocks = ((long)B * (N / 2) + GR::ROW_PPW - 1) / GR::ROW_PPW;
        launch(blocks, GR::ROW_THREADS, sizeof(cx<R>) * RowShape<GR>::template lds_elems<R>(),
               [&](const CpuCtx &ctx, long b, unsigned char *lds) {
                   row_body<N, R, false, GR>(ctx, a, reinterpret_cast<cx<R> *>(lds), b);
               });
    }

    // J(x, y) into nl: two-state column pass (mode 8) and a row pass.
    void jacobian(cx<R> *x, cx<R> *y, std::vector<cx<R>> &mx, std::vector<cx<R>> &out) {
        StageArgs<R> a = args(1.0, nullptr, 0);
        a.w0 = x;
        a.w1 = y;
        a.mix = mx.data();
        col_any(8, a);
        row_on(mx.data(), out.data());
    }

    // kind 0: drift of w0; 1: tangent along aux; 2: adjoint applied to aux.
    void calibration(int kind, std::vector<cx<R>> &aux, std::vector<cx<R>> &out) {
        std::vector<cx<R>> mA((size_t)B * 4 * G::MIXF), mB(mA.size()), nA((size_t)B * G::MIXF),
            nB(nA.size()), nC(nA.size());
        StageArgs<R> a = args(1.0, nullptr, 0);
        a.out = out.data();
        if (kind == 0) {
            jacobian(w0.data(), w0.data(), mA, nA);
            a.nl = nA.data();
            a.w1 = w0.data();
            col_any(9, a);
        } else if (kind == 1) {
            jacobian(aux.data(), w0.data(), mA, nA);
            jacobian(w0.data(), aux.data(), mB, nB);
            a.nl = nA.data();
            a.nl2 = nB.data();
            a.w1 = aux.data();
            col_any(9, a);
        } else {
            jacobian(w0.data(), aux.data(), mA, nA);
            StageArgs<R> m0 = args(1.0, nullptr, 0);
            m0.mix = mB.data();
            col_any(0, m0);
            ProductArgs<R> pa{B, mA.data(), mB.data(), nB.data(), nC.data(), tw_row.data()};
            const long blocks = ((long)B * (N / 2) + GR::ROW_PPW - 1) / GR::ROW_PPW;
            launch(blocks, GR::ROW_THREADS,
                   sizeof(cx<R>) * GR::ROW_PPW * GroupFFT<N, GR::E, R>::LDS_ELEMS,
                   [&](const CpuCtx &ctx, long b, unsigned char *lds) {
                       row_body_products<N, R, GR>(ctx, pa, reinterpret_cast<cx<R> *>(lds), b);
                   });
            a.nl = nA.data();
            a.nl2 = nB.data();
            a.nl3 = nC.data();
            a.w1 = aux.data();
            col_any(10, a);
        }
    }

    // SSPRK(3,3) steps with device-style noise (modes 5, 6, 7).
    void ssp_steps(double dt, int count) {
        for (int s = 0; s < count; ++s) {
            col(0, dt, nullptr, 0);
            make_noise(dt);
            for (int stage = 1; stage <= 3; ++stage) {
                row(stage == 1);
                col(4 + stage, dt, nullptr, 0);
            }
            ++step;
        }
    }

    void row(bool with_cfl) {
        RowArgs<R> a{B, mix.data(), nl.data(), tw_row.data(), with_cfl ? cfl.data() : nullptr};
        const long blocks = ((long)B * (N / 2) + GR::ROW_PPW - 1) / GR::ROW_PPW;
        if constexpr (DIRECT) {
            launch(blocks, GR::ROW_THREADS,
                   sizeof(cx<R>) * GR::ROW_PPW * GroupFFT<N, GR::E, R>::LDS_ELEMS,
                   [&](const CpuCtx &ctx, long b, unsigned char *lds) {
                       if (with_cfl)
                           row_body_direct<N, R, true, GR>(ctx, a, reinterpret_cast<cx<R> *>(lds), b);
                       else
                           row_body_direct<N, R, false, GR>(ctx, a, reinterpret_cast<cx<R> *>(lds), b);
                   });
            return;
        }
        launch(blocks, GR::ROW_THREADS, sizeof(cx<R>) * RowShape<GR>::template lds_elems<R>(),
               [&](const CpuCtx &ctx, long b, unsigned char *lds) {
                   if (with_cfl)
                       row_body<N, R, true, GR>(ctx, a, reinterpret_cast<cx<R> *>(lds), b);
                   else
                       row_body<N, R, false, GR>(ctx, a, reinterpret_cast<cx<R> *>(lds), b);
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

template <int N, typename R, class GC = DefaultGeometry<N, R>, class GR = DefaultGeometry<N, R>, bool DIRECT = false>
int run(int B, double nu, double alpha, double sigma, double amp, int kf, int kmin, int kmax,
        uint64_t seed, double dt, int steps, int grf, const double *init, const double *forcing,
        int forcing_shared, double *out, double *diag) {
    Sim<N, R, GC, GR, DIRECT> s(B, nu, alpha, sigma, amp, kf, kmin, kmax, seed);
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

// Batch of 1-D transforms through GroupFFT (E values per lane; two or three
// passes by the split): lines[L][N] in natural order.
template <int N, typename R, int E = Shape<N>::E> int fft_lines(int lines, int inverse,
                                                                 double *data) {
    using F = GroupFFT<N, E, R>;
    constexpr int T = N / E;
    const auto tw = twiddles<R>(N, E);
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

// One product through gemm_wave, one wavefront per 32 x 32 tile.  Operands
// arrive as FP32 host arrays and are rounded to T; C (and pre) are returned
// as FP32 after rounding to the output type.
template <typename T>
int gemm_any(const int *dims, const long *st, const float *A, size_t na, const float *B,
             size_t nb, float *C, size_t nc, const float *bias, const float *aux, size_t nx,
             float *pre, float alpha, float beta, int act, int c_bf16) {
    std::vector<T> a(na), b(nb), x(nx);
    for (size_t i = 0; i < na; ++i)
        a[i] = (T)A[i];
    for (size_t i = 0; i < nb; ++i)
        b[i] = (T)B[i];
    for (size_t i = 0; i < nx; ++i)
        x[i] = (T)aux[i];
    std::vector<unsigned char> c(nc * 4), p(nc * 4);
    for (size_t i = 0; i < nc; ++i) {
        if (c_bf16) {
            reinterpret_cast<__bf16 *>(c.data())[i] = (__bf16)C[i];
        } else {
            reinterpret_cast<float *>(c.data())[i] = C[i];
        }
    }
    GemmArgs<T> g{};
    g.M = dims[0];
    g.N = dims[1];
    g.K = dims[2];
    g.nb1 = dims[3];
    g.nb2 = dims[4];
    g.A = a.data();
    g.a_m = st[0];
    g.a_k = st[1];
    g.a_b1 = st[2];
    g.a_b2 = st[3];
    g.B = b.data();
    g.b_n = st[4];
    g.b_k = st[5];
    g.b_b1 = st[6];
    g.b_b2 = st[7];
    g.C = c.data();
    g.c_m = st[8];
    g.c_n = st[9];
    g.c_b1 = st[10];
    g.c_b2 = st[11];
    g.c_type = c_bf16 ? OUT_BF16 : OUT_F32;
    g.alpha = alpha;
    g.beta = beta;
    g.bias = bias;
    g.aux = nx ? x.data() : nullptr;
    g.x_m = st[12];
    g.x_n = st[13];
    g.x_b1 = st[14];
    g.x_b2 = st[15];
    g.aux_type = sizeof(T) == 2 ? OUT_BF16 : OUT_F32;
    g.pre = pre ? p.data() : nullptr;
    g.act = act;
    const long tiles = (long)g.nb1 * g.nb2 * ((g.M + 31) / 32) * ((g.N + 31) / 32);
    launch(tiles, 64, GemmTile<T>::LDS_PER_WAVE * sizeof(T),
           [&](const CpuCtx &ctx, long tile, unsigned char *lds) {
               gemm_wave<T>(ctx, g, tile, ctx.lane, reinterpret_cast<T *>(lds));
           });
    for (size_t i = 0; i < nc; ++i) {
        C[i] = c_bf16 ? (float)reinterpret_cast<__bf16 *>(c.data())[i]
                      : reinterpret_cast<float *>(c.data())[i];
        if (pre)
            pre[i] = c_bf16 ? (float)reinterpret_cast<__bf16 *>(p.data())[i]
                            : reinterpret_cast<float *>(p.data())[i];
    }
    return 0;
}

template <typename T>
int proj_head_any(long M, int C, int H, const float *X, const float *Q1, const float *q1,
                  const float *Q2, const float *q2, float *out) {
    std::vector<T> x((size_t)M * C), q((size_t)H * C);
    for (size_t i = 0; i < x.size(); ++i)
        x[i] = (T)X[i];
    for (size_t i = 0; i < q.size(); ++i)
        q[i] = (T)Q1[i];
    HeadArgs<T> a{x.data(), q.data(), q1, Q2, q2, out, M, C, H};
    launch((M + 31) / 32, 64, GemmTile<T>::LDS_PER_WAVE * sizeof(T),
           [&](const CpuCtx &ctx, long wave, unsigned char *lds) {
               proj_head_wave<T>(ctx, a, wave, ctx.lane, reinterpret_cast<T *>(lds));
           });
    return 0;
}

// Calibration drifts (kind 0 drift, 1 tangent along aux, 2 adjoint of aux) or
// (kind 3) `steps` SSPRK(3,3) steps, of B members from reference Fields,
// in the default launch shape.
template <int N, typename R>
int calib(int kind, int B, double nu, double alpha, double sigma, double amp, int kf, int kmin,
          int kmax, uint64_t seed, double dt, int steps, const double *state, const double *aux,
          double *out) {
    Sim<N, R> s(B, nu, alpha, sigma, amp, kf, kmin, kmax, seed);
    const size_t nn = (size_t)N * N;
    std::vector<cx<R>> x((size_t)B * Geometry<N>::SPEC), o(x.size());
    for (int m = 0; m < B; ++m) {
        full_to_half<R>(N, state + 2 * nn * m, s.w0.data() + (size_t)m * Geometry<N>::SPEC);
        if (aux)
            full_to_half<R>(N, aux + 2 * nn * m, x.data() + (size_t)m * Geometry<N>::SPEC);
    }
    if (kind == 3) {
        s.ssp_steps(dt, steps);
        o = s.w0;
    } else {
        s.calibration(kind, x, o);
    }
    for (int m = 0; m < B; ++m)
        half_to_full<R>(N, o.data() + (size_t)m * Geometry<N>::SPEC, out + 2 * nn * m);
    return 0;
}

} // namespace

extern "C" {

// OU forcing kernel body (kernels.hpp ou_body) over B members (per-member sigma
// and tau) for `calls`
// calls: modes[c] (0 start, 1 transition over hs[c], 2 reseed from
// keys + c * B); after each call the coefficients [B][npairs] and, for
// transitions, the midpoint forcing in engine order go to F_hist / noise_hist.
int emu_ou(int B, int npairs, const int *map, const uint64_t *keys, const double *sigma,
           const double *tau, int calls, const int *modes, const double *hs, double *F_hist,
           double *noise_hist) {
    std::vector<uint64_t> mt((size_t)(MT_N + 1) * B);
    std::vector<cx<double>> F((size_t)B * npairs), noise((size_t)B * npairs);
    for (int c = 0; c < calls; ++c) {
        OuArgs<double> a{B,        npairs,   mt.data(), F.data(), map,         sigma,
                         keys + (long)c * B, tau, hs[c], modes[c], modes[c] == 1 ? noise.data() : nullptr,
                         nullptr,  1};
        launch(B, OU_LANES, sizeof(uint64_t) * MT_N + 64,
               [&](const CpuCtx &ctx, long b, unsigned char *lds) {
                   auto *words = reinterpret_cast<uint64_t *>(lds);
                   auto *pos = reinterpret_cast<int *>(lds + sizeof(uint64_t) * MT_N);
                   ou_body<double>(ctx, a, words, pos, (int)b);
               });
        std::memcpy(F_hist + (size_t)c * 2 * B * npairs, F.data(), sizeof(double) * 2 * B * npairs);
        std::memcpy(noise_hist + (size_t)c * 2 * B * npairs, noise.data(),
                    sizeof(double) * 2 * B * npairs);
    }
    return 0;
}

// The device NormalStream (kernels.hpp normal_stream_body) on `lanes` CPU
// lanes: a stream seeded like std::mt19937_64(seed), then `calls` calls of
// counts[c] normals each, written one after another to out.
int emu_normal_stream(uint64_t seed, int lanes, int calls, const long *counts, double *out) {
    std::vector<uint64_t> mt(MT_N + 1);
    mt[0] = seed;
    for (int i = 1; i < MT_N; ++i)
        mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
    mt[MT_N] = MT_N;
    double sp[2] = {0, 0};
    long offset = 0;
    for (int c = 0; c < calls; ++c) {
        launch(1, lanes, sizeof(uint64_t) * MT_N, [&](const CpuCtx &ctx, long, unsigned char *lds) {
            normal_stream_body(ctx, mt.data(), sp, counts[c], out + offset,
                               reinterpret_cast<uint64_t *>(lds), lanes);
        });
        offset += counts[c];
    }
    return 0;
}

int emu_proj_head(int bf16, long M, int C, int H, const float *X, const float *Q1,
                  const float *q1, const float *Q2, const float *q2, float *out) {
    if (bf16)
        return proj_head_any<__bf16>(M, C, H, X, Q1, q1, Q2, q2, out);
    return proj_head_any<float>(M, C, H, X, Q1, q1, Q2, q2, out);
}

// Fused surrogate layer passes (fno_fused.hpp) over S samples; operands
// arrive as FP32 and are rounded to BF16, outputs return as FP32.
static std::vector<__bf16> to_bf16(const float *x, size_t n) {
    std::vector<__bf16> o(x ? n : 0);
    for (size_t i = 0; i < o.size(); ++i)
        o[i] = (__bf16)x[i];
    return o;
}

// feat (with Lw, Lb, nu, amp, kf) non-null: layer 0's pass, v lifted inside
// the pass from the grid feature (channel 0 of [S][G][G][CIN]) and the
// samples' viscosity and forcing features (v is then not read).
int emu_fno_forward(long S, const float *v, const float *Fx, const float *Fy2, const float *feat,
                    const float *Lw, const float *Lb, const double *nu, const double *amp,
                    int kf, float *V) {
    using namespace fno;
    const bool lifted = feat != nullptr;
    const auto bv = to_bf16(lifted ? nullptr : v, (size_t)S * SAMPLE), bx = to_bf16(Fx, 2 * KX * G),
               by = to_bf16(Fy2, 2 * KY * 2 * G), bf = to_bf16(feat, (size_t)S * G * G * CIN),
               bw = to_bf16(Lw, (size_t)C * CIN);
    std::vector<__bf16> out((size_t)S * SPECTRAL);
    FnoForwardArgs a{lifted ? nullptr : bv.data(), bx.data(), by.data(), out.data(), S,
                     FnoLift{lifted ? bf.data() : nullptr, lifted ? bw.data() : nullptr, Lb, nu,
                             amp, kf}};
    launch(S, THREADS, (lifted ? FWD_LDS_LIFT : FWD_LDS) * sizeof(__bf16),
           [&](const CpuCtx &ctx, long b, unsigned char *l) {
               if (lifted)
                   fno_forward_body<true>(ctx, a, reinterpret_cast<__bf16 *>(l), b);
               else
                   fno_forward_body<false>(ctx, a, reinterpret_cast<__bf16 *>(l), b);
           });
    for (size_t i = 0; i < out.size(); ++i)
        V[i] = (float)out[i];
    return 0;
}

int emu_fno_inverse(long S, const float *U, const float *Gy2, const float *Gx, const float *v,
                    const float *W, const float *bias, int act, const float *feat,
                    const float *Lw, const float *Lb, const double *nu, const double *amp,
                    int kf, float *out) {
    using namespace fno;
    const bool lifted = feat != nullptr;
    const auto bu = to_bf16(U, (size_t)S * SPECTRAL), by = to_bf16(Gy2, 2 * G * 2 * KY),
               bx = to_bf16(Gx, G * 2 * KX), bv = to_bf16(lifted ? nullptr : v, (size_t)S * SAMPLE),
               bw = to_bf16(W, C * C), bf = to_bf16(feat, (size_t)S * G * G * CIN),
               bl = to_bf16(Lw, (size_t)C * CIN);
    std::vector<__bf16> o((size_t)S * SAMPLE);
    FnoInverseArgs a{bu.data(), by.data(), bx.data(), lifted ? nullptr : bv.data(), bw.data(),
                     bias, o.data(), S, act,
                     FnoLift{lifted ? bf.data() : nullptr, lifted ? bl.data() : nullptr, Lb, nu,
                             amp, kf}};
    launch(S * NBLKI, THREADS, (lifted ? INV_LDS_LIFT : INV_LDS) * sizeof(__bf16),
           [&](const CpuCtx &ctx, long b, unsigned char *l) {
               if (lifted)
                   fno_inverse_body<true>(ctx, a, reinterpret_cast<__bf16 *>(l), b);
               else
                   fno_inverse_body<false>(ctx, a, reinterpret_cast<__bf16 *>(l), b);
           });
    for (size_t i = 0; i < o.size(); ++i)
        out[i] = (float)o[i];
    return 0;
}

int emu_fno_head(long M, const float *X, const float *Q1, const float *q1, const float *Q2,
                 const float *q2, float *out) {
    using namespace fno;
    std::vector<__bf16> x((size_t)M * C), q((size_t)FNO_H * C);
    for (size_t i = 0; i < x.size(); ++i)
        x[i] = (__bf16)X[i];
    for (size_t i = 0; i < q.size(); ++i)
        q[i] = (__bf16)Q1[i];
    FnoHeadArgs a{x.data(), q.data(), q1, Q2, q2, out, M};
    launch((M + 31) / 32, 64, 2 * FNO_H * sizeof(float),
           [&](const CpuCtx &ctx, long wave, unsigned char *l) {
               float *hq = reinterpret_cast<float *>(l);
               for (int e = ctx.lane; e < FNO_H; e += 64) {
                   hq[e] = a.q1[e];
                   hq[FNO_H + e] = a.Q2[e];
               }
               ctx.sync();
               fno_head_wave(ctx, a, wave, ctx.lane, hq);
           });
    return 0;
}

// Fused state -> grid (FP32 output [S][G][G]) and grid -> increment passes.
int emu_state_to_grid(long S, int K, const float *states, const float *g1, const float *g2,
                      float alpha, float *out) {
    using namespace fno;
    const long spec2 = 2L * (K + 1) * (2 * K + 1);
    StateGridArgs a{states, spec2, K, 2 * K + 1, g1, g2, out, 0, G, 1, (long)G * G, alpha, S};
    launch(S, SG_THREADS, SG_LDS * sizeof(float), [&](const CpuCtx &ctx, long b, unsigned char *l) {
        fno_state_to_grid_body(ctx, a, reinterpret_cast<float *>(l), b);
    });
    return 0;
}

int emu_grid_to_increment(long S, const float *grid, const float *f1, const float *f2,
                          float *incr) {
    using namespace fno;
    GridIncrementArgs a{grid, f1, f2, incr, S};
    launch(S, SG_THREADS, GS_LDS * sizeof(float), [&](const CpuCtx &ctx, long b, unsigned char *l) {
        fno_grid_to_increment_body(ctx, a, reinterpret_cast<float *>(l), b);
    });
    return 0;
}

int emu_fno_mix(long S, int modes, const float *V, const float *R, float *U) {
    using namespace fno;
    const size_t nv = (size_t)S * modes * 2 * C, nr = (size_t)modes * 4 * C * C;
    std::vector<__bf16> v(nv), rr(nr), u(nv);
    for (size_t i = 0; i < nv; ++i)
        v[i] = (__bf16)V[i];
    for (size_t i = 0; i < nr; ++i)
        rr[i] = (__bf16)R[i];
    FnoMixArgs a{v.data(), rr.data(), u.data(), S, modes};
    launch(modes, MIX_THREADS, 0, [&](const CpuCtx &ctx, long b, unsigned char *) {
        fno_mix_body(ctx, a, b);
    });
    for (size_t i = 0; i < nv; ++i)
        U[i] = (float)u[i];
    return 0;
}

int emu_gelu_sg(long n, const float *x, float *y) {
    for (long i = 0; i < n; ++i)
        y[i] = gelu_sg(x[i]);
    return 0;
}

int emu_gemm(int bf16, const int *dims, const long *strides, const float *A, size_t na,
             const float *B, size_t nb, float *C, size_t nc, const float *bias, const float *aux,
             size_t nx, float *pre, float alpha, float beta, int act, int c_bf16) {
    if (bf16)
        return gemm_any<__bf16>(dims, strides, A, na, B, nb, C, nc, bias, aux, nx, pre, alpha,
                                beta, act, c_bf16);
    return gemm_any<float>(dims, strides, A, na, B, nb, C, nc, bias, aux, nx, pre, alpha, beta,
                           act, c_bf16);
}

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

// FP32 engine with column variant cv and the default row shape (rv == 0), or
// the default column shape and row variant rv (cv == 0); variants.hpp.
int emu_run_variant(int n, int cv, int rv, int B, double nu, double alpha, double sigma,
                    double amp, int kf, int kmin, int kmax, uint64_t seed, double dt, int steps,
                    int grf, const double *init, double *out, double *diag) {
    if (rv == 0) {
#define HZL_X(NN, IDX, E, CT, W, S)                                                            \
    if (n == NN && cv == IDX)                                                                  \
        return run<NN, float, Geometry<NN, E, CT, 0, S>, Geometry<NN>>(                        \
            B, nu, alpha, sigma, amp, kf, kmin, kmax, seed, dt, steps, grf, init, nullptr, 1, out, \
            diag);
        HZL_COL_VARIANTS(HZL_X)
#undef HZL_X
    }
    if (cv == 0) {
#define HZL_X(NN, IDX, E, PPW, W, D, S)                                                        \
    if (n == NN && rv == IDX)                                                                  \
        return run<NN, float, Geometry<NN>, Geometry<NN, E, 64, PPW, S>, (D != 0)>(            \
            B, nu, alpha, sigma, amp, kf, kmin, kmax, seed, dt, steps, grf, init, nullptr, 1, out, \
            diag);
        HZL_ROW_VARIANTS(HZL_X)
#undef HZL_X
    }
    return -1;
}

int emu_calibration(int n, int precision, int kind, int B, double nu, double alpha,
                    double sigma, double amp, int kf, int kmin, int kmax, uint64_t seed,
                    double dt, int steps, const double *state, const double *aux, double *out) {
#define HZL_EMU_CAL(NN)                                                                        \
    case NN:                                                                                   \
        return precision == 0 ? calib<NN, float>(kind, B, nu, alpha, sigma, amp, kf, kmin,     \
                                                 kmax, seed, dt, steps, state, aux, out)       \
                              : calib<NN, double>(kind, B, nu, alpha, sigma, amp, kf, kmin,    \
                                                  kmax, seed, dt, steps, state, aux, out);
    switch (n) {
        HZL_EMU_CAL(16)
        HZL_EMU_CAL(32)
        HZL_EMU_CAL(64)
    }
#undef HZL_EMU_CAL
    return -1;
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

// The same with E values per lane: the splits of the launch-shape variants
// (three-pass where E < n / E).
int emu_fft_lines_split(int n, int E, int precision, int lines, int inverse, double *data) {
#define HZL_EMU_FFT(NN, EE)                                                                    \
    if (n == NN && E == EE)                                                                    \
        return precision == 0 ? fft_lines<NN, float, EE>(lines, inverse, data)                 \
                              : fft_lines<NN, double, EE>(lines, inverse, data);
    HZL_EMU_FFT(64, 8)
    HZL_EMU_FFT(128, 8)
    HZL_EMU_FFT(128, 32)
    HZL_EMU_FFT(256, 8)
    HZL_EMU_FFT(256, 32)
    HZL_EMU_FFT(512, 8)
#undef HZL_EMU_FFT
    return -1;
}
}
