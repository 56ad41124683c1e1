// kernels.hpp -- bodies of the batched spectral RK4 kernels.
//
// One member's state lives in HBM in the retained half plane of the strict 2/3
// square mask (reference mask: src/solver.cpp:126-135):
//     kx = c   in [0, K],  ky = j - K  for j in [0, 2K],  K = (N-1)/3,
// stored [member][c][j] (KP = K+1 columns of J = 2K+1 entries, ky fastest).
// The kx = 0 column keeps both signs of ky (they are conjugates).
//
// One right-hand side evaluation (reference Solver::rhs, src/solver.cpp:186-203)
// is split into two kernels that meet in "mixed" space (kx spectral, y
// physical), stored tiled as [member][field][y/16][c][y%16] so that both a
// column (fixed c, all y) and a 16-row tile (all c, 16 y) are read with full
// 128-byte segments:
//
//   col kernel  (one group of T threads per column c):
//     pass C of stage s:  forward y-FFT of the nonlinear column -> N_hat,
//                         RHS = -N_hat/N^2 - (nu k^2 + alpha) w + F on the mask,
//                         RK4 accumulate / stage update (src/solver.cpp:223-234)
//     pass A of stage s+1: psi = w/k^2; inverse y-FFTs of psi, ky psi, w, ky w
//   row kernel  (one group of T threads per pair of rows):
//     u = i(ky psi)~, v = -i kx psi~, wx = i kx w~, wy = i (ky w)~
//     two packed complex inverse x-FFTs per row (z1 = u + i wx, z2 = v + i wy),
//     product u wx + v wy, one packed forward x-FFT for the two rows,
//     keep kx in [0, K] -> nonlinear mixed field.
//
// A whole RK4 step is 4 x (row, col) launches; the col launch of stage 4 also
// produces the mixed fields of the next step's stage 1.
#pragma once
#include "fft_lds.hpp"
#include <cstdint>
#include <type_traits>

namespace hzl_amd {

// ------------------------------------------------------------- shapes ----
// Default split of the grid's transforms: E values per lane (two passes when
// E >= N / E, three when E < N / E) and CT lanes per column workgroup.
template <int N> struct Shape;
template <> struct Shape<16> {
    static constexpr int E = 4, CT = 64;
};
template <> struct Shape<32> {
    static constexpr int E = 8, CT = 64;
};
template <> struct Shape<64> {
    static constexpr int E = 16, CT = 64;
};
// 128: three passes over 16 lanes of 8 values; 0.60 -> 0.47 ms per step of
// 1024 members against the 8-lane, 16-value split
template <> struct Shape<128> {
    static constexpr int E = 8, CT = 128;
};
template <> struct Shape<256> {
    static constexpr int E = 16, CT = 64;
};
// 512: three passes over 64 lanes of 8 values (about 115 VGPRs in FP32)
// rather than two over 16 lanes of 32 (about 250, one or two waves per SIMD);
// four columns per workgroup
template <> struct Shape<512> {
    static constexpr int E = 8, CT = 256;
};

// Launch shape of the column and row kernels at grid N: EV values per lane
// (an N = T x EV two-pass transform), CT threads per column workgroup, PPW
// row pairs per row workgroup (0: the default), and NT: bit 0 streams the
// mixed and nonlinear fields with non-temporal loads and stores; bit 1 keeps
// twiddles in the table (L1) rather than in registers where that brings a
// kernel under the 168-register budget of three waves per SIMD: the column
// stages that hold three state arrays (RK4 stages 3 and 4) and the row pass.  Shape<N> gives the
// default; other shapes are variants the autotuner times.  (Streaming the RK4
// state, stage and accumulator arrays as well was measured 30% slower: later
// stages re-read them from cache.)
template <int N, int EV = Shape<N>::E, int CT = Shape<N>::CT, int PPW = 0, int NT = 0>
struct Geometry {
    static constexpr int NN = N;
    static constexpr bool STREAM = (NT & 1) != 0;
    static constexpr bool TABLE_TW = (NT & 2) != 0;
    static constexpr int E = EV;
    static constexpr int T = N / E;
    static constexpr int K = (N - 1) / 3;
    static constexpr int KP = K + 1;
    static constexpr int J = 2 * K + 1;
    static constexpr int TILES = N / 16;
    // column kernel: CT lanes per workgroup, G columns
    static constexpr int COL_THREADS = CT;
    static constexpr int COL_G = COL_THREADS / T;
    // row kernel: PPW row pairs per workgroup
    static constexpr int ROW_PPW = PPW ? PPW : ((64 / T) > 8 ? (64 / T) : 8);
    static constexpr int ROW_THREADS = ROW_PPW * T;
    static constexpr long SPEC = (long)KP * J; // complex entries per member state
    static constexpr long MIXF = (long)KP * N; // complex entries per mixed field
};

// Default launch shape of precision R at grid N.  An FP64 value takes two
// registers, so from 128^2 up FP64 takes the three-pass split (8 values per
// lane) where FP32 holds 16 or 32: the two-pass FP64 kernels there needed 256
// VGPRs (one wave per SIMD) and the row pass spilled to scratch.
template <int N, typename R> struct DefaultShape {
    static constexpr bool WIDE = sizeof(R) == 8 && N >= 128;
    static constexpr int E = WIDE ? 8 : Shape<N>::E;
    static constexpr int CT = WIDE ? (N == 512 ? 256 : 128) : Shape<N>::CT;
};
template <int N, typename R>
using DefaultGeometry = Geometry<N, DefaultShape<N, R>::E, DefaultShape<N, R>::CT>;

// 1/x: correctly rounded division in FP64; the hardware reciprocal (1 ulp)
// in FP32, where it feeds psi = w / |k|^2.
HZL_HD double recip(double x) {
    return 1.0 / x;
}
HZL_HD float recip(float x) {
#if __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// Twiddle argument of a group's transforms: the lane's twiddles in registers
// when the split is square (E == T: one set serves forward and inverse) and
// REG allows it, else the table itself.
template <class FFT, typename R, bool REG = true>
using TwArg = std::conditional_t<FFT::Q == 1 && REG, typename FFT::Twiddles, const cx<R> *>;

template <class FFT, typename R, bool REG = true>
HZL_HD TwArg<FFT, R, REG> lane_twiddles(const cx<R> *tw, int t) {
    if constexpr (FFT::Q == 1 && REG)
        return FFT::load_twiddles(tw, t);
    else
        return tw;
}

// Loads and stores of fields that are written once and read once per stage
// (mixed and nonlinear fields): with STREAM, non-temporal hints so they do not
// displace cached data (the column pass's mixed-field stores, whole 128-byte
// lines per group, and the loads of both passes).
template <bool STREAM, typename R> HZL_HD void stream_store(cx<R> *p, cx<R> v) {
#if __HIP_DEVICE_COMPILE__
    if constexpr (STREAM) {
        typedef R v2 __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(v2{v.x, v.y}, reinterpret_cast<v2 *>(p));
        return;
    }
#endif
    *p = v;
}

template <bool STREAM, typename R> HZL_HD cx<R> stream_load(const cx<R> *p) {
#if __HIP_DEVICE_COMPILE__
    if constexpr (STREAM) {
        typedef R v2 __attribute__((ext_vector_type(2)));
        const v2 x = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(p));
        return {x.x, x.y};
    }
#endif
    return *p;
}

HZL_HD long tiled(int KP, int c, int y) {
    return (long)(y >> 4) * (KP * 16) + c * 16 + (y & 15);
}

// ------------------------------------------------------------ streams ----
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
HZL_HD void philox4x32_10(uint32_t (&c)[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// SplitMix64-style logical stream identity (reference src/solver.cpp:61-66).
HZL_HD uint64_t stream_seed(uint64_t base, uint64_t index) {
    uint64_t x = base + 0x9e3779b97f4a7c15ULL * (index + 1);
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

constexpr uint32_t TAG_NOISE = 0x4E4F4953u; // "NOIS"
constexpr uint32_t TAG_GRF = 0x47524631u;   // "GRF1"
constexpr double PI = 3.14159265358979323846;

// Two standard normals (double) from one Philox block: 53-bit open-interval
// uniforms and Box-Muller (cos, sin).
HZL_HD void philox_normal2(uint64_t key, uint32_t c0, uint64_t c12, uint32_t tag, double &g0,
                           double &g1) {
    uint32_t c[4] = {c0, (uint32_t)c12, (uint32_t)(c12 >> 32), tag};
    philox4x32_10(c, (uint32_t)key, (uint32_t)(key >> 32));
    const uint64_t a = (((uint64_t)c[0] << 32) | c[1]) >> 11;
    const uint64_t b = (((uint64_t)c[2] << 32) | c[3]) >> 11;
    const double u1 = ((double)a + 0.5) * 0x1.0p-53, u2 = ((double)b + 0.5) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1)), th = 2.0 * PI * u2;
    g0 = r * cos(th);
    g1 = r * sin(th);
}

// ----------------------------------------------------------- forcing ----
// Shell white noise of one step, generated once per step for every member by
// a small kernel (so the FP64 Philox/Box-Muller code stays out of the FFT
// kernels): noise[m * npairs + p] = sigma_m / sqrt(2 dt) (g0 + i g1) with
// (g0, g1) from Philox keyed by member m, counter (p, step, "NOIS").
template <typename R> struct NoiseArgs {
    int B, npairs;
    const double *sigma;  // per member white-noise amplitude
    const uint64_t *keys; // per member Philox key
    double dt;
    uint64_t step;        // step index, or the offset added to *step_base
    cx<R> *noise;
    const uint64_t *step_base = nullptr; // device step counter (graph replays)
};

template <typename R> HZL_HD void noise_body(const NoiseArgs<R> &a, long i) {
    if (i >= (long)a.B * a.npairs)
        return;
    const int m = (int)(i / a.npairs), p = (int)(i % a.npairs);
    const double s = a.sigma[m];
    cx<R> v = {0, 0};
    if (s > 0) {
        double g0, g1;
        const uint64_t step = a.step_base ? *a.step_base + a.step : a.step;
        philox_normal2(a.keys[m], (uint32_t)p, step, TAG_NOISE, g0, g1);
        const double scale = s / sqrt(2.0 * a.dt);
        v = {(R)(scale * g0), (R)(scale * g1)};
    }
    a.noise[i] = v;
}

// ------------------------------------------------ OU shell forcing ----
// The reference's forcing process (src/solver.cpp:41-54,84-105,174-176,
// 217-254): per member an mt19937_64 stream read through NormalStream
// (Box-Muller, first draw r cos, second r sin) and one FP64 OU coefficient
// per shell pair, in the reference's pair order (forcing_pairs,
// src/solver.cpp:67-83: grid index y*n + x of the pair's first mode, rising).
// Every use draws two normals per pair, so a stream never holds a cached
// second normal between calls and its position stays even: pair j of a call
// takes mt outputs pos + 2j and pos + 2j + 1, and the lanes of the member's
// workgroup draw in parallel.  Tempering only reads the state words, so a call
// that stays inside the current block reads the 2 P words it draws and writes
// back the position alone.  A call that crosses the block end stages the 312
// words in LDS and regenerates them (twist) in two parallel phases, words
// 0..155 from the old words and 156..311 from the new words 0..155.
constexpr int OU_LANES = 64;
constexpr int MT_N = 312;

template <typename R> struct OuArgs {
    int B, npairs;
    uint64_t *mt;          // [B][MT_N + 1]: state words, then the position
    cx<double> *F;         // [B][npairs] coefficients, reference pair order
    const int *map;        // [npairs] engine pair index, ~index when conjugated
    const double *sigma;   // per member
    const uint64_t *keys;  // per member seeds (modes 0 and 2)
    const double *tau;     // per member OU time constant
    double h;              // transition interval (mode 1)
    int mode;              // 0: seed + stationary start, 1: transition, 2: reseed,
                           // 3: publish the current coefficients to noise (no draws)
    cx<R> *noise;          // modes 1, 3: forcing [B][npairs] (engine order) or null
    const double *stats;   // mode 2: only members of filters that resampled
    int P;                 // particles per filter (mode 2)
};

HZL_HD uint64_t mt_temper(uint64_t y) {
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    return y ^ (y >> 43);
}

HZL_HD uint64_t mt_mix(uint64_t a, uint64_t b) {
    const uint64_t x = (a & 0xFFFFFFFF80000000ULL) | (b & 0x7FFFFFFFULL);
    return (x >> 1) ^ ((x & 1) ? 0xB5026F5AA96619E9ULL : 0ULL);
}

// NormalStream's open-interval uniform ((x >> 11) + 0.5) 2^-53.
HZL_HD double mt_open_uniform(uint64_t x) {
    return ((double)(x >> 11) + 0.5) * 0x1.0p-53;
}

template <typename R, class Ctx>
HZL_HD void ou_body(const Ctx &ctx, const OuArgs<R> &a, uint64_t *s, int *shared_pos, int m) {
    const int lane = ctx.tid();
    uint64_t *words = a.mt + (long)m * (MT_N + 1);
    if (a.stats && a.stats[8 * (m / a.P) + 3] == 0)
        return; // filter did not resample (uniform across the workgroup)
    if (a.mode == 3) {
        for (int j = lane; j < a.npairs; j += OU_LANES) {
            const cx<double> v = a.F[(long)m * a.npairs + j];
            const int p = a.map[j] >= 0 ? a.map[j] : ~a.map[j];
            a.noise[(long)m * a.npairs + p] =
                a.map[j] >= 0 ? cx<R>{(R)v.x, (R)v.y} : cx<R>{(R)v.x, (R)(-v.y)};
        }
        return;
    }
    if (a.mode == 0 || a.mode == 2) {
        if (lane == 0) {
            s[0] = a.keys[m];
            for (int i = 1; i < MT_N; ++i)
                s[i] = 6364136223846793005ULL * (s[i - 1] ^ (s[i - 1] >> 62)) + (uint64_t)i;
            *shared_pos = MT_N;
        }
    } else {
        const int p0 = (int)words[MT_N];
        if (p0 + 2 * a.npairs <= MT_N) {
            // inside the current block: draw straight from the state words
            const double sg = a.sigma[m], tau = a.tau[m], decay = exp(-a.h / tau),
                         scale = sg * sqrt(-tau * expm1(-2 * a.h / tau) / 4);
            for (int j = lane; j < a.npairs; j += OU_LANES) {
                const double u1 = mt_open_uniform(mt_temper(words[p0 + 2 * j]));
                const double u2 = mt_open_uniform(mt_temper(words[p0 + 2 * j + 1]));
                const double r = sqrt(-2 * log(u1)), angle = 2 * 3.14159265358979323846 * u2;
                cx<double> &f = a.F[(long)m * a.npairs + j];
                const cx<double> v = {decay * f.x + scale * (r * cos(angle)),
                                      decay * f.y + scale * (r * sin(angle))};
                f = v;
                if (a.noise) {
                    const int p = a.map[j] >= 0 ? a.map[j] : ~a.map[j];
                    a.noise[(long)m * a.npairs + p] =
                        a.map[j] >= 0 ? cx<R>{(R)v.x, (R)v.y} : cx<R>{(R)v.x, (R)(-v.y)};
                }
            }
            ctx.sync(); // every lane has read the position before it moves
            if (lane == 0)
                words[MT_N] = (uint64_t)(p0 + 2 * a.npairs);
            return;
        }
        for (int i = lane; i < MT_N; i += OU_LANES)
            s[i] = words[i];
        if (lane == 0)
            *shared_pos = p0;
    }
    ctx.sync();
    int pos = *shared_pos;
    if (a.mode != 2) {
        const double sg = a.sigma[m], tau = a.tau[m];
        double decay = 0, scale;
        if (a.mode == 0) {
            scale = sg * sqrt(tau / 4);
        } else {
            decay = exp(-a.h / tau);
            scale = sg * sqrt(-tau * expm1(-2 * a.h / tau) / 4);
        }
        int done = 0;
        while (done < a.npairs) {
            if (pos >= MT_N) {
                uint64_t nv[(MT_N / 2 + OU_LANES - 1) / OU_LANES];
                int c = 0;
                for (int i = lane; i < MT_N / 2; i += OU_LANES)
                    nv[c++] = s[i + MT_N / 2] ^ mt_mix(s[i], s[i + 1]);
                ctx.sync();
                c = 0;
                for (int i = lane; i < MT_N / 2; i += OU_LANES)
                    s[i] = nv[c++];
                ctx.sync();
                c = 0;
                for (int i = MT_N / 2 + lane; i < MT_N; i += OU_LANES)
                    nv[c++] = s[i - MT_N / 2] ^ mt_mix(s[i], s[(i + 1) % MT_N]);
                ctx.sync();
                c = 0;
                for (int i = MT_N / 2 + lane; i < MT_N; i += OU_LANES)
                    s[i] = nv[c++];
                ctx.sync();
                pos = 0;
            }
            const int count = (a.npairs - done) < (MT_N - pos) / 2 ? (a.npairs - done)
                                                                      : (MT_N - pos) / 2;
            for (int l = lane; l < count; l += OU_LANES) {
                const int j = done + l;
                const double u1 = mt_open_uniform(mt_temper(s[pos + 2 * l]));
                const double u2 = mt_open_uniform(mt_temper(s[pos + 2 * l + 1]));
                const double r = sqrt(-2 * log(u1)), angle = 2 * 3.14159265358979323846 * u2;
                const double re = r * cos(angle), im = r * sin(angle);
                cx<double> &f = a.F[(long)m * a.npairs + j];
                cx<double> v;
                if (a.mode == 0)
                    v = {scale * re, scale * im};
                else
                    v = {decay * f.x + scale * re, decay * f.y + scale * im};
                f = v;
                if (a.noise) {
                    const int p = a.map[j] >= 0 ? a.map[j] : ~a.map[j];
                    a.noise[(long)m * a.npairs + p] =
                        a.map[j] >= 0 ? cx<R>{(R)v.x, (R)v.y} : cx<R>{(R)v.x, (R)(-v.y)};
                }
            }
            pos += 2 * count;
            done += count;
        }
    }
    for (int i = lane; i < MT_N; i += OU_LANES)
        words[i] = s[i];
    if (lane == 0)
        words[MT_N] = (uint64_t)pos;
}

template <typename R> struct ForcingModel {
    const cx<R> *noise;  // [B][npairs] shell noise of the current step, or null
    int npairs;
    const double *amp;   // per member Kolmogorov amplitude
    const int *pair_of;  // (kmax+1) x (2 kmax + 1): canonical pair index or -1
    int kmin, kmax, kf;
};

// Held forcing of mode (c, ky) for member m:
//   amp sin(kf y) -> -i amp/2 at (0, kf), +i amp/2 at (0, -kf);
//   shell pairs kmin <= |k| <= kmax -> the step's noise, with the kx = 0,
//   ky < 0 entries the conjugates of their ky > 0 partners.
template <typename R>
HZL_HD cx<R> model_forcing(const ForcingModel<R> &f, int m, int c, int ky) {
    cx<R> F = {0, 0};
    if (c == 0 && f.kf > 0 && (ky == f.kf || ky == -f.kf))
        F.y = (R)(ky > 0 ? -0.5 * f.amp[m] : 0.5 * f.amp[m]);
    const int k2 = c * c + ky * ky;
    if (f.noise && k2 >= f.kmin * f.kmin && k2 <= f.kmax * f.kmax) {
        const bool flip = c == 0 && ky < 0;
        const int b = flip ? -ky : ky;
        const int p = f.pair_of[c * (2 * f.kmax + 1) + (b + f.kmax)];
        if (p >= 0) {
            const cx<R> v = f.noise[(long)m * f.npairs + p];
            F.x += v.x;
            F.y += flip ? -v.y : v.y;
        }
    }
    return F;
}

// -------------------------------------------------------- column pass ----
template <typename R> struct StageArgs {
    int B; // members
    cx<R> *w0, *w1, *acc;
    const cx<R> *nl;
    cx<R> *mix;
    const cx<R> *tw;
    const double *nu, *alpha; // per member
    ForcingModel<R> forcing;
    const cx<R> *ext;     // host-provided held forcing [.][KP][J] or null
    long long ext_stride; // complex entries between members (0: shared)
    R h_half, h_full, h_sixth;
    // drift outputs (modes 9, 10): further nonlinear mixed fields and the
    // retained-half-plane result
    const cx<R> *nl2 = nullptr, *nl3 = nullptr;
    cx<R> *out = nullptr;
};

// Column pass modes.
//   0      pass A only: mixed fields of w0
//   1..4   RK4 stage MODE (src/solver.cpp:223-234)
//   5..7   SSPRK(3,3) stage MODE - 4 (src/solver.cpp:235-246): u1 -> w1,
//          u2 -> acc, the new state -> w0
//   8      pass A of two states: psi~, (ky psi)~ of w0 and w~, (ky w)~ of w1,
//          so the row pass forms J(w0, w1) = u(w0) . grad(w1)
//   9      out = -FFT_y(nl [+ nl2]) / N^2 - (nu k^2 + alpha) w1 on the mask
//          (drift and its tangent, src/solver.cpp:255-290); no pass A
//   10     out = FFT_y(nl) / N^2 + i ky / k^2 FFT_y(nl2) / N^2
//          - i kx / k^2 FFT_y(nl3) / N^2 - (nu k^2 + alpha) w1 on the mask
//          (the adjoint drift, src/solver.cpp:291-323); no pass A
template <int N, int MODE, typename R, class G = Geometry<N>, class Ctx>
HZL_HD void col_body(const Ctx &ctx, const StageArgs<R> &a, cx<R> *lds, long block) {
    static_assert(G::NN == N, "launch shape of another grid");
    using FFT = GroupFFT<N, G::E, R>;
    constexpr int E = G::E, T = G::T, K = G::K, KP = G::KP, J = G::J;
    const int tid = ctx.tid();
    const int g = tid / T, t = tid % T;
    const long gc = block * G::COL_G + g;
    const bool live = gc < (long)a.B * KP;
    const int m = live ? (int)(gc / KP) : 0, c = live ? (int)(gc % KP) : 0;
    cx<R> *L = lds + g * FFT::LDS_ELEMS;
    const long soff = (long)m * G::SPEC + (long)c * J;
    constexpr bool REG_TW = !(G::TABLE_TW && (MODE == 3 || MODE == 4));
    const TwArg<FFT, R, REG_TW> twa = lane_twiddles<FFT, R, REG_TW>(a.tw, t);
    cx<R> v[E];
    cx<R> vb[MODE == 8 ? E : 1]; // the w-source state of mode 8

    if constexpr (MODE == 0 || MODE == 8) {
#pragma unroll
        for (int idx = 0; idx < E; ++idx) {
            const int k = FFT::index(t, idx);
            const int ky = k < N / 2 ? k : k - N;
            const bool in = !FFT::slot_outside(idx, K) && live && ky >= -K && ky <= K;
            v[idx] = in ? a.w0[soff + ky + K] : cx<R>{0, 0};
            if constexpr (MODE == 8)
                vb[idx] = in ? a.w1[soff + ky + K] : cx<R>{0, 0};
        }
    } else if constexpr (MODE == 9 || MODE == 10) {
        constexpr int NF = MODE == 10 ? 3 : 1;
        const cx<R> *src[3] = {a.nl, a.nl2, a.nl3};
        const R inv_nn = R(1) / (R(N) * R(N));
        const R nu = live ? (R)a.nu[m] : R(0), alpha = live ? (R)a.alpha[m] : R(0);
        cx<R> f[NF][E];
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            const cx<R> *nl = src[q] + (long)m * G::MIXF;
            const cx<R> *nl2 = MODE == 9 && a.nl2 ? a.nl2 + (long)m * G::MIXF : nullptr;
#pragma unroll
            for (int j = 0; j < E; ++j) {
                cx<R> x = live ? nl[tiled(KP, c, t + T * j)] : cx<R>{0, 0};
                if (nl2 && live)
                    x = cadd(x, nl2[tiled(KP, c, t + T * j)]);
                f[q][j] = x;
            }
            FFT::forward(ctx, f[q], L, t, twa);
        }
#pragma unroll
        for (int idx = 0; idx < E; ++idx) {
            if (FFT::slot_outside(idx, K))
                continue;
            const int k = FFT::index(t, idx);
            const int ky = k < N / 2 ? k : k - N;
            if (live && ky >= -K && ky <= K) {
                const long off = soff + ky + K;
                const int k2 = c * c + ky * ky;
                cx<R> r = {0, 0};
                if (k2 != 0) {
                    const cx<R> x = a.w1[off];
                    const R rate = nu * (R)k2 + alpha;
                    const cx<R> n1 = cscale(f[0][idx], inv_nn);
                    if constexpr (MODE == 9) {
                        r = {-n1.x - rate * x.x, -n1.y - rate * x.y};
                    } else {
                        const R q = R(1) / (R)k2;
                        const cx<R> n2 = cscale(f[1][idx], inv_nn);
                        const cx<R> n3 = cscale(f[NF - 1][idx], inv_nn);
                        const R sy = (R)ky * q, sx = (R)c * q;
                        // + i sy n2 - i sx n3
                        r = {n1.x - sy * n2.y + sx * n3.y - rate * x.x,
                             n1.y + sy * n2.x - sx * n3.x - rate * x.y};
                    }
                }
                a.out[off] = r;
            }
        }
        return;
    } else {
        const cx<R> *nl = a.nl + (long)m * G::MIXF;
#pragma unroll
        for (int j = 0; j < E; ++j)
            v[j] = live ? stream_load<G::STREAM>(nl + tiled(KP, c, t + T * j)) : cx<R>{0, 0};
        // All state loads of the column are issued together (one memory round
        // trip for the column, not one per slot).  Stage 2 rebuilds
        // w2 = w_n + dt/2 k1 from acc = k1 exactly as stage 1 formed it, so
        // stage 1 never stores w1.
        cx<R> wn[E], ws[E], sa[E];
        auto load_state = [&]() {
#pragma unroll
            for (int idx = 0; idx < E; ++idx) {
                const int k = FFT::index(t, idx);
                const int ky = k < N / 2 ? k : k - N;
                wn[idx] = ws[idx] = sa[idx] = {0, 0};
                if (!FFT::slot_outside(idx, K) && live && ky >= -K && ky <= K) {
                    const long off = soff + ky + K;
                    wn[idx] = a.w0[off];
                    if constexpr (MODE == 3 || MODE == 4 || MODE == 6)
                        ws[idx] = a.w1[off];
                    if constexpr (MODE == 7)
                        ws[idx] = a.acc[off];
                    if constexpr (MODE >= 2 && MODE <= 4)
                        sa[idx] = a.acc[off];
                }
            }
        };
#ifdef HZL_COL_PREFETCH
        load_state();
#endif
        FFT::forward(ctx, v, L, t, twa);
#ifndef HZL_COL_PREFETCH
        load_state();
#endif
        const R inv_nn = R(1) / (R(N) * R(N));
        const R nu = live ? (R)a.nu[m] : R(0), alpha = live ? (R)a.alpha[m] : R(0);
        if constexpr (MODE == 2) {
#pragma unroll
            for (int idx = 0; idx < E; ++idx)
                ws[idx] = cadd(wn[idx], cscale(sa[idx], a.h_half));
        }
#pragma unroll
        for (int idx = 0; idx < E; ++idx) {
            if (FFT::slot_outside(idx, K)) {
                v[idx] = {0, 0};
                continue;
            }
            const int k = FFT::index(t, idx);
            const int ky = k < N / 2 ? k : k - N;
            cx<R> next = {0, 0};
            if (live && ky >= -K && ky <= K) {
                const long off = soff + ky + K;
                const int k2 = c * c + ky * ky;
                const cx<R> w_n = wn[idx];
                const cx<R> w_s = (MODE == 1 || MODE == 5) ? w_n : ws[idx];
                cx<R> rhs = {0, 0};
                if (k2 != 0) {
                    cx<R> F;
                    if (a.ext)
                        F = a.ext[(long)m * a.ext_stride + (long)c * J + ky + K];
                    else
                        F = model_forcing(a.forcing, m, c, ky);
                    const R rate = nu * (R)k2 + alpha;
                    const cx<R> nh = cscale(v[idx], inv_nn);
                    rhs = {-nh.x - rate * w_s.x + F.x, -nh.y - rate * w_s.y + F.y};
                }
                if constexpr (MODE == 1) {
                    a.acc[off] = rhs;
                    next = cadd(w_n, cscale(rhs, a.h_half));
                } else if constexpr (MODE == 2) {
                    a.acc[off] = cadd(sa[idx], cscale(rhs, R(2)));
                    next = cadd(w_n, cscale(rhs, a.h_half));
                    a.w1[off] = next;
                } else if constexpr (MODE == 3) {
                    a.acc[off] = cadd(sa[idx], cscale(rhs, R(2)));
                    next = cadd(w_n, cscale(rhs, a.h_full));
                    a.w1[off] = next;
                } else if constexpr (MODE == 4) {
                    next = cadd(w_n, cscale(cadd(sa[idx], rhs), a.h_sixth));
                    a.w0[off] = next;
                } else if constexpr (MODE == 5) {
                    next = cadd(w_n, cscale(rhs, a.h_full)); // u1
                    a.w1[off] = next;
                } else if constexpr (MODE == 6) {
                    const cx<R> u1 = ws[idx];
                    next = cadd(cscale(w_n, R(0.75)),
                                cscale(cadd(u1, cscale(rhs, a.h_full)), R(0.25))); // u2
                    a.acc[off] = next;
                } else {
                    const cx<R> u2 = ws[idx];
                    const R third = R(3), two_thirds = (R)(2.0 / 3.0);
                    const cx<R> e = cadd(u2, cscale(rhs, a.h_full));
                    next = {w_n.x / third + two_thirds * e.x, w_n.y / third + two_thirds * e.y};
                    a.w0[off] = next;
                }
            }
            v[idx] = next;
        }
    }

    // pass A: mixed fields psi~, (ky psi)~, w~, (ky w)~ of the next stage state
    cx<R> *mix = a.mix + (long)m * 4 * G::MIXF;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        cx<R> u[E];
#pragma unroll
        for (int idx = 0; idx < E; ++idx) {
            const int k = FFT::index(t, idx);
            const int ky = k < N / 2 ? k : k - N;
            const int k2 = c * c + ky * ky;
            const R inv = k2 > 0 ? recip((R)k2) : R(0);
            R s = R(1);
            if (f == 0)
                s = inv;
            else if (f == 1)
                s = inv * (R)ky;
            else if (f == 3)
                s = (R)ky;
            cx<R> src = v[idx];
            if constexpr (MODE == 8)
                if (f >= 2)
                    src = vb[idx];
            u[idx] = cscale(src, s);
        }
        FFT::inverse(ctx, u, L, t, twa);
        if (live) {
#pragma unroll
            for (int n2 = 0; n2 < E; ++n2)
                stream_store<G::STREAM>(mix + f * G::MIXF + tiled(KP, c, t + T * n2), u[n2]);
        }
    }
}

// ----------------------------------------------------------- row pass ----
template <typename R> struct RowArgs {
    int B;
    const cx<R> *mix;
    cx<R> *nl;
    const cx<R> *tw;
    double *cfl; // per row pair running max of |u|+|v| (CFL variant only)
};

template <class GEO> struct RowShape {
    using G = GEO;
    static constexpr int N = G::NN;
    static constexpr int ROWS = 2 * G::ROW_PPW;      // rows per workgroup
    static constexpr int TILES = ROWS / 16;          // 16-row tiles per workgroup
    static constexpr int PITCH = 17;                 // staged kx-row pitch (bank-conflict free)
    static constexpr int CHUNK = TILES * G::KP * PITCH; // complex entries of one staged field
    template <typename R> static constexpr int lds_elems() {
        constexpr int fft = G::ROW_PPW * GroupFFT<N, G::E, R>::LDS_ELEMS;
        return 2 * CHUNK > fft ? 2 * CHUNK : fft;
    }
};

// Stage field f of the workgroup's rows into dst, layout [tile][kx][PITCH]
// (the global [tile][kx][16] tile layout with padded rows), with 2-entry
// vector loads by consecutive lanes.
// A field's share of the workgroup's 16-row tile held in registers by one
// thread: ITERS pairs of consecutive complex entries.
template <class GEO, typename R> struct RowTile {
    using G = GEO;
    using RS = RowShape<G>;
    static constexpr int N = G::NN;
    struct alignas(4 * sizeof(R)) pair2 {
        cx<R> u, v;
    };
    static constexpr int PER_TILE = G::KP * 16, ELEMS = RS::TILES * PER_TILE;
    static constexpr int STEP = 2 * G::ROW_THREADS, ITERS = (ELEMS + STEP - 1) / STEP;
    pair2 v[ITERS];

    // Issue all loads of field f (global tile layout [tile][kx][16]).
    template <class Ctx> HZL_HD void load(const Ctx &ctx, const RowArgs<R> &a, long row0, int f) {
        const long total_rows = (long)a.B * N;
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int e = 2 * ctx.tid() + it * STEP;
            const int ti = e / PER_TILE, rem = e % PER_TILE;
            const long grow = row0 + 16 * ti;
            if (e < ELEMS && grow < total_rows) {
                const long mm = grow / N;
                const int yy = (int)(grow % N);
                const cx<R> *src = a.mix + (mm * 4 + f) * G::MIXF + (long)(yy >> 4) * PER_TILE + rem;
                v[it].u = stream_load<G::STREAM>(src);
                v[it].v = stream_load<G::STREAM>(src + 1);
            }
        }
    }

    // Write to LDS with padded kx rows: layout [tile][kx][PITCH].
    template <class Ctx> HZL_HD void store(const Ctx &ctx, const RowArgs<R> &a, long row0,
                                           cx<R> *dst) const {
        const long total_rows = (long)a.B * N;
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int e = 2 * ctx.tid() + it * STEP;
            const int ti = e / PER_TILE, rem = e % PER_TILE;
            if (e < ELEMS && row0 + 16 * ti < total_rows) {
                cx<R> *d = dst + ti * (G::KP * RS::PITCH) + (rem >> 4) * RS::PITCH + (rem & 15);
                d[0] = v[it].u;
                d[1] = v[it].v;
            }
        }
    }
};

// Offset of (kx, local row ly) in a staged field.
template <class G> HZL_HD int staged(int kx, int ly) {
    using RS = RowShape<G>;
    return (ly >> 4) * (G::KP * RS::PITCH) + kx * RS::PITCH + (ly & 15);
}

// Entry k of the packed spectrum z = a + i b of two real rows with x spectra
// A and B at x wave number kx = k (pos) or kx = N - k (the conjugate half).
template <typename R> HZL_HD cx<R> pack_entry(cx<R> A, cx<R> Bv, int k, bool pos) {
    if (k == 0) // DC terms of real signals: real parts only
        return {A.x, Bv.x};
    if (pos)
        return {A.x - Bv.y, A.y + Bv.x};
    return {A.x + Bv.y, Bv.x - A.y};
}

// Entry k of the packed row spectrum z1 = u + i wx (FIRST: p = (ky psi)~,
// q = w~) or z2 = v + i wy (p = psi~, q = (ky w)~) at x wave number kx = k
// (pos) or kx = N - k (the conjugate half).
template <typename R, bool FIRST>
HZL_HD cx<R> z_entry(cx<R> p, cx<R> q, int k, int kx, bool pos) {
    const R fx = (R)kx;
    // FIRST: a = u = i P,        b = wx = i kx Q
    // else : a = v = -i kx P,    b = wy = i Q
    const cx<R> A = FIRST ? mul_i(p) : mul_mi(cscale(p, fx));
    const cx<R> Bv = FIRST ? mul_i(cscale(q, fx)) : mul_i(q);
    return pack_entry(A, Bv, k, pos);
}

// One row-pair group: build the packed spectra z1 = u + i wx (fields B, C) or
// z2 = v + i wy (fields A, D) of local row ly from the staged fields P, Q.
template <class G, typename R, bool FIRST>
HZL_HD void row_spectrum(cx<R> (&z)[G::E], const cx<R> *P, const cx<R> *Q, int t,
                         int ly) {
    constexpr int N = G::NN;
    using FFT = GroupFFT<N, G::E, R>;
    constexpr int K = G::K;
#pragma unroll
    for (int idx = 0; idx < G::E; ++idx) {
        const int k = FFT::index(t, idx);
        const bool pos = k <= K, neg = k >= N - K;
        z[idx] = {0, 0};
        if (!FFT::slot_outside(idx, K) && (pos || neg)) {
            const int kx = pos ? k : N - k;
            const int o = staged<G>(kx, ly);
            z[idx] = z_entry<R, FIRST>(P[o], Q[o], k, kx, pos);
        }
    }
}

// The same spectra for both rows y0, y0 + 1 of a pair straight from the tiled
// mixed fields in HBM: (kx, y0) and (kx, y0 + 1) are adjacent, so one 16-byte
// load per field and slot serves both rows.
template <class G, typename R, bool FIRST>
HZL_HD void direct_spectra(cx<R> (&za)[G::E], cx<R> (&zb)[G::E], const cx<R> *P,
                           const cx<R> *Q, int t, int y0) {
    constexpr int N = G::NN;
    using FFT = GroupFFT<N, G::E, R>;
    constexpr int K = G::K;
    struct alignas(4 * sizeof(R)) pair2 {
        cx<R> u, v;
    };
#pragma unroll
    for (int idx = 0; idx < G::E; ++idx) {
        const int k = FFT::index(t, idx);
        const bool pos = k <= K, neg = k >= N - K;
        za[idx] = zb[idx] = {0, 0};
        if (!FFT::slot_outside(idx, K) && (pos || neg)) {
            const int kx = pos ? k : N - k;
            const long o = tiled(G::KP, kx, y0);
            const pair2 p = *reinterpret_cast<const pair2 *>(P + o);
            const pair2 q = *reinterpret_cast<const pair2 *>(Q + o);
            za[idx] = z_entry<R, FIRST>(p.u, q.u, k, kx, pos);
            zb[idx] = z_entry<R, FIRST>(p.v, q.v, k, kx, pos);
        }
    }
}

// Packed forward transform of a row pair's products pa, pb, unpacked to
// kx in [0, K] and stored as the nonlinear mixed field; with CFL, the group's
// maximum of |u| + |v| joins the pair's running maximum.
template <int N, typename R, bool CFL, class G, class Ctx, class TW>
HZL_HD void row_finish(const Ctx &ctx, const RowArgs<R> &a, cx<R> *L, int t, const TW &twa,
                       const R (&pa)[G::E], const R (&pb)[G::E], R vmax, bool live, int m,
                       int y0, long gp) {
    using FFT = GroupFFT<N, G::E, R>;
    constexpr int E = G::E, T = G::T, K = G::K, KP = G::KP;
    cx<R> q[E];
#pragma unroll
    for (int j = 0; j < E; ++j)
        q[j] = {pa[j], pb[j]};
    FFT::forward(ctx, q, L, t, twa);
#pragma unroll
    for (int idx = 0; idx < E; ++idx)
        if (!FFT::slot_outside(idx, K)) // only |kx| <= K is unpacked
            L[FFT::index(t, idx)] = q[idx];
    ctx.wsync();
    cx<R> *nl = a.nl + (long)m * G::MIXF;
#pragma unroll
    for (int idx = 0; idx < E; ++idx) {
        const int k = FFT::index(t, idx);
        if (live && k <= K) {
            const cx<R> qa = L[k], qb = L[(N - k) % N];
            const cx<R> p0 = {R(0.5) * (qa.x + qb.x), R(0.5) * (qa.y - qb.y)};
            const cx<R> p1 = {R(0.5) * (qa.y + qb.y), R(0.5) * (qb.x - qa.x)};
            const long o = tiled(KP, k, y0);
            // ordinary stores: a lane writes 16 bytes of a 128-byte line whose other
            // parts come from the other row pairs, which L2 merges (non-temporal
            // stores measured 1.05x the bytes here)
            nl[o] = p0;
            nl[o + 1] = p1;
        }
    }
    // the group's region is free again only after every lane's unpack reads
    ctx.wsync();
    if constexpr (CFL) {
        R *red = reinterpret_cast<R *>(L);
        red[t] = vmax;
        ctx.wsync();
        if (live && t == 0) {
            R mx = R(0);
            for (int i = 0; i < T; ++i) {
                const R v = red[i];
                mx = (v > mx || v != v) ? v : mx;
            }
            const double prev = a.cfl[gp];
            const double cur = (double)mx;
            a.cfl[gp] = (cur > prev || cur != cur) ? cur : prev;
        }
    }
}

template <int N, typename R, bool CFL, class G = Geometry<N>, class Ctx>
HZL_HD void row_body(const Ctx &ctx, const RowArgs<R> &a, cx<R> *lds, long block) {
    static_assert(G::NN == N, "launch shape of another grid");
    using FFT = GroupFFT<N, G::E, R>;
    using RS = RowShape<G>;
    constexpr int E = G::E, T = G::T, K = G::K, KP = G::KP, PPW = G::ROW_PPW;
    const int tid = ctx.tid();
    const int g = tid / T, t = tid % T; // group-contiguous lanes: conflict-free exchanges
    const long row0 = block * RS::ROWS;
    const long gp = block * PPW + g;
    const bool live = gp < (long)a.B * (N / 2);
    const int m = live ? (int)((2 * gp) / N) : 0, y0 = live ? (int)((2 * gp) % N) : 0;
    const int ly = 2 * g; // local row of y0
    cx<R> *S0 = lds, *S1 = lds + RS::CHUNK;
    cx<R> *L = lds + g * FFT::LDS_ELEMS;
    const TwArg<FFT, R, !G::TABLE_TW> twa = lane_twiddles<FFT, R, !G::TABLE_TW>(a.tw, t);

    R pa[E], pb[E], ua[CFL ? E : 1], ub[CFL ? E : 1];
    R vmax = R(0);

    // u wx from (ky psi)~ and w~; the loads of psi~ and (ky w)~ are issued
    // right after, so they arrive while the first pair is transformed.
    {
        RowTile<G, R> tb, tc;
        tb.load(ctx, a, row0, 1);
        tc.load(ctx, a, row0, 2);
        tb.store(ctx, a, row0, S0);
        tc.store(ctx, a, row0, S1);
    }
    RowTile<G, R> ta, td;
#ifdef HZL_ROW_PREFETCH
    ta.load(ctx, a, row0, 0);
    td.load(ctx, a, row0, 3);
#endif
    ctx.sync();
    {
        cx<R> za[E], zb[E];
        row_spectrum<G, R, true>(za, S0, S1, t, ly);
        row_spectrum<G, R, true>(zb, S0, S1, t, ly + 1);
        ctx.sync();
        FFT::inverse(ctx, za, L, t, twa);
#pragma unroll
        for (int j = 0; j < E; ++j) {
            pa[j] = za[j].x * za[j].y;
            if constexpr (CFL)
                ua[j] = za[j].x < 0 ? -za[j].x : za[j].x;
        }
        FFT::inverse(ctx, zb, L, t, twa);
#pragma unroll
        for (int j = 0; j < E; ++j) {
            pb[j] = zb[j].x * zb[j].y;
            if constexpr (CFL)
                ub[j] = zb[j].x < 0 ? -zb[j].x : zb[j].x;
        }
    }
    // v wy from psi~ and (ky w)~ (the staging area doubles as the groups'
    // exchange areas: every wavefront must be done with them first)
    ctx.sync();
#ifndef HZL_ROW_PREFETCH
    ta.load(ctx, a, row0, 0);
    td.load(ctx, a, row0, 3);
#endif
    ta.store(ctx, a, row0, S0);
    td.store(ctx, a, row0, S1);
    ctx.sync();
    {
        cx<R> za[E], zb[E];
        row_spectrum<G, R, false>(za, S0, S1, t, ly);
        row_spectrum<G, R, false>(zb, S0, S1, t, ly + 1);
        ctx.sync();
        FFT::inverse(ctx, za, L, t, twa);
#pragma unroll
        for (int j = 0; j < E; ++j) {
            pa[j] += za[j].x * za[j].y;
            if constexpr (CFL) {
                R s = ua[j] + (za[j].x < 0 ? -za[j].x : za[j].x);
                if (!(s <= R(1e30)))
                    s = (R)__builtin_huge_valf(); // non-finite velocity poisons the maximum
                vmax = s > vmax ? s : vmax;
            }
        }
        FFT::inverse(ctx, zb, L, t, twa);
#pragma unroll
        for (int j = 0; j < E; ++j) {
            pb[j] += zb[j].x * zb[j].y;
            if constexpr (CFL) {
                R s = ub[j] + (zb[j].x < 0 ? -zb[j].x : zb[j].x);
                if (!(s <= R(1e30)))
                    s = (R)__builtin_huge_valf();
                vmax = s > vmax ? s : vmax;
            }
        }
    }
    row_finish<N, R, CFL, G>(ctx, a, L, t, twa, pa, pb, vmax, live, m, y0, gp);
}

// Row pass without staging: each row-pair group loads its spectra straight
// from HBM (direct_spectra), so the workgroup needs no barrier and LDS only
// for the groups' transform exchanges.  Same arithmetic as row_body.
template <int N, typename R, bool CFL, class G, class Ctx>
HZL_HD void row_body_direct(const Ctx &ctx, const RowArgs<R> &a, cx<R> *lds, long block) {
    static_assert(G::NN == N, "launch shape of another grid");
    using FFT = GroupFFT<N, G::E, R>;
    constexpr int E = G::E, T = G::T, PPW = G::ROW_PPW;
    const int tid = ctx.tid();
    const int g = tid / T, t = tid % T;
    const long gp = block * PPW + g;
    const bool live = gp < (long)a.B * (N / 2);
    const int m = live ? (int)((2 * gp) / N) : 0, y0 = live ? (int)((2 * gp) % N) : 0;
    cx<R> *L = lds + g * FFT::LDS_ELEMS;
    const TwArg<FFT, R> twa = lane_twiddles<FFT>(a.tw, t);
    const cx<R> *mix = a.mix + (long)m * 4 * G::MIXF;
    R pa[E], pb[E], ua[CFL ? E : 1], ub[CFL ? E : 1];
    R vmax = R(0);
    {
        cx<R> za[E], zb[E];
        direct_spectra<G, R, true>(za, zb, mix + G::MIXF, mix + 2 * G::MIXF, t, y0);
        FFT::inverse(ctx, za, L, t, twa);
#pragma unroll
        for (int j = 0; j < E; ++j) {
            pa[j] = za[j].x * za[j].y;
            if constexpr (CFL)
                ua[j] = za[j].x < 0 ? -za[j].x : za[j].x;
        }
        FFT::inverse(ctx, zb, L, t, twa);
#pragma unroll
        for (int j = 0; j < E; ++j) {
            pb[j] = zb[j].x * zb[j].y;
            if constexpr (CFL)
                ub[j] = zb[j].x < 0 ? -zb[j].x : zb[j].x;
        }
    }
    {
        cx<R> za[E], zb[E];
        direct_spectra<G, R, false>(za, zb, mix, mix + 3 * G::MIXF, t, y0);
        FFT::inverse(ctx, za, L, t, twa);
#pragma unroll
        for (int j = 0; j < E; ++j) {
            pa[j] += za[j].x * za[j].y;
            if constexpr (CFL) {
                R s = ua[j] + (za[j].x < 0 ? -za[j].x : za[j].x);
                if (!(s <= R(1e30)))
                    s = (R)__builtin_huge_valf();
                vmax = s > vmax ? s : vmax;
            }
        }
        FFT::inverse(ctx, zb, L, t, twa);
#pragma unroll
        for (int j = 0; j < E; ++j) {
            pb[j] += zb[j].x * zb[j].y;
            if constexpr (CFL) {
                R s = ub[j] + (zb[j].x < 0 ? -zb[j].x : zb[j].x);
                if (!(s <= R(1e30)))
                    s = (R)__builtin_huge_valf();
                vmax = s > vmax ? s : vmax;
            }
        }
    }
    row_finish<N, R, CFL, G>(ctx, a, L, t, twa, pa, pb, vmax, live, m, y0, gp);
}

// Products of the adjoint drift (src/solver.cpp:291-323).  For each row pair:
// the physical cotangent lam (field 2 of mixA) and the physical gradient
// (wx, wy) of the state (fields 2, 3 of mixB, the state's w~ and (ky w)~),
// then lam wx and lam wy, forward x-transformed into the mixed fields out2,
// out3.  Spectra are loaded straight from HBM, one row at a time.
template <typename R> struct ProductArgs {
    int B;
    const cx<R> *mixA, *mixB; // [member][4][mixed field]
    cx<R> *out2, *out3;       // [member][mixed field]
    const cx<R> *tw;
};

template <int N, typename R, class G, class Ctx>
HZL_HD void row_body_products(const Ctx &ctx, const ProductArgs<R> &a, cx<R> *lds, long block) {
    static_assert(G::NN == N, "launch shape of another grid");
    using FFT = GroupFFT<N, G::E, R>;
    constexpr int E = G::E, T = G::T, K = G::K, KP = G::KP, PPW = G::ROW_PPW;
    const int tid = ctx.tid();
    const int g = tid / T, t = tid % T;
    const long gp = block * PPW + g;
    const bool live = gp < (long)a.B * (N / 2);
    const int m = live ? (int)((2 * gp) / N) : 0, y0 = live ? (int)((2 * gp) % N) : 0;
    cx<R> *L = lds + g * FFT::LDS_ELEMS;
    const TwArg<FFT, R> twa = lane_twiddles<FFT>(a.tw, t);
    const cx<R> *lam = a.mixA + (long)m * 4 * G::MIXF + 2 * G::MIXF;
    const cx<R> *w = a.mixB + (long)m * 4 * G::MIXF + 2 * G::MIXF, *kw = w + G::MIXF;
    R p2[2][E], p3[2][E];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        // zA = lam + i wx, zB = wy (+ i 0) of row y0 + r
        cx<R> zA[E], zB[E];
#pragma unroll
        for (int idx = 0; idx < E; ++idx) {
            const int k = FFT::index(t, idx);
            const bool pos = k <= K, neg = k >= N - K;
            zA[idx] = zB[idx] = {0, 0};
            if (!FFT::slot_outside(idx, K) && (pos || neg)) {
                const int kx = pos ? k : N - k;
                const long o = tiled(KP, kx, y0 + r);
                const cx<R> wx = mul_i(cscale(w[o], (R)kx)), wy = mul_i(kw[o]);
                zA[idx] = pack_entry(lam[o], wx, k, pos);
                zB[idx] = pack_entry(wy, cx<R>{0, 0}, k, pos);
            }
        }
        FFT::inverse(ctx, zA, L, t, twa);
        FFT::inverse(ctx, zB, L, t, twa);
#pragma unroll
        for (int j = 0; j < E; ++j) {
            p2[r][j] = zA[j].x * zA[j].y;
            p3[r][j] = zA[j].x * zB[j].x;
        }
    }
    const RowArgs<R> r2{a.B, nullptr, a.out2, a.tw, nullptr};
    const RowArgs<R> r3{a.B, nullptr, a.out3, a.tw, nullptr};
    row_finish<N, R, false, G>(ctx, r2, L, t, twa, p2[0], p2[1], R(0), live, m, y0, gp);
    row_finish<N, R, false, G>(ctx, r3, L, t, twa, p3[0], p3[1], R(0), live, m, y0, gp);
}

// ---------------------------------------------------------------- GRF ----
// Gaussian random field on canonical retained modes (kx > 0, or kx = 0 and
// ky > 0), enumerated kx-major then ky: amplitude a(k) = k^{5/2} exp(-k^2/32),
// value a(k) (g0 + i g1)/sqrt(2), normalised so (A/2) sum_full |w|^2 = 1.
HZL_HD int grf_canonical_index(int K, int c, int ky) {
    const int J = 2 * K + 1;
    if (c == 0)
        return ky > 0 ? ky - 1 : -1;
    return K + (c - 1) * J + (ky + K);
}

HZL_HD cx<double> grf_mode(uint64_t key, int K, int c, int ky) {
    int b = ky;
    bool flip = false;
    if (c == 0 && b < 0) {
        b = -b;
        flip = true;
    }
    const int q = grf_canonical_index(K, c, b);
    if (q < 0)
        return {0.0, 0.0};
    const double k2 = (double)(c * c + b * b), k = sqrt(k2);
    const double amp = pow(k, 2.5) * exp(-k2 / 32.0) / sqrt(2.0);
    double g0, g1;
    philox_normal2(key, (uint32_t)q, 0, TAG_GRF, g0, g1);
    return {amp * g0, flip ? -amp * g1 : amp * g1};
}

template <typename R> struct GrfArgs {
    int K;
    cx<R> *w0; // [B][KP][J]
    const uint64_t *keys;
};

// One workgroup of `nt` threads per member; `red` holds nt doubles.
template <typename R, class Ctx>
HZL_HD void grf_body(const Ctx &ctx, const GrfArgs<R> &a, double *red, int nt, long member) {
    const int K = a.K, KP = K + 1, J = 2 * K + 1;
    const long total = (long)KP * J;
    const int tid = ctx.tid();
    const uint64_t key = a.keys[member];
    double s = 0;
    for (long e = tid; e < total; e += nt) {
        const int c = (int)(e / J), ky = (int)(e % J) - K;
        if (c > 0 || ky > 0) {
            const cx<double> w = grf_mode(key, K, c, ky);
            s += w.x * w.x + w.y * w.y;
        }
    }
    red[tid] = s;
    ctx.sync();
    for (int w = nt / 2; w > 0; w /= 2) {
        if (tid < w)
            red[tid] += red[tid + w];
        ctx.sync();
    }
    const double scale = 1.0 / sqrt(4.0 * PI * PI * red[0]);
    cx<R> *out = a.w0 + member * total;
    for (long e = tid; e < total; e += nt) {
        const int c = (int)(e / J), ky = (int)(e % J) - K;
        const cx<double> w = grf_mode(key, K, c, ky);
        out[e] = {(R)(w.x * scale), (R)(w.y * scale)};
    }
    ctx.sync();
}

// -------------------------------------------------------- diagnostics ----
template <typename R> struct DiagArgs {
    int K;
    const cx<R> *w0;
    const double *nu;
    ForcingModel<R> forcing; // forcing of the last step (injection)
    int with_forcing;
    const double *cfl; // per pair maxima or null
    int pairs;        // N/2
    int n;
    double *out; // [B][6]: energy, enstrophy, dissipation, injection, cfl_rate, unused
};

// The reference NormalStream (src/solver.cpp:41-54) continued on the device
// by one workgroup of nt lanes: `count` normals of the stream whose state is
// mt[0..311] (mt19937_64 words), mt[312] (position) and sp = {has_spare,
// spare}, written to out[0..count) in stream order, the state advanced.
// Normals come in Box-Muller pairs (r cos, r sin) from two consecutive outputs;
// a stream that holds a kept r sin returns it first, and an odd remainder
// keeps the last pair's r sin.  Every pair takes two outputs, so the position
// stays even and a pair never straddles a block of 312; the block's pairs are
// drawn in parallel (tempering only reads the words) and the block is
// regenerated in two parallel phases, words 0..155 from old words, then
// 156..311 from the new words 0..155, the sequential recurrence's order.
// s: 312 words of workgroup scratch.
template <class Ctx>
HZL_HD void normal_stream_body(const Ctx &ctx, uint64_t *mt, double *sp, long count, double *out,
                               uint64_t *s, int nt) {
    const int lane = ctx.tid();
    if (count <= 0)
        return;
    for (int i = lane; i < MT_N; i += nt)
        s[i] = mt[i];
    int pos = (int)mt[MT_N];
    const long first = sp[0] != 0 ? 1 : 0; // the kept r sin goes first
    if (first && lane == 0)
        out[0] = sp[1];
    const long rest = count - first, pairs = (rest + 1) / 2;
    ctx.sync();
    long done = 0;
    while (done < pairs) {
        if (pos >= MT_N) {
            uint64_t nv[(MT_N / 2 + 31) / 32];
            int c = 0;
            for (int i = lane; i < MT_N / 2; i += nt)
                nv[c++] = s[i + MT_N / 2] ^ mt_mix(s[i], s[i + 1]);
            ctx.sync();
            c = 0;
            for (int i = lane; i < MT_N / 2; i += nt)
                s[i] = nv[c++];
            ctx.sync();
            c = 0;
            for (int i = MT_N / 2 + lane; i < MT_N; i += nt)
                nv[c++] = s[i - MT_N / 2] ^ mt_mix(s[i], s[(i + 1) % MT_N]);
            ctx.sync();
            c = 0;
            for (int i = MT_N / 2 + lane; i < MT_N; i += nt)
                s[i] = nv[c++];
            ctx.sync();
            pos = 0;
        }
        const long avail = (MT_N - pos) / 2, take = pairs - done < avail ? pairs - done : avail;
        for (long p = lane; p < take; p += nt) {
            const double u1 = mt_open_uniform(mt_temper(s[pos + 2 * p]));
            const double u2 = mt_open_uniform(mt_temper(s[pos + 2 * p + 1]));
            const double r = sqrt(-2 * log(u1)), angle = 2 * 3.14159265358979323846 * u2;
            const long k = first + 2 * (done + p);
            out[k] = r * cos(angle);
            if (k + 1 < count)
                out[k + 1] = r * sin(angle);
            else
                sp[1] = r * sin(angle); // odd remainder: kept for the next call
        }
        pos += 2 * (int)take;
        done += take;
    }
    ctx.sync();
    for (int i = lane; i < MT_N; i += nt)
        mt[i] = s[i];
    if (lane == 0) {
        mt[MT_N] = (uint64_t)pos;
        sp[0] = (rest % 2 == 1) ? 1.0 : 0.0;
    }
}

// floor(sqrt(x)) and ceil(sqrt(x)) of a nonnegative integer, exact.
HZL_HD long isqrt_floor(long x) {
    long r = (long)sqrt((double)x);
    while (r * r > x)
        --r;
    while ((r + 1) * (r + 1) <= x)
        ++r;
    return r;
}
HZL_HD long isqrt_ceil(long x) {
    const long r = isqrt_floor(x);
    return r * r == x ? r : r + 1;
}

// Shell energy spectrum of one member (Solver::energy_spectrum,
// src/solver.cpp:426-433): bin s = floor(|k| + 1/2) of n bins collects
// 2 pi^2 |w_k|^2 / k^2 over the full plane, i.e. over the retained half plane
// with kx > 0 counted twice.  For integer |k|^2 the bin test is exact in
// integers: k lies in bin s >= 1 iff s^2 - s + 1 <= |k|^2 <= s^2 + s.  Lane t
// owns bins t, t + nt, ... and walks each annulus column by column in a fixed
// order, so the sums are the same on every run.
template <typename R, class Ctx>
HZL_HD void spectrum_body(const Ctx &ctx, int K, int n, const cx<R> *w, double *out, int nt) {
    const int J = 2 * K + 1;
    const double two_pi2 = 2 * 3.14159265358979323846 * 3.14159265358979323846;
    for (int s = ctx.tid(); s < n; s += nt) {
        double acc = 0;
        const long lo = (long)s * s - s + 1, hi = (long)s * s + s;
        for (int c = 0; s > 0 && c <= K && (long)c * c <= hi; ++c) {
            const long c2 = (long)c * c;
            const long a = lo - c2 <= 0 ? 0 : isqrt_ceil(lo - c2);
            const long b = isqrt_floor(hi - c2) < K ? isqrt_floor(hi - c2) : K;
            const double mult = c > 0 ? 2.0 : 1.0;
            for (long m = a; m <= b; ++m)
                for (int sign = 1; sign >= (m == 0 ? 1 : -1); sign -= 2) {
                    const long ky = sign * m, k2 = c2 + ky * ky;
                    const cx<R> v = w[(long)c * J + ky + K];
                    const double re = (double)v.x, im = (double)v.y;
                    acc += mult * two_pi2 * (re * re + im * im) / (double)k2;
                }
        }
        out[s] = acc;
    }
}

// Energy, enstrophy, dissipation and injection in FP64 over the half plane
// (kx > 0 entries count twice, the stored kx = 0 column once), reference
// definitions src/solver.cpp:381-397.  `red` holds 3*nt doubles.
template <typename R, class Ctx>
HZL_HD void diag_body(const Ctx &ctx, const DiagArgs<R> &a, double *red, int nt, long member) {
    const int K = a.K, KP = K + 1, J = 2 * K + 1;
    const long total = (long)KP * J;
    const int tid = ctx.tid();
    const cx<R> *w = a.w0 + member * total;
    double e = 0, z = 0, inj = 0;
    for (long i = tid; i < total; i += nt) {
        const int c = (int)(i / J), ky = (int)(i % J) - K;
        const int k2 = c * c + ky * ky;
        if (k2 == 0)
            continue;
        const double mult = c > 0 ? 2.0 : 1.0;
        const double wr = (double)w[i].x, wi = (double)w[i].y;
        const double norm = wr * wr + wi * wi;
        e += mult * norm / k2;
        z += mult * norm;
        if (a.with_forcing) {
            const cx<R> F = model_forcing(a.forcing, (int)member, c, ky);
            inj += mult * (wr * (double)F.x + wi * (double)F.y) / k2;
        }
    }
    red[tid] = e;
    red[nt + tid] = z;
    red[2 * nt + tid] = inj;
    ctx.sync();
    for (int s = nt / 2; s > 0; s /= 2) {
        if (tid < s) {
            red[tid] += red[tid + s];
            red[nt + tid] += red[nt + tid + s];
            red[2 * nt + tid] += red[2 * nt + tid + s];
        }
        ctx.sync();
    }
    if (tid == 0) {
        const double area = 4 * PI * PI;
        double *o = a.out + member * 6;
        o[0] = area / 2 * red[0];
        o[1] = area / 2 * red[nt];
        o[2] = 2 * a.nu[member] * o[1];
        o[3] = area * red[2 * nt];
        double mx = 0;
        if (a.cfl)
            for (int i = 0; i < a.pairs; ++i) {
                const double v = a.cfl[member * a.pairs + i];
                mx = (v > mx || v != v) ? v : mx;
            }
        o[4] = mx * a.n / (2 * PI);
        o[5] = 0;
    }
    ctx.sync();
}

} // namespace hzl_amd

namespace hzl_amd {

// ------------------------------------------------------ particle filter ----
// Bootstrap particle filter pieces (reference src/filter.cpp:38-126), FP64.
template <typename R> struct PfArgs {
    int B, N, K;
    const cx<R> *mix;      // mixed fields of the current states (field 2 = w~)
    int nobs;
    const int *ox, *oy;    // observation grid indices
    const double *ov, *osd;
    const double *cs, *sn; // cos / sin (2 pi j / N), j < N
    double *wstate;        // running weights (1/B after resampling)
    double *wnorm;         // normalised weights of this group
    double *logw;
    double *pred;          // [B][nobs] predictions (optional)
};

// Point value of member m's physical vorticity at grid (x, y) from the mixed
// field: w(x, y) = Re w~(0, y) + 2 sum_{kx=1..K} Re(w~(kx, y) e^{2 pi i kx x / N}).
// N is a power of two and 0 <= x < N (the observations are checked when
// staged), so the table index kx x mod N is a mask of kx x < N^2.
template <typename R>
HZL_HD double point_value(const PfArgs<R> &a, const cx<R> *w, int x, int y) {
    const int KP = a.K + 1, mask = a.N - 1;
    double s = 0;
    // unrolled so that several terms' loads are in flight at once; the sum
    // keeps its order
#pragma unroll 8
    for (int kx = 0; kx <= a.K; ++kx) {
        const cx<R> c = w[tiled(KP, kx, y)];
        const int j = (kx * x) & mask;
        const double t = (double)c.x * a.cs[j] - (double)c.y * a.sn[j];
        s += kx == 0 ? t : 2 * t;
    }
    return s;
}

// One workgroup of nt threads per particle m: Gaussian log-weight increment
// log w_prev - sum_o [r^2/2 + log sd + log(2 pi)/2] (src/filter.cpp:66-74).
template <typename R, class Ctx>
HZL_HD void pf_loglik_body(const Ctx &ctx, const PfArgs<R> &a, double *red, int nt, long m) {
    const int tid = ctx.tid();
    const cx<R> *w = a.mix + m * 4 * (long)(a.K + 1) * a.N + 2 * (long)(a.K + 1) * a.N;
    const double half_log_2pi = 0.5 * log(2 * PI);
    double acc = 0;
    for (int o = tid; o < a.nobs; o += nt) {
        const double v = point_value(a, w, a.ox[o], a.oy[o]);
        if (a.pred)
            a.pred[m * a.nobs + o] = v;
        const double r = (v - a.ov[o]) / a.osd[o];
        acc += 0.5 * r * r + log(a.osd[o]) + half_log_2pi;
    }
    red[tid] = acc;
    ctx.sync();
    for (int s = nt / 2; s > 0; s /= 2) {
        if (tid < s)
            red[tid] += red[tid + s];
        ctx.sync();
    }
    if (tid == 0) {
        const double wp = a.wstate[m];
        a.logw[m] = (wp > 0 ? log(wp) : -__builtin_huge_val()) - red[0];
    }
    ctx.sync();
}

struct PfControl {
    int P;                   // particles per filter; filter f owns members [f P, (f+1) P)
    const double *logw;
    double *wstate, *wnorm;
    double *stats;           // per filter, 8 each: [0] increment [1] ESS [2] max weight
                             // [3] resampled [4] total log-lik [5] resampling events
                             // [6] uniforms used
    int *ancestors;          // global member indices
    const double *uniforms;  // per filter, max_uniforms pre-drawn resampling uniforms
    int max_uniforms;
    uint64_t *keys;          // member Philox keys (children are reseeded)
    const uint64_t *seeds;   // per filter seed
    uint64_t group;
    int is_last;
    const uint64_t *cycle = nullptr; // device (group, is_last) replacing the two
                                     // fields above (graph replays)
};

// Particles per filter whose resampling sums fit in a normalising
// workgroup's LDS (32 KB).
constexpr int kPfLocalWeights = 4096;

// One workgroup per filter: log-sum-exp normalisation, ESS, systematic
// resampling below P/2 unless this is the last group (src/filter.cpp:76-120).
// cw: workgroup-local room for `cap` doubles; when P <= cap the weights, their
// running sums and the ancestor search stay there (the running sum is one
// thread's chain of P dependent additions, in the reference's order, so its
// operands should not wait on device memory), else in wstate.
template <class Ctx>
HZL_HD void pf_normalize_body(const Ctx &ctx, const PfControl &a, double *red, int nt, int f,
                              double *cw = nullptr, int cap = 0) {
    const int tid = ctx.tid(), P = a.P;
    const long base = (long)f * P;
    const double *logw = a.logw + base;
    double *wnorm = a.wnorm + base, *wstate = a.wstate + base, *stats = a.stats + 8L * f;
    int *ancestors = a.ancestors + base;
    uint64_t *keys = a.keys + base;
    const uint64_t group = a.cycle ? a.cycle[0] : a.group;
    const bool is_last = a.cycle ? a.cycle[1] != 0 : a.is_last != 0;
    const double *uniforms = a.uniforms + (long)f * a.max_uniforms;
    double mx = -__builtin_huge_val();
    for (int i = tid; i < P; i += nt)
        mx = logw[i] > mx ? logw[i] : mx;
    red[tid] = mx;
    ctx.sync();
    for (int s = nt / 2; s > 0; s /= 2) {
        if (tid < s)
            red[tid] = red[tid + s] > red[tid] ? red[tid + s] : red[tid];
        ctx.sync();
    }
    mx = red[0];
    ctx.sync();
    const bool finite = mx > -__builtin_huge_val() && mx < __builtin_huge_val();
    double tot = 0;
    if (finite)
        for (int i = tid; i < P; i += nt)
            tot += exp(logw[i] - mx);
    red[tid] = tot;
    ctx.sync();
    for (int s = nt / 2; s > 0; s /= 2) {
        if (tid < s)
            red[tid] += red[tid + s];
        ctx.sync();
    }
    const double inc = finite ? mx + log(red[0]) : -__builtin_huge_val();
    ctx.sync();
    double *const local = cw && P <= cap ? cw : nullptr;
    double sq = 0, mw = 0;
    for (int i = tid; i < P; i += nt) {
        const double w = finite ? exp(logw[i] - inc) : 0.0;
        wnorm[i] = w;
        if (local)
            local[i] = w;
        sq += w * w;
        mw = w > mw ? w : mw;
    }
    red[tid] = sq;
    red[nt + tid] = mw;
    ctx.sync();
    for (int s = nt / 2; s > 0; s /= 2) {
        if (tid < s) {
            red[tid] += red[tid + s];
            red[nt + tid] = red[nt + tid + s] > red[nt + tid] ? red[nt + tid + s] : red[nt + tid];
        }
        ctx.sync();
    }
    const double ess = 1 / red[0];
    const bool resample = finite && ess < (double)P / 2 && !is_last &&
                          (int)stats[6] < a.max_uniforms;
    ctx.sync();
    if (tid == 0) {
        stats[0] = inc;
        stats[1] = ess;
        stats[2] = red[nt];
        stats[3] = resample ? 1 : 0;
        stats[4] += inc;
        if (resample) {
            // one uniform per resampling event; the running sums of the
            // reference's scan (cum_0 = w_0, cum_k = cum_{k-1} + w_k), kept in
            // wstate until it is rewritten below
            red[0] = uniforms[(int)stats[6]];
            stats[6] += 1;
            stats[5] += 1;
            // unrolled so that the loads of several weights are in flight
            // ahead of the chain of additions (whose order is the reference's)
            double cum = 0;
            if (local) {
#pragma unroll 8
                for (int i = 0; i < P; ++i) {
                    cum += local[i];
                    local[i] = cum;
                }
            } else {
#pragma unroll 8
                for (int i = 0; i < P; ++i) {
                    cum += wnorm[i];
                    wstate[i] = cum;
                }
            }
        }
    }
    ctx.sync();
    if (resample) {
        // child i descends from the first particle whose running sum reaches
        // u/P + i/P: the strict `position > cumulative` scan of
        // src/filter.cpp:102-120, by binary search
        const double start = red[0] / (double)P;
        const double *sums = local ? local : wstate;
        for (int i = tid; i < P; i += nt) {
            const double pos = start + (double)i / (double)P;
            int lo = 0, hi = P - 1;
            while (lo < hi) {
                const int mid = (lo + hi) / 2;
                if (pos > sums[mid])
                    lo = mid + 1;
                else
                    hi = mid;
            }
            ancestors[i] = (int)base + lo;
        }
    }
    ctx.sync();
    for (int i = tid; i < P; i += nt) {
        wstate[i] = resample ? 1.0 / (double)P : wnorm[i];
        if (resample)
            keys[i] = stream_seed(a.seeds[f], (group + 1) * (uint64_t)P + (uint64_t)i);
        else
            ancestors[i] = (int)base + i;
    }
    ctx.sync();
}

// Weighted mean of the retained states with the group's normalised weights
// (src/filter.cpp:94-100), in two launches: sums over the particles
// [c per, (c + 1) per) of chunk c in order, then the chunks in order (the
// reference's running sum regrouped: results agree to FP64 rounding).  part
// holds 2 doubles per (chunk, mode).  per: 32, or more when that would give
// more than 65535 chunks (the grid's second dimension).
HZL_HD int pf_mean_chunk(long B) {
    const long per = (B + 65534) / 65535;
    return per > 32 ? (int)per : 32;
}

template <typename R>
HZL_HD void pf_mean_partial_body(int B, long spec, int per, const cx<R> *w0, const double *wn,
                                 double *part, long e, int c) {
    if (e >= spec)
        return;
    const int lo = c * per, hi = lo + per < B ? lo + per : B;
    double re = 0, im = 0;
#pragma unroll 8
    for (int m = lo; m < hi; ++m) {
        const cx<R> v = w0[(long)m * spec + e];
        re += wn[m] * (double)v.x;
        im += wn[m] * (double)v.y;
    }
    part[2 * ((long)c * spec + e)] = re;
    part[2 * ((long)c * spec + e) + 1] = im;
}

HZL_HD void pf_mean_finish_body(long spec, int chunks, const double *part, double *mean, long e) {
    if (e >= spec)
        return;
    double re = 0, im = 0;
#pragma unroll 8
    for (int c = 0; c < chunks; ++c) {
        re += part[2 * ((long)c * spec + e)];
        im += part[2 * ((long)c * spec + e) + 1];
    }
    mean[2 * e] = re;
    mean[2 * e + 1] = im;
}

} // namespace hzl_amd
