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

namespace hzl_amd {

// ------------------------------------------------------------- shapes ----
template <int N> struct Shape;
template <> struct Shape<16> {
    static constexpr int E = 4;
};
template <> struct Shape<32> {
    static constexpr int E = 8;
};
template <> struct Shape<64> {
    static constexpr int E = 16;
};
template <> struct Shape<128> {
    static constexpr int E = 16;
};
template <> struct Shape<256> {
    static constexpr int E = 16;
};
template <> struct Shape<512> {
    static constexpr int E = 32;
};

template <int N> struct Geometry {
    static constexpr int E = Shape<N>::E;
    static constexpr int T = N / E;
    static constexpr int K = (N - 1) / 3;
    static constexpr int KP = K + 1;
    static constexpr int J = 2 * K + 1;
    static constexpr int TILES = N / 16;
    // column kernel: one 64-lane workgroup, G columns
    static constexpr int COL_THREADS = 64;
    static constexpr int COL_G = COL_THREADS / T;
    // row kernel: PPW row pairs per workgroup
    static constexpr int ROW_PPW = (64 / T) > 8 ? (64 / T) : 8;
    static constexpr int ROW_THREADS = ROW_PPW * T;
    static constexpr long SPEC = (long)KP * J; // complex entries per member state
    static constexpr long MIXF = (long)KP * N; // complex entries per mixed field
};

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
struct ForcingModel {
    const double *sigma;  // per member white-noise amplitude
    const double *amp;    // per member Kolmogorov amplitude
    const uint64_t *keys; // per member Philox key
    const int *pair_of;   // (kmax+1) x (2 kmax + 1): canonical pair index or -1
    int kmin, kmax, kf;
    double dt;
};

// Held forcing of mode (c, ky) for member m at step `step`:
//   amp sin(kf y) -> -i amp/2 at (0, kf), +i amp/2 at (0, -kf);
//   shell pairs kmin <= |k| <= kmax -> sigma/sqrt(2 dt) (g0 + i g1), with the
//   kx = 0, ky < 0 entries the conjugates of their ky > 0 partners.
HZL_HD cx<double> model_forcing(const ForcingModel &f, int m, int c, int ky, uint64_t step) {
    cx<double> F = {0.0, 0.0};
    if (c == 0 && f.kf > 0 && (ky == f.kf || ky == -f.kf))
        F.y += ky > 0 ? -0.5 * f.amp[m] : 0.5 * f.amp[m];
    const int k2 = c * c + ky * ky;
    const double s = f.sigma[m];
    if (s > 0 && k2 >= f.kmin * f.kmin && k2 <= f.kmax * f.kmax) {
        int b = ky;
        bool flip = false;
        if (c == 0 && b < 0) {
            b = -b;
            flip = true;
        }
        const int p = f.pair_of[c * (2 * f.kmax + 1) + (b + f.kmax)];
        if (p >= 0) {
            double g0, g1;
            philox_normal2(f.keys[m], (uint32_t)p, step, TAG_NOISE, g0, g1);
            const double scale = s / sqrt(2.0 * f.dt);
            F.x += scale * g0;
            F.y += flip ? -scale * g1 : scale * g1;
        }
    }
    return F;
}

// -------------------------------------------------------- column pass ----
template <typename R> struct StageArgs {
    int B;    // members
    int mode; // 0: pass A only (from w0); 1..4: RK4 stage
    cx<R> *w0, *w1, *acc;
    const cx<R> *nl;
    cx<R> *mix;
    const cx<R> *tw;
    const double *nu, *alpha; // per member
    ForcingModel forcing;
    const cx<R> *ext;     // host-provided held forcing [.][KP][J] or null
    long long ext_stride; // complex entries between members (0: shared)
    R h_half, h_full, h_sixth;
    uint64_t step;
};

template <int N, typename R, class Ctx>
HZL_HD void col_body(const Ctx &ctx, const StageArgs<R> &a, cx<R> *lds, long block) {
    using G = Geometry<N>;
    using FFT = GroupFFT<N, G::E, R>;
    constexpr int E = G::E, T = G::T, K = G::K, KP = G::KP, J = G::J;
    const int tid = ctx.tid();
    const int g = tid / T, t = tid % T;
    const long gc = block * G::COL_G + g;
    const bool live = gc < (long)a.B * KP;
    const int m = live ? (int)(gc / KP) : 0, c = live ? (int)(gc % KP) : 0;
    cx<R> *L = lds + g * FFT::LDS_ELEMS;
    const long soff = (long)m * G::SPEC + (long)c * J;
    cx<R> v[E];

    if (a.mode == 0) {
#pragma unroll
        for (int idx = 0; idx < E; ++idx) {
            const int k = FFT::index(t, idx);
            const int ky = k < N / 2 ? k : k - N;
            v[idx] = (live && ky >= -K && ky <= K) ? a.w0[soff + ky + K] : cx<R>{0, 0};
        }
    } else {
        const cx<R> *nl = a.nl + (long)m * G::MIXF;
#pragma unroll
        for (int j = 0; j < E; ++j)
            v[j] = live ? nl[tiled(KP, c, t + T * j)] : cx<R>{0, 0};
        FFT::forward(ctx, v, L, t, a.tw);
        const R inv_nn = R(1) / (R(N) * R(N));
        const R nu = live ? (R)a.nu[m] : R(0), alpha = live ? (R)a.alpha[m] : R(0);
#pragma unroll
        for (int idx = 0; idx < E; ++idx) {
            const int k = FFT::index(t, idx);
            const int ky = k < N / 2 ? k : k - N;
            cx<R> next = {0, 0};
            if (live && ky >= -K && ky <= K) {
                const long off = soff + ky + K;
                const int k2 = c * c + ky * ky;
                const cx<R> wn = a.w0[off];
                const cx<R> ws = a.mode == 1 ? wn : a.w1[off];
                cx<R> rhs = {0, 0};
                if (k2 != 0) {
                    cx<R> F;
                    if (a.ext) {
                        F = a.ext[(long)m * a.ext_stride + (long)c * J + ky + K];
                    } else {
                        const cx<double> f = model_forcing(a.forcing, m, c, ky, a.step);
                        F = {(R)f.x, (R)f.y};
                    }
                    const R rate = nu * (R)k2 + alpha;
                    const cx<R> nh = cscale(v[idx], inv_nn);
                    rhs = {-nh.x - rate * ws.x + F.x, -nh.y - rate * ws.y + F.y};
                }
                if (a.mode == 1) {
                    a.acc[off] = rhs;
                    next = cadd(wn, cscale(rhs, a.h_half));
                    a.w1[off] = next;
                } else if (a.mode == 2) {
                    const cx<R> s = a.acc[off];
                    a.acc[off] = cadd(s, cscale(rhs, R(2)));
                    next = cadd(wn, cscale(rhs, a.h_half));
                    a.w1[off] = next;
                } else if (a.mode == 3) {
                    const cx<R> s = a.acc[off];
                    a.acc[off] = cadd(s, cscale(rhs, R(2)));
                    next = cadd(wn, cscale(rhs, a.h_full));
                    a.w1[off] = next;
                } else {
                    const cx<R> s = a.acc[off];
                    next = cadd(wn, cscale(cadd(s, rhs), a.h_sixth));
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
            const R inv = k2 > 0 ? R(1) / (R)k2 : R(0);
            R s = R(1);
            if (f == 0)
                s = inv;
            else if (f == 1)
                s = inv * (R)ky;
            else if (f == 3)
                s = (R)ky;
            u[idx] = cscale(v[idx], s);
        }
        FFT::inverse(ctx, u, L, t, a.tw);
        if (live) {
#pragma unroll
            for (int n2 = 0; n2 < E; ++n2)
                mix[f * G::MIXF + tiled(KP, c, t + T * n2)] = u[n2];
        }
    }
}

// ----------------------------------------------------------- row pass ----
template <typename R> struct RowArgs {
    int B;
    const cx<R> *mix;
    cx<R> *nl;
    const cx<R> *tw;
    float *cfl; // per row pair running max of |u|+|v| (null: off)
};

template <int N, typename R, class Ctx>
HZL_HD void row_body(const Ctx &ctx, const RowArgs<R> &a, cx<R> *lds, long block) {
    using G = Geometry<N>;
    using FFT = GroupFFT<N, G::E, R>;
    constexpr int E = G::E, T = G::T, K = G::K, KP = G::KP, PPW = G::ROW_PPW;
    const int tid = ctx.tid();
    const int t = tid / PPW, g = tid % PPW; // pair-fastest lanes: coalesced tile rows
    const long gp = block * PPW + g;
    const bool live = gp < (long)a.B * (N / 2);
    const int m = live ? (int)((2 * gp) / N) : 0, y0 = live ? (int)((2 * gp) % N) : 0;
    cx<R> *L = lds + g * FFT::LDS_ELEMS;
    const cx<R> *mix = a.mix + (long)m * 4 * G::MIXF;

    R p[2][E];
    R vmax = R(0);
#pragma unroll
    for (int row = 0; row < 2; ++row) {
        const int y = y0 + row;
        cx<R> z1[E], z2[E];
#pragma unroll
        for (int idx = 0; idx < E; ++idx) {
            const int k = FFT::index(t, idx);
            z1[idx] = {0, 0};
            z2[idx] = {0, 0};
            const bool pos = k <= K, neg = k >= N - K;
            if (live && (pos || neg)) {
                const int kx = pos ? k : N - k;
                const long o = tiled(KP, kx, y);
                const cx<R> A = mix[o], Bv = mix[G::MIXF + o], C = mix[2 * G::MIXF + o],
                            D = mix[3 * G::MIXF + o];
                const R fx = (R)kx;
                const cx<R> U = mul_i(Bv), V = mul_mi(cscale(A, fx)), WX = mul_i(cscale(C, fx)),
                            WY = mul_i(D);
                if (k == 0) {
                    // DC terms of real signals: keep the real parts only
                    z1[idx] = {U.x, R(0)};
                    z2[idx] = {R(0), WY.x};
                } else if (pos) {
                    z1[idx] = {U.x - WX.y, U.y + WX.x};
                    z2[idx] = {V.x - WY.y, V.y + WY.x};
                } else {
                    z1[idx] = {U.x + WX.y, WX.x - U.y};
                    z2[idx] = {V.x + WY.y, WY.x - V.y};
                }
            }
        }
        FFT::inverse(ctx, z1, L, t, a.tw);
        FFT::inverse(ctx, z2, L, t, a.tw);
#pragma unroll
        for (int j = 0; j < E; ++j) {
            p[row][j] = z1[j].x * z1[j].y + z2[j].x * z2[j].y;
            R s = (z1[j].x < 0 ? -z1[j].x : z1[j].x) + (z2[j].x < 0 ? -z2[j].x : z2[j].x);
            if (!(s <= R(3.0e38)))
                s = R(3.0e38) * R(10); // non-finite velocity poisons the maximum
            vmax = s > vmax ? s : vmax;
        }
    }
    cx<R> q[E];
#pragma unroll
    for (int j = 0; j < E; ++j)
        q[j] = {p[0][j], p[1][j]};
    FFT::forward(ctx, q, L, t, a.tw);
#pragma unroll
    for (int idx = 0; idx < E; ++idx)
        L[FFT::index(t, idx)] = q[idx];
    ctx.sync();
    cx<R> *nl = a.nl + (long)m * G::MIXF;
#pragma unroll
    for (int idx = 0; idx < E; ++idx) {
        const int k = FFT::index(t, idx);
        if (live && k <= K) {
            const cx<R> qa = L[k], qb = L[(N - k) % N];
            const cx<R> pa = {R(0.5) * (qa.x + qb.x), R(0.5) * (qa.y - qb.y)};
            const cx<R> pb = {R(0.5) * (qa.y + qb.y), R(0.5) * (qb.x - qa.x)};
            const long o = tiled(KP, k, y0);
            nl[o] = pa;
            nl[o + 1] = pb;
        }
    }
    ctx.sync();
    if (a.cfl) {
        // reduce |u|+|v| over the T threads of this pair through LDS
        R *red = reinterpret_cast<R *>(lds);
        red[g * T + t] = vmax;
        ctx.sync();
        if (live && t == 0) {
            R mx = R(0);
            for (int i = 0; i < T; ++i)
                mx = red[g * T + i] > mx ? red[g * T + i] : mx;
            const float prev = a.cfl[gp];
            const float cur = (float)mx;
            a.cfl[gp] = (cur > prev || cur != cur) ? cur : prev;
        }
        ctx.sync();
    }
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
    ForcingModel forcing;
    uint64_t step; // forcing of this step is used for the injection
    int with_forcing;
    const float *cfl; // per pair maxima or null
    int pairs;        // N/2
    int n;
    double *out; // [B][6]: energy, enstrophy, dissipation, injection, cfl_rate, unused
};

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
            const cx<double> F = model_forcing(a.forcing, (int)member, c, ky, a.step);
            inj += mult * (wr * F.x + wi * F.y) / k2;
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
