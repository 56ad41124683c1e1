// enkf.hip -- stochastic (perturbed-observation) ensemble Kalman filter on the
// engine's member states, restating reference ensemble_kalman_filter
// (src/filter.cpp:204-356) on the retained half plane.
//
// Two analysis modes over one observation group:
//   serial (reference order): scalar observations one after another; for each,
//     point predictions h_i of every member, mean / variance, the state-
//     prediction covariance cov = sum_i x_i (h_i - hbar) / (M - 1), and
//     x_i += cov (y + sd eps_i - h_i) / (var + sd^2), followed by the
//     reference's conjugate-pair symmetrisation (src/solver.cpp:403-425).
//     All of it in FP64 arithmetic on device, state stored in R.
//   batch (FP32 ensembles): all m observations at once,
//     X_a = X + P S^-1 (D - H X),  P = X'^T (HX)' / (M - 1),
//     S = (HX)'^T (HX)' / (M - 1) + diag(sd^2),
//     with the four dense products on MFMA (FP32 operands, FP32 accumulation)
//     and the m x m Cholesky factorisation and inverse in FP64 on the host.
// Serial mode optionally tapers each covariance with the Gaspari-Cohn
// function of the periodic grid distance to the observation, in physical
// space (src/filter.cpp:290-308).
// Observation perturbations eps come from the reference NormalStream
// (mt19937_64 seeded stream_seed(seed, M), Box-Muller with a cached second
// draw; src/solver.cpp:41-54), consumed observation-major, member-minor as in
// src/filter.cpp:311-313.  Point predictions are direct evaluations of the
// retained Fourier sum at the grid point (exactly the physical field there).
#include "engine.hpp"
#include "gemm_launch.hpp"
#include "tables.hpp"

#include <cmath>
#include <random>
#include <stdexcept>
#include <vector>

namespace hzl_amd {

// Deterministic tree sum over a 256-thread block (result valid in thread 0
// and broadcast through red[0]).
__device__ double block_sum(double v, double *red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            red[t] += red[t + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// pred[o * stride + m] = w_m(x_o, y_o) for block (m, o): the retained sum
// sum_{c,ky} (c > 0 ? 2 : 1) Re(w[c][ky] ex_o[c] ey_o[ky]).
template <typename R>
__global__ void __launch_bounds__(256)
    k_enkf_predict(const cx<R> *w, long spec, int K, const cx<double> *ex, const cx<double> *ey,
                   double *pred, long stride) {
    __shared__ double red[256];
    const int m = blockIdx.x, o = blockIdx.y, J = 2 * K + 1;
    const cx<R> *wm = w + (long)m * spec;
    const cx<double> *exo = ex + (long)o * (K + 1), *eyo = ey + (long)o * J;
    double acc = 0;
    for (long e = threadIdx.x; e < spec; e += 256) {
        const int c = (int)(e / J), j = (int)(e % J);
        const cx<double> ph = cmul(exo[c], eyo[j]);
        acc += (c ? 2.0 : 1.0) * ((double)wm[e].x * ph.x - (double)wm[e].y * ph.y);
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0)
        pred[(long)o * stride + m] = s;
}

// One block: mean and variance (1/(M-1)) of the predictions, prescaled
// anomalies delta_i = (h_i - hbar)/(M-1) and gains g_i = (y + sd eps_i - h_i)/(var + sd^2).
__global__ void __launch_bounds__(256)
    k_enkf_innov(const double *pred, int B, double value, double sd, const double *eps,
                 double *delta, double *gain) {
    __shared__ double red[256];
    double s = 0;
    for (int i = threadIdx.x; i < B; i += 256)
        s += pred[i];
    const double mean = block_sum(s, red) / B;
    double v = 0;
    for (int i = threadIdx.x; i < B; i += 256) {
        const double d = pred[i] - mean;
        v += d * d;
    }
    const double var = block_sum(v, red) / (B - 1);
    const double den = var + sd * sd;
    for (int i = threadIdx.x; i < B; i += 256) {
        delta[i] = (pred[i] - mean) / (B - 1);
        gain[i] = (value + sd * eps[i] - pred[i]) / den;
    }
}

// partial[c][e] = sum over the members of chunk c of w_i[e] delta_i.
template <typename R>
__global__ void k_enkf_cov_partial(const cx<R> *w, int B, long spec, const double *delta,
                                   int chunks, cx<double> *partial) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    if (e >= spec)
        return;
    const int per = (B + chunks - 1) / chunks, lo = c * per, hi = lo + per < B ? lo + per : B;
    double re = 0, im = 0;
    for (int i = lo; i < hi; ++i) {
        const cx<R> v = w[(long)i * spec + e];
        re += (double)v.x * delta[i];
        im += (double)v.y * delta[i];
    }
    partial[(long)c * spec + e] = {re, im};
}

__global__ void k_enkf_cov_final(const cx<double> *partial, int chunks, long spec,
                                 cx<double> *cov) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= spec)
        return;
    double re = 0, im = 0;
    for (int c = 0; c < chunks; ++c) {
        re += partial[(long)c * spec + e].x;
        im += partial[(long)c * spec + e].y;
    }
    cov[e] = {re, im};
}

// Block per member: w += cov g_m with the kx = 0 column symmetrised
// (w(0,ky) = (a + conj(b)) / 2, w(0,-ky) = conj of it), then, when ex is given,
// the member's prediction of the next observation from the updated state.
template <typename R>
__global__ void __launch_bounds__(256)
    k_enkf_update(cx<R> *w, long spec, int K, const cx<double> *cov, const double *gain,
                  const cx<double> *ex, const cx<double> *ey, double *pred_next) {
    __shared__ double red[256];
    const int m = blockIdx.x, J = 2 * K + 1;
    cx<R> *wm = w + (long)m * spec;
    const double g = gain[m];
    double acc = 0;
    for (long e = threadIdx.x; e < spec; e += 256) {
        const int c = (int)(e / J), j = (int)(e % J), ky = j - K;
        if (c == 0) {
            if (ky <= 0)
                continue; // (0,0) stays zero; ky < 0 is written by its partner
            const long p = K - ky;
            const double ar = (double)wm[e].x + cov[e].x * g, ai = (double)wm[e].y + cov[e].y * g;
            const double br = (double)wm[p].x + cov[p].x * g, bi = (double)wm[p].y + cov[p].y * g;
            const double sr = 0.5 * (ar + br), si = 0.5 * (ai - bi);
            wm[e] = {(R)sr, (R)si};
            wm[p] = {(R)sr, (R)-si};
            if (ex) {
                const cx<double> pa = cmul(ex[0], ey[j]), pb = cmul(ex[0], ey[p]);
                acc += ((double)(R)sr * pa.x - (double)(R)si * pa.y) +
                       ((double)(R)sr * pb.x + (double)(R)si * pb.y);
            }
        } else {
            const double vr = (double)wm[e].x + cov[e].x * g, vi = (double)wm[e].y + cov[e].y * g;
            wm[e] = {(R)vr, (R)vi};
            if (ex) {
                const cx<double> ph = cmul(ex[c], ey[j]);
                acc += 2.0 * ((double)(R)vr * ph.x - (double)(R)vi * ph.y);
            }
        }
    }
    if (ex) {
        const double s = block_sum(acc, red);
        if (threadIdx.x == 0)
            pred_next[m] = s;
    }
}

// Member mean per mode in two passes: partial[c][e] = sum over the members
// of chunk c of w_i[e] / B, then mean[e] = sum_c partial[c][e]
// (src/filter.cpp:254-258 sums the same terms in member order).
template <typename R>
__global__ void k_enkf_mean_partial(const cx<R> *w, int B, long spec, int chunks,
                                    cx<double> *partial) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    if (e >= spec)
        return;
    const int per = (B + chunks - 1) / chunks, lo = c * per, hi = lo + per < B ? lo + per : B;
    double re = 0, im = 0;
    for (int i = lo; i < hi; ++i) {
        re += (double)w[(long)i * spec + e].x / B;
        im += (double)w[(long)i * spec + e].y / B;
    }
    partial[(long)c * spec + e] = {re, im};
}

// ---- OU forcing state (the reference's latent forcing, src/filter.cpp:237-319):
// F[B][P] coefficients per shell pair, one thread per pair.
// Inflation F = mean + lambda (F - mean), mean = sum over members of F / B.
__global__ void __launch_bounds__(256) k_ou_inflate(cx<double> *F, int B, int P, double lambda) {
    const int j = (int)(blockIdx.x * 256 + threadIdx.x);
    if (j >= P)
        return;
    cx<double> mu = {0, 0};
    for (int i = 0; i < B; ++i) {
        mu.x += F[(long)i * P + j].x / B;
        mu.y += F[(long)i * P + j].y / B;
    }
    for (int i = 0; i < B; ++i) {
        cx<double> &f = F[(long)i * P + j];
        f = {mu.x + lambda * (f.x - mu.x), mu.y + lambda * (f.y - mu.y)};
    }
}

// fcov[j] = sum over members of (F_ij - mean_j) delta_i (delta prescaled by 1/(M-1)).
__global__ void __launch_bounds__(256) k_ou_cov(const cx<double> *F, int B, int P,
                                                const double *delta, cx<double> *fcov) {
    const int j = (int)(blockIdx.x * 256 + threadIdx.x);
    if (j >= P)
        return;
    cx<double> mu = {0, 0}, c = {0, 0};
    for (int i = 0; i < B; ++i) {
        mu.x += F[(long)i * P + j].x / B;
        mu.y += F[(long)i * P + j].y / B;
    }
    for (int i = 0; i < B; ++i) {
        c.x += (F[(long)i * P + j].x - mu.x) * delta[i];
        c.y += (F[(long)i * P + j].y - mu.y) * delta[i];
    }
    fcov[j] = c;
}

// y = 0 over n complex entries (a fill kernel, safe under graph capture).
__global__ void __launch_bounds__(256) k_zero_cx(cx<double> *y, long n) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e < n)
        y[e] = {0, 0};
}

// Pair covariance into a half-plane field (zero elsewhere), for the taper.
__global__ void __launch_bounds__(256) k_ou_scatter(const cx<double> *fcov, int P,
                                                    const int *off, int K, cx<double> *field) {
    const int j = (int)(blockIdx.x * 256 + threadIdx.x);
    if (j >= P)
        return;
    const int J = 2 * K + 1, o = off[j] >= 0 ? off[j] : ~off[j];
    const cx<double> v = off[j] >= 0 ? fcov[j] : cx<double>{fcov[j].x, -fcov[j].y};
    field[o] = v;
    if (o / J == 0) // kx = 0: the conjugate mode is stored too
        field[(o / J) * J + (2 * K - o % J)] = {v.x, -v.y};
}

// Tapered pair covariance back from the half-plane field, projected as
// update_latent_state projects the forcing (src/solver.cpp:412-416): a pair
// is the mean of its mode and its partner's conjugate.
__global__ void __launch_bounds__(256) k_ou_gather(const cx<double> *field, int P,
                                                   const int *off, int K, cx<double> *fcov) {
    const int j = (int)(blockIdx.x * 256 + threadIdx.x);
    if (j >= P)
        return;
    const int J = 2 * K + 1, o = off[j] >= 0 ? off[j] : ~off[j];
    cx<double> v = field[o];
    if (o / J == 0) {
        const cx<double> m = field[2 * K - o % J];
        v = {0.5 * (v.x + m.x), 0.5 * (v.y - m.y)};
    }
    fcov[j] = off[j] >= 0 ? v : cx<double>{v.x, -v.y};
}

// F_ij += fcov_j g_i (the member's scaled innovation).
__global__ void __launch_bounds__(256) k_ou_update(cx<double> *F, int B, int P,
                                                   const cx<double> *fcov, const double *gain) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)B * P)
        return;
    const int i = (int)(e / P), j = (int)(e % P);
    const double g = gain[i];
    F[e] = {F[e].x + fcov[j].x * g, F[e].y + fcov[j].y * g};
}

// Multiplicative anomaly inflation w = mean + lambda (w - mean).
template <typename R>
__global__ void k_enkf_inflate(cx<R> *w, int B, long spec, const cx<double> *mean,
                               double lambda) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)B * spec)
        return;
    const long e = i % spec;
    const cx<double> mu = mean[e];
    w[i] = {(R)(mu.x + lambda * ((double)w[i].x - mu.x)),
            (R)(mu.y + lambda * ((double)w[i].y - mu.y))};
}

// partial[c][e] = (kx > 0 ? 2 : 1) sum over the members of chunk c of
// |w_i[e] - mean[e]|^2: summed over chunks and modes, the full-field sum of
// the reference's posterior variance (src/filter.cpp:343-347).
template <typename R>
__global__ void k_enkf_spread_partial(const cx<R> *w, int B, long spec, int K,
                                      const cx<double> *mean, int chunks, double *partial) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    if (e >= spec)
        return;
    const int per = (B + chunks - 1) / chunks, lo = c * per, hi = lo + per < B ? lo + per : B;
    double s = 0;
    for (int i = lo; i < hi; ++i) {
        const double dr = (double)w[(long)i * spec + e].x - mean[e].x;
        const double di = (double)w[(long)i * spec + e].y - mean[e].y;
        s += dr * dr + di * di;
    }
    partial[(long)c * spec + e] = (e / (2 * K + 1) > 0 ? 2.0 : 1.0) * s;
}

// out[e] = sum_s part[s * count + e]
__global__ void k_sum_slices(const float *part, int slices, long count, float *out) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= count)
        return;
    float s = 0;
    for (int k = 0; k < slices; ++k)
        s += part[(long)k * count + e];
    out[e] = s;
}

// Gaspari-Cohn fifth-order compactly supported correlation (Gaspari & Cohn
// 1999, eq. 4.10) of half-support `radius`, the taper of src/filter.cpp:25-36.
HZL_HD double gaspari_cohn(double d, double radius) {
    if (radius <= 0)
        return 1;
    const double x = d / radius;
    if (x >= 2)
        return 0;
    if (x <= 1)
        return 1 - 5.0 / 3 * x * x + 5.0 / 8 * x * x * x + 0.5 * x * x * x * x -
               0.25 * x * x * x * x * x;
    return 4 - 5 * x + 5.0 / 3 * x * x + 5.0 / 8 * x * x * x - 0.5 * x * x * x * x +
           x * x * x * x * x / 12 - 2.0 / (3 * x);
}

// Localisation of the covariance (retained half plane, FP64) around the
// observation (xo, yo), as src/filter.cpp:290-308: to physical space,
// multiply by the taper of the periodic grid distance, back to spectral.
// Four separable DFT passes over the (K+1) x n mixed field; cs/sn are
// cos/sin(2 pi j / n).
__global__ void k_loc_inv_y(const cx<double> *cov, int K, int n, const double *cs,
                            const double *sn, cx<double> *mixed) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)(K + 1) * n)
        return;
    const int c = (int)(i / n), y = (int)(i % n), J = 2 * K + 1;
    double re = 0, im = 0;
    for (int j = 0; j < J; ++j) {
        const int t = (int)((((long)(j - K) * y) % n + n) % n);
        const cx<double> v = cov[(long)c * J + j];
        re += v.x * cs[t] - v.y * sn[t];
        im += v.x * sn[t] + v.y * cs[t];
    }
    mixed[i] = {re, im};
}

__global__ void k_loc_inv_x_taper(const cx<double> *mixed, int K, int n, const double *cs,
                                  const double *sn, int xo, int yo, double radius,
                                  double *phys) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)n * n)
        return;
    const int y = (int)(i / n), x = (int)(i % n);
    double v = 0;
    for (int c = 0; c <= K; ++c) {
        const int t = (int)(((long)c * x) % n);
        const cx<double> m = mixed[(long)c * n + y];
        v += (c ? 2.0 : 1.0) * (m.x * cs[t] - m.y * sn[t]);
    }
    const int ax = x > xo ? x - xo : xo - x, ay = y > yo ? y - yo : yo - y;
    const double dx = ax < n - ax ? ax : n - ax, dy = ay < n - ay ? ay : n - ay;
    phys[i] = v * gaspari_cohn(sqrt(dx * dx + dy * dy), radius);
}

__global__ void k_loc_fwd_x(const double *phys, int K, int n, const double *cs, const double *sn,
                            cx<double> *mixed) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)(K + 1) * n)
        return;
    const int c = (int)(i / n), y = (int)(i % n);
    double re = 0, im = 0;
    for (int x = 0; x < n; ++x) {
        const int t = (int)(((long)c * x) % n);
        const double v = phys[(long)y * n + x];
        re += v * cs[t];
        im -= v * sn[t];
    }
    mixed[i] = {re, im};
}

__global__ void k_loc_fwd_y(const cx<double> *mixed, int K, int n, const double *cs,
                            const double *sn, cx<double> *cov) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const int J = 2 * K + 1;
    if (e >= (long)(K + 1) * J)
        return;
    const int c = (int)(e / J), j = (int)(e % J);
    double re = 0, im = 0;
    for (int y = 0; y < n; ++y) {
        const int t = (int)((((long)(j - K) * y) % n + n) % n);
        const cx<double> v = mixed[(long)c * n + y];
        re += v.x * cs[t] + v.y * sn[t];
        im += v.y * cs[t] - v.x * sn[t];
    }
    const double z = 1.0 / ((double)n * n);
    cov[e] = (c == 0 && j == K) ? cx<double>{0, 0} : cx<double>{re * z, im * z};
}

// Batch mode: per observation column o (block), HA[i][o] = HX[i][o] - mean_o
// and Dm[o][i] = y_o + sd_o eps[o][i] - HX[i][o].
__global__ void __launch_bounds__(256)
    k_enkf_batch_anomalies(const float *HX, int B, int m, const double *value, const double *sd,
                           const double *eps, float *HA, float *Dm) {
    __shared__ double red[256];
    const int o = blockIdx.x;
    double s = 0;
    for (int i = threadIdx.x; i < B; i += 256)
        s += HX[(long)i * m + o];
    const double mean = block_sum(s, red) / B;
    for (int i = threadIdx.x; i < B; i += 256) {
        const double h = HX[(long)i * m + o];
        HA[(long)i * m + o] = (float)(h - mean);
        Dm[(long)o * B + i] = (float)(value[o] + sd[o] * eps[(long)o * B + i] - h);
    }
}

// Batch mode: symmetrise the kx = 0 column of every member after the update.
template <typename R> __global__ void k_enkf_symmetrise(cx<R> *w, int B, long spec, int K) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)B * (K + 1))
        return;
    const long m = i / (K + 1);
    const int ky = (int)(i % (K + 1));
    cx<R> *col = w + m * spec;
    if (ky == 0) {
        col[K] = {0, 0};
        return;
    }
    const cx<R> a = col[K + ky], b = col[K - ky];
    const R sr = (R)0.5 * (a.x + b.x), si = (R)0.5 * (a.y - b.y);
    col[K + ky] = {sr, si};
    col[K - ky] = {sr, (R)-si};
}

// The reference NormalStream (src/solver.cpp:41-54).
struct NormalStream {
    std::mt19937_64 engine;
    bool has_spare = false;
    double spare = 0;
    explicit NormalStream(uint64_t seed) : engine(seed) {}
    double uniform() {
        return ((double)(engine() >> 11) + 0.5) * 0x1.0p-53;
    }
    double operator()() {
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        const double r = std::sqrt(-2 * std::log(uniform())), angle = 2 * PI * uniform();
        spare = r * std::sin(angle);
        has_spare = true;
        return r * std::cos(angle);
    }
};

struct EnkfState : ModuleState {
    int mode;
    NormalStream normals;
    Buf<cx<double>> ex, ey, partial, cov, mean;
    Buf<double> pred, pred2, delta, gain, eps, spread, value, sd, spart, cs, sn, phys;
    Buf<cx<double>> mixed, fcov, ffield;
    // batch mode (FP32)
    Buf<float> H, HX, HA, P, S, Sinv, Dm, Z, HXpart;
    EnkfState(int mode_, uint64_t seed) : mode(mode_), normals(seed) {}
};

namespace {
template <typename T> void upload(Buf<T> &b, const T *h, size_t n, hipStream_t s) {
    b.alloc(n);
    hcheck(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, s), "EnKF upload");
}

} // namespace

// Cholesky factorisation and inverse of an SPD matrix (FP64, in place: a
// becomes the inverse).
inline void spd_inverse(std::vector<double> &a, int m) {
    std::vector<double> L((size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j) {
        double d = a[(size_t)j * m + j];
        for (int k = 0; k < j; ++k)
            d -= L[(size_t)j * m + k] * L[(size_t)j * m + k];
        if (!(d > 0))
            throw std::runtime_error("EnKF innovation covariance is not positive definite");
        L[(size_t)j * m + j] = std::sqrt(d);
        for (int i = j + 1; i < m; ++i) {
            double s = a[(size_t)i * m + j];
            for (int k = 0; k < j; ++k)
                s -= L[(size_t)i * m + k] * L[(size_t)j * m + k];
            L[(size_t)i * m + j] = s / L[(size_t)j * m + j];
        }
    }
    // inverse of L (lower), then A^-1 = L^-T L^-1
    std::vector<double> Li((size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j) {
        Li[(size_t)j * m + j] = 1.0 / L[(size_t)j * m + j];
        for (int i = j + 1; i < m; ++i) {
            double s = 0;
            for (int k = j; k < i; ++k)
                s -= L[(size_t)i * m + k] * Li[(size_t)k * m + j];
            Li[(size_t)i * m + j] = s / L[(size_t)i * m + i];
        }
    }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0;
            for (int k = i; k < m; ++k)
                s += Li[(size_t)k * m + i] * Li[(size_t)k * m + j];
            a[(size_t)i * m + j] = a[(size_t)j * m + i] = s;
        }
}

template <typename R> struct Enkf {
    EngineBase &eng;
    EnkfState &st;
    EngineViews ev;
    long spec;
    int K, J;
    cx<R> *w;
    Enkf(EngineBase &e, EnkfState &s) : eng(e), st(s), ev(e.views()) {
        K = ev.K;
        J = 2 * K + 1;
        spec = (long)(K + 1) * J;
        w = reinterpret_cast<cx<R> *>(ev.w0);
    }

    static constexpr int CHUNKS = 64; // member chunks of the two-pass reductions

    void mean() {
        st.mean.alloc((size_t)spec);
        st.partial.alloc((size_t)CHUNKS * spec);
        hipLaunchKernelGGL(k_enkf_mean_partial<R>, dim3((unsigned)((spec + 255) / 256), CHUNKS),
                           dim3(256), 0, ev.stream, w, ev.B, spec, CHUNKS, st.partial.p);
        hipLaunchKernelGGL(k_enkf_cov_final, grid1d(spec), dim3(256), 0, ev.stream, st.partial.p,
                           CHUNKS, spec, st.mean.p);
    }

    // posterior spread (after mean()), src/filter.cpp:343-347
    double spread() {
        st.spart.alloc((size_t)CHUNKS * spec);
        hipLaunchKernelGGL(k_enkf_spread_partial<R>, dim3((unsigned)((spec + 255) / 256), CHUNKS),
                           dim3(256), 0, ev.stream, w, ev.B, spec, K, st.mean.p, CHUNKS,
                           st.spart.p);
        std::vector<double> h((size_t)CHUNKS * spec);
        hcheck(hipMemcpyAsync(h.data(), st.spart.p, h.size() * sizeof(double),
                              hipMemcpyDeviceToHost, ev.stream),
               "spread download");
        hcheck(hipStreamSynchronize(ev.stream), "spread");
        double s = 0;
        for (double v : h)
            s += v;
        return std::sqrt(s / (ev.B - 1));
    }

    void mean_field(double *field) {
        std::vector<cx<double>> h((size_t)spec);
        hcheck(hipMemcpyAsync(h.data(), st.mean.p, spec * sizeof(cx<double>),
                              hipMemcpyDeviceToHost, ev.stream),
               "mean download");
        hcheck(hipStreamSynchronize(ev.stream), "mean");
        half_to_full<double>(ev.n, h.data(), field);
    }

    void assimilate(size_t nobs, const int *x, const int *y, const double *value,
                    const double *sd, double inflation, double radius, double *record,
                    double *mean_out) {
        const int B = ev.B, n = ev.n;
        // point-evaluation tables of every observation
        std::vector<cx<double>> ex(nobs * (K + 1)), ey(nobs * J);
        for (size_t o = 0; o < nobs; ++o) {
            for (int c = 0; c <= K; ++c) {
                const double a = 2 * PI * (double)((long)c * x[o] % n) / n;
                ex[o * (K + 1) + c] = {std::cos(a), std::sin(a)};
            }
            for (int j = 0; j < J; ++j) {
                const long ky = j - K;
                const double a = 2 * PI * (double)((((ky * y[o]) % n) + n) % n) / n;
                ey[o * J + j] = {std::cos(a), std::sin(a)};
            }
        }
        upload(st.ex, ex.data(), ex.size(), ev.stream);
        upload(st.ey, ey.data(), ey.size(), ev.stream);
        if (inflation > 1) {
            mean();
            hipLaunchKernelGGL(k_enkf_inflate<R>, grid1d((long)B * spec), dim3(256), 0,
                               ev.stream, w, B, spec, st.mean.p, inflation);
            if (ev.ou_F)
                hipLaunchKernelGGL(k_ou_inflate, grid1d(ev.npairs), dim3(256), 0, ev.stream,
                                   ev.ou_F, B, ev.npairs, inflation);
        }
        // forecast sensor RMSE of the ensemble mean, before any update
        // (src/filter.cpp:254-264)
        mean();
        st.pred.alloc(std::max<size_t>(nobs, (size_t)B));
        hipLaunchKernelGGL(k_enkf_predict<double>, dim3(1, (unsigned)nobs), dim3(256), 0,
                           ev.stream, st.mean.p, spec, K, st.ex.p, st.ey.p, st.pred.p, 1L);
        std::vector<double> fc(nobs);
        hcheck(hipMemcpyAsync(fc.data(), st.pred.p, nobs * sizeof(double), hipMemcpyDeviceToHost,
                              ev.stream),
               "forecast download");
        hcheck(hipStreamSynchronize(ev.stream), "forecast");
        double err = 0;
        for (size_t o = 0; o < nobs; ++o)
            err += (fc[o] - value[o]) * (fc[o] - value[o]);
        record[0] = std::sqrt(err / (double)nobs);
        // perturbations, observation-major then member
        std::vector<double> eps(nobs * (size_t)B);
        for (auto &e : eps)
            e = st.normals();
        upload(st.eps, eps.data(), eps.size(), ev.stream);
        if (st.mode == HZL_ENKF_SERIAL)
            serial(nobs, x, y, value, sd, radius);
        else
            batch(nobs, value, sd);
        hcheck(hipGetLastError(), "EnKF kernels");
        eng.invalidate_mixed();
        mean();
        record[1] = spread();
        if (mean_out)
            mean_field(mean_out);
    }

    void localize(int xo, int yo, double radius) {
        localize(st.cov.p, xo, yo, radius);
    }

    // taper the half-plane FP64 field cov in place
    void localize(cx<double> *cov, int xo, int yo, double radius) {
        const int n = ev.n;
        if (!st.cs.p || st.cs.n < (size_t)n) {
            std::vector<double> c(n), sv(n);
            for (int j = 0; j < n; ++j) {
                c[j] = std::cos(2 * PI * j / n);
                sv[j] = std::sin(2 * PI * j / n);
            }
            upload(st.cs, c.data(), n, ev.stream);
            upload(st.sn, sv.data(), n, ev.stream);
        }
        const long mix = (long)(K + 1) * n;
        st.mixed.alloc((size_t)mix);
        st.phys.alloc((size_t)n * n);
        hipLaunchKernelGGL(k_loc_inv_y, grid1d(mix), dim3(256), 0, ev.stream, cov, K, n,
                           st.cs.p, st.sn.p, st.mixed.p);
        hipLaunchKernelGGL(k_loc_inv_x_taper, grid1d((long)n * n), dim3(256), 0, ev.stream,
                           st.mixed.p, K, n, st.cs.p, st.sn.p, xo, yo, radius, st.phys.p);
        hipLaunchKernelGGL(k_loc_fwd_x, grid1d(mix), dim3(256), 0, ev.stream, st.phys.p, K, n,
                           st.cs.p, st.sn.p, st.mixed.p);
        hipLaunchKernelGGL(k_loc_fwd_y, grid1d(spec), dim3(256), 0, ev.stream, st.mixed.p, K, n,
                           st.cs.p, st.sn.p, cov);
    }

    void serial(size_t nobs, const int *x, const int *y, const double *value, const double *sd,
                double radius) {
        const int B = ev.B, chunks = CHUNKS;
        st.pred.alloc((size_t)B);
        st.pred2.alloc((size_t)B);
        st.delta.alloc((size_t)B);
        st.gain.alloc((size_t)B);
        st.partial.alloc((size_t)chunks * spec);
        st.cov.alloc((size_t)spec);
        double *cur = st.pred.p, *next = st.pred2.p;
        hipLaunchKernelGGL(k_enkf_predict<R>, dim3((unsigned)B, 1), dim3(256), 0, ev.stream, w,
                           spec, K, st.ex.p, st.ey.p, cur, 0L);
        for (size_t o = 0; o < nobs; ++o) {
            hipLaunchKernelGGL(k_enkf_innov, dim3(1), dim3(256), 0, ev.stream, cur, B, value[o],
                               sd[o], st.eps.p + o * (size_t)B, st.delta.p, st.gain.p);
            hipLaunchKernelGGL(k_enkf_cov_partial<R>, dim3((unsigned)((spec + 255) / 256), chunks),
                               dim3(256), 0, ev.stream, w, B, spec, st.delta.p, chunks,
                               st.partial.p);
            hipLaunchKernelGGL(k_enkf_cov_final, grid1d(spec), dim3(256), 0, ev.stream,
                               st.partial.p, chunks, spec, st.cov.p);
            if (radius > 0)
                localize(x[o], y[o], radius);
            if (ev.ou_F) {
                // the forcing state: its covariance with the prediction (all
                // members before any update), tapered like the state's
                const int P = ev.npairs;
                st.fcov.alloc((size_t)P);
                hipLaunchKernelGGL(k_ou_cov, grid1d(P), dim3(256), 0, ev.stream, ev.ou_F, B, P,
                                   st.delta.p, st.fcov.p);
                if (radius > 0) {
                    st.ffield.alloc((size_t)spec);
                    hipLaunchKernelGGL(k_zero_cx, grid1d(spec), dim3(256), 0, ev.stream,
                                       st.ffield.p, spec);
                    hipLaunchKernelGGL(k_ou_scatter, grid1d(P), dim3(256), 0, ev.stream,
                                       st.fcov.p, P, ev.ou_off, K, st.ffield.p);
                    localize(st.ffield.p, x[o], y[o], radius);
                    hipLaunchKernelGGL(k_ou_gather, grid1d(P), dim3(256), 0, ev.stream,
                                       st.ffield.p, P, ev.ou_off, K, st.fcov.p);
                }
                hipLaunchKernelGGL(k_ou_update, grid1d((long)B * P), dim3(256), 0, ev.stream,
                                   ev.ou_F, B, P, st.fcov.p, st.gain.p);
            }
            const bool more = o + 1 < nobs;
            hipLaunchKernelGGL(k_enkf_update<R>, dim3((unsigned)B), dim3(256), 0, ev.stream, w,
                               spec, K, st.cov.p, st.gain.p,
                               more ? st.ex.p + (o + 1) * (K + 1) : nullptr,
                               more ? st.ey.p + (o + 1) * J : nullptr, next);
            std::swap(cur, next);
        }
    }

    void batch(size_t nobs, const double *value, const double *sd);
};

template <> void Enkf<double>::batch(size_t, const double *, const double *) {
    throw std::invalid_argument("batch EnKF analysis needs an FP32 ensemble");
}

template <> void Enkf<float>::batch(size_t nobs, const double *value, const double *sd) {
    const int B = ev.B, m = (int)nobs;
    const long D = 2 * spec; // real components per member
    // observation operator rows: Re(w ph) = w.x ph.x - w.y ph.y, weighted
    std::vector<float> H((size_t)m * D);
    std::vector<cx<double>> ex((size_t)m * (K + 1)), ey((size_t)m * J);
    hcheck(hipMemcpyAsync(ex.data(), st.ex.p, ex.size() * sizeof(cx<double>),
                          hipMemcpyDeviceToHost, ev.stream),
           "tables");
    hcheck(hipMemcpyAsync(ey.data(), st.ey.p, ey.size() * sizeof(cx<double>),
                          hipMemcpyDeviceToHost, ev.stream),
           "tables");
    hcheck(hipStreamSynchronize(ev.stream), "tables");
    for (int o = 0; o < m; ++o)
        for (long e = 0; e < spec; ++e) {
            const int c = (int)(e / J), j = (int)(e % J);
            const cx<double> a = ex[(size_t)o * (K + 1) + c], b = ey[(size_t)o * J + j];
            const double wgt = c ? 2.0 : 1.0, pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
            H[(size_t)o * D + 2 * e] = (float)(wgt * pr);
            H[(size_t)o * D + 2 * e + 1] = (float)(-wgt * pi);
        }
    upload(st.H, H.data(), H.size(), ev.stream);
    upload(st.value, value, nobs, ev.stream);
    upload(st.sd, sd, nobs, ev.stream);
    st.HX.alloc((size_t)B * m);
    st.HA.alloc((size_t)B * m);
    st.Dm.alloc((size_t)m * B);
    st.P.alloc((size_t)D * m);
    st.S.alloc((size_t)m * m);
    st.Sinv.alloc((size_t)m * m);
    st.Z.alloc((size_t)m * B);
    const Span<const float> X{reinterpret_cast<const float *>(w), (size_t)B * D};
    const Span<void> Xw{reinterpret_cast<float *>(w), (size_t)B * D};
    // HX = X H^T, split along the long K = 2 spec axis: full slices of KC, a
    // remainder slice, then a sum over slices
    {
        const long KC = 256, full = D / KC, rest = D - full * KC, slices = full + (rest ? 1 : 0);
        const long count = (long)B * m;
        st.HXpart.alloc((size_t)(slices * count));
        if (full) {
            Gemm<float> g1(B, m, (int)KC, X, D, 1, cs(st.H), D, 1, vs(st.HXpart), m, 1, OUT_F32);
            g1.batch((int)full, KC, KC, count);
            g1.run(ev.stream);
        }
        if (rest) {
            const Span<const float> Xr{X.p + full * KC, X.n - full * KC};
            Gemm<float> g1r(B, m, (int)rest, Xr, D, 1, cs(st.H, full * KC), D, 1,
                            vs(st.HXpart, full * count), m, 1, OUT_F32);
            g1r.run(ev.stream);
        }
        hipLaunchKernelGGL(k_sum_slices, grid1d(count), dim3(256), 0, ev.stream, st.HXpart.p,
                           (int)slices, count, st.HX.p);
    }
    hipLaunchKernelGGL(k_enkf_batch_anomalies, dim3((unsigned)m), dim3(256), 0, ev.stream,
                       st.HX.p, B, m, st.value.p, st.sd.p, st.eps.p, st.HA.p, st.Dm.p);
    // P = X^T HA / (M - 1)
    Gemm<float> g2((int)D, m, B, X, 1, D, cs(st.HA), 1, m, vs(st.P), m, 1, OUT_F32);
    g2.a.alpha = 1.0f / (float)(B - 1);
    g2.run(ev.stream);
    // S = HA^T HA / (M - 1)
    Gemm<float> g3(m, m, B, cs(st.HA), 1, m, cs(st.HA), 1, m, vs(st.S), m, 1, OUT_F32);
    g3.a.alpha = 1.0f / (float)(B - 1);
    g3.run(ev.stream);
    std::vector<float> Sf((size_t)m * m);
    hcheck(hipMemcpyAsync(Sf.data(), st.S.p, Sf.size() * sizeof(float), hipMemcpyDeviceToHost,
                          ev.stream),
           "S download");
    hcheck(hipStreamSynchronize(ev.stream), "S");
    std::vector<double> S((size_t)m * m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j)
            S[(size_t)i * m + j] = 0.5 * ((double)Sf[(size_t)i * m + j] + Sf[(size_t)j * m + i]) +
                                   (i == j ? sd[i] * sd[i] : 0.0);
    spd_inverse(S, m);
    for (size_t i = 0; i < S.size(); ++i)
        Sf[i] = (float)S[i];
    upload(st.Sinv, Sf.data(), Sf.size(), ev.stream);
    // Z = S^-1 Dm
    Gemm<float> g4(m, B, m, cs(st.Sinv), m, 1, cs(st.Dm), 1, B, vs(st.Z), B, 1, OUT_F32);
    g4.run(ev.stream);
    // X += Z^T P^T
    Gemm<float> g5(B, (int)D, m, cs(st.Z), 1, B, cs(st.P), m, 1, Xw, D, 1, OUT_F32);
    g5.a.beta = 1.f;
    g5.run(ev.stream);
    hipLaunchKernelGGL(k_enkf_symmetrise<float>, grid1d((long)B * (K + 1)), dim3(256), 0,
                       ev.stream, w, B, spec, K);
}

} // namespace hzl_amd

// ---------------------------------------------------------------- C-ABI --
using hzl_amd::engine_of;
using hzl_amd::guarded;

namespace {
hzl_amd::EnkfState &enkf_of(hzl_amd::EngineBase &e) {
    auto *s = dynamic_cast<hzl_amd::EnkfState *>(e.enkf.get());
    if (!s)
        throw std::invalid_argument("ensemble Kalman filter not started");
    return *s;
}
} // namespace

extern "C" {

int hzl_enkf_begin(hzl_ensemble_t h, uint64_t seed, int mode) {
    return guarded([&] {
        hzl_amd::EngineBase &e = engine_of(h);
        const hzl_amd::EngineViews v = e.views();
        if (v.B < 2)
            throw std::invalid_argument("EnKF requires at least two ensemble members");
        if (mode != HZL_ENKF_SERIAL && mode != HZL_ENKF_BATCH)
            throw std::invalid_argument("EnKF mode must be HZL_ENKF_SERIAL or HZL_ENKF_BATCH");
        if (mode == HZL_ENKF_BATCH && v.precision != HZL_FP32)
            throw std::invalid_argument("batch EnKF analysis needs an FP32 ensemble");
        // the OU forcing is a latent state of the serial analysis, as in the
        // reference (src/filter.cpp:237-253,273-321); the batch analysis
        // updates the vorticity only
        if (mode == HZL_ENKF_BATCH && e.forcing_process() == HZL_FORCING_OU)
            throw std::invalid_argument("batch EnKF analysis runs with the white forcing process");
        e.enkf = std::make_unique<hzl_amd::EnkfState>(
            mode, hzl_amd::stream_seed(seed, (uint64_t)v.B));
    });
}

int hzl_enkf_assimilate(hzl_ensemble_t h, double dt, size_t steps, size_t nobs, const int *x,
                        const int *y, const double *value, const double *sd, double inflation,
                        double localization_radius, double *record, double *mean_field) {
    return guarded([&] {
        hzl_amd::EngineBase &e = engine_of(h);
        hzl_amd::EnkfState &st = enkf_of(e);
        if (!record)
            throw std::invalid_argument("null argument");
        if (nobs == 0 || !x || !y || !value || !sd)
            throw std::invalid_argument("observation group is empty");
        if (!std::isfinite(inflation) || inflation < 1 || !std::isfinite(localization_radius) ||
            localization_radius < 0)
            throw std::invalid_argument(
                "EnKF inflation must be >= 1 and localization radius nonnegative");
        if (localization_radius > 0 && st.mode == HZL_ENKF_BATCH)
            throw std::invalid_argument("EnKF localization needs the serial analysis");
        const hzl_amd::EngineViews v = e.views();
        if (st.mode == HZL_ENKF_BATCH && nobs > 1024)
            throw std::invalid_argument("batch EnKF analysis takes at most 1024 observations");
        for (size_t o = 0; o < nobs; ++o)
            if (x[o] < 0 || y[o] < 0 || x[o] >= v.n || y[o] >= v.n || !std::isfinite(value[o]) ||
                !std::isfinite(sd[o]) || sd[o] <= 0)
                throw std::invalid_argument("observations must be finite, in bounds, and have "
                                            "positive uncertainty");
        if (steps)
            e.step(dt, steps);
        if (v.precision == HZL_FP32)
            hzl_amd::Enkf<float>(e, st).assimilate(nobs, x, y, value, sd, inflation,
                                                   localization_radius, record, mean_field);
        else
            hzl_amd::Enkf<double>(e, st).assimilate(nobs, x, y, value, sd, inflation,
                                                    localization_radius, record, mean_field);
    });
}

int hzl_ensemble_mean_spread(hzl_ensemble_t h, double *spread, double *mean_field) {
    return guarded([&] {
        hzl_amd::EngineBase &e = engine_of(h);
        const hzl_amd::EngineViews v = e.views();
        if (v.B < 2)
            throw std::invalid_argument("spread needs at least two members");
        hzl_amd::EnkfState scratch(HZL_ENKF_SERIAL, 0);
        double s = 0;
        if (v.precision == HZL_FP32) {
            hzl_amd::Enkf<float> k(e, scratch);
            k.mean();
            s = k.spread();
            if (mean_field)
                k.mean_field(mean_field);
        } else {
            hzl_amd::Enkf<double> k(e, scratch);
            k.mean();
            s = k.spread();
            if (mean_field)
                k.mean_field(mean_field);
        }
        if (spread)
            *spread = s;
    });
}

} // extern "C"
