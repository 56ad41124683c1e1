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
//     and S = L L^T factorised and solved in FP64 by one workgroup.
// Serial mode optionally tapers each covariance with the Gaspari-Cohn
// function of the periodic grid distance to the observation, in physical
// space (src/filter.cpp:290-308).
// Observation perturbations eps come from the reference NormalStream
// (mt19937_64 seeded stream_seed(seed, M), Box-Muller with a cached second
// draw; src/solver.cpp:41-54), consumed observation-major, member-minor as in
// src/filter.cpp:311-313, drawn on the device (normal_stream_body).  Point
// predictions are direct evaluations of the retained Fourier sum at the grid
// point (exactly the physical field there).
//
// A cycle (forecast steps, inflation, forecast sensor RMSE, perturbations,
// analysis, posterior mean and spread) is one launch sequence on the engine's
// stream reading the observation group from device memory: after one pinned
// upload of the group it runs, or replays as a captured HIP graph, with one
// host synchronisation at its end for the two record values.
#include "engine.hpp"
#include "gemm_launch.hpp"
#include "tables.hpp"

#include <cmath>
#include <cstring>
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
    k_enkf_innov(const double *pred, int B, const double *value_o, const double *sd_o,
                 const double *eps, double *delta, double *gain) {
    __shared__ double red[256];
    const double value = *value_o, sd = *sd_o;
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
                                  const double *sn, const int *xo_p, const int *yo_p,
                                  double radius, double *phys) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)n * n)
        return;
    const int xo = *xo_p, yo = *yo_p;
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

// Point-evaluation tables of the staged observations: ex[o][c] =
// e^{2 pi i (c x_o mod n) / n}, ey[o][j] = e^{2 pi i (ky y_o mod n) / n},
// ky = j - K, read from cos / sin tables of 2 pi t / n.
__global__ void k_enkf_phase(const int *ox, const int *oy, int nobs, int K, int n,
                             const double *cs, const double *sn, cx<double> *ex,
                             cx<double> *ey) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const int J = 2 * K + 1;
    if (i >= (long)nobs * (K + 1 + J))
        return;
    const int o = (int)(i / (K + 1 + J)), r = (int)(i % (K + 1 + J));
    if (r <= K) {
        const int t = (int)(((long)r * ox[o]) % n);
        ex[(long)o * (K + 1) + r] = {cs[t], sn[t]};
    } else {
        const long ky = r - (K + 1) - K;
        const int t = (int)(((ky * oy[o]) % n + n) % n);
        ey[(long)o * J + (r - K - 1)] = {cs[t], sn[t]};
    }
}

// rec[0] = sqrt(mean over o of (pred[o] - value[o])^2): the ensemble mean's
// forecast sensor RMSE (src/filter.cpp:254-264); clears the cycle's flags.
__global__ void __launch_bounds__(256)
    k_enkf_rmse(const double *pred, const double *value, int nobs, double *rec) {
    __shared__ double red[256];
    double s = 0;
    for (int o = threadIdx.x; o < nobs; o += 256) {
        const double d = pred[o] - value[o];
        s += d * d;
    }
    const double t = block_sum(s, red);
    if (threadIdx.x == 0) {
        rec[0] = sqrt(t / nobs);
        rec[2] = 0; // the factorisation's failure flag of this cycle
        rec[3] = 0;
    }
}

// *out = sqrt(sum of part[0..count) / (B - 1)): the posterior spread from the
// per-chunk, per-mode squared anomalies (src/filter.cpp:343-347).
__global__ void __launch_bounds__(256)
    k_enkf_spread_final(const double *part, long count, int B, double *out) {
    __shared__ double red[256];
    double s = 0;
    for (long e = threadIdx.x; e < count; e += 256)
        s += part[e];
    const double t = block_sum(s, red);
    if (threadIdx.x == 0)
        *out = sqrt(t / (B - 1));
}

// One workgroup continues the filter's NormalStream (normal_stream_body).
__global__ void __launch_bounds__(256)
    k_normal_stream(uint64_t *mt, double *spare, long count, double *out) {
    __shared__ uint64_t s[MT_N];
    normal_stream_body(DevCtx{}, mt, spare, count, out, s, 256);
}

// Batch mode: observation operator rows H[o][2e], H[o][2e + 1] with
// Re(w e^{i phi}) = w.x cos - w.y sin, weighted 2 for kx > 0 modes.
__global__ void k_enkf_build_H(const cx<double> *ex, const cx<double> *ey, int m, int K,
                               float *H) {
    const int J = 2 * K + 1;
    const long spec = (long)(K + 1) * J, i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)m * spec)
        return;
    const int o = (int)(i / spec);
    const long e = i % spec;
    const int c = (int)(e / J), j = (int)(e % J);
    const cx<double> a = ex[(long)o * (K + 1) + c], b = ey[(long)o * J + j];
    const double wgt = c ? 2.0 : 1.0, pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
    H[(long)o * 2 * spec + 2 * e] = (float)(wgt * pr);
    H[(long)o * 2 * spec + 2 * e + 1] = (float)(-wgt * pi);
}

// Batch mode, one workgroup: L L^T = (S + S^T) / 2 + diag(sd^2) in FP64
// (lower triangle of L, row-major m x m; right-looking, one column per step).
// flag[0] is set when a pivot is not positive.
__global__ void __launch_bounds__(256)
    k_enkf_cholesky(const float *S, const double *sd, int m, double *L, double *flag) {
    __shared__ double pivot;
    const int t = threadIdx.x;
    for (long idx = t; idx < (long)m * m; idx += 256) {
        const int i = (int)(idx / m), j = (int)(idx % m);
        L[idx] = j <= i ? 0.5 * ((double)S[idx] + (double)S[(long)j * m + i]) +
                              (i == j ? sd[i] * sd[i] : 0.0)
                        : 0.0;
    }
    __syncthreads();
    for (int k = 0; k < m; ++k) {
        if (t == 0) {
            const double d = L[(long)k * m + k];
            if (!(d > 0))
                flag[0] = 1;
            pivot = d > 0 ? sqrt(d) : 1.0;
            L[(long)k * m + k] = pivot;
        }
        __syncthreads();
        const double p = pivot;
        for (int i = k + 1 + t; i < m; i += 256)
            L[(long)i * m + k] /= p;
        __syncthreads();
        for (int i = k + 1 + t; i < m; i += 256) {
            const double lik = L[(long)i * m + k];
            for (int j = k + 1; j <= i; ++j)
                L[(long)i * m + j] -= lik * L[(long)j * m + k];
        }
        __syncthreads();
    }
}

// Batch mode, one thread per member column: Z[:, i] = S^-1 Dm[:, i] by
// forward and back substitution with L (FP64, work in zw[m][B]).
__global__ void __launch_bounds__(256)
    k_enkf_chol_solve(const double *L, int m, const float *Dm, int B, double *zw, float *Z) {
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i >= B)
        return;
    for (int r = 0; r < m; ++r) {
        double s = (double)Dm[(long)r * B + i];
        for (int k = 0; k < r; ++k)
            s -= L[(long)r * m + k] * zw[(long)k * B + i];
        zw[(long)r * B + i] = s / L[(long)r * m + r];
    }
    for (int r = m - 1; r >= 0; --r) {
        double s = zw[(long)r * B + i];
        for (int k = r + 1; k < m; ++k)
            s -= L[(long)k * m + r] * zw[(long)k * B + i];
        zw[(long)r * B + i] = s / L[(long)r * m + r];
        Z[(long)r * B + i] = (float)zw[(long)r * B + i];
    }
}

// Filter state kept between observation groups: the perturbation stream
// (device mt19937_64 words, position, kept second normal), the device buffers
// of every cycle, and the captured cycle graph.
struct EnkfState : ModuleState {
    int mode;
    Buf<uint64_t> mt;   // 312 words + position
    Buf<double> spare;  // has_spare, spare
    Buf<cx<double>> ex, ey, partial, cov, mean;
    Buf<double> pred, pred2, delta, gain, eps, spart, cs, sn, phys, obs, rec;
    Buf<cx<double>> mixed, fcov, ffield;
    // batch mode (FP32 products, FP64 factorisation)
    Buf<float> H, HX, HA, P, S, Dm, Z, HXpart;
    Buf<double> L, zw;
    // pinned staging of one observation group and of the record values
    double *host_obs = nullptr, *host_rec = nullptr;
    size_t host_obs_cap = 0;
    hipEvent_t uploaded = nullptr; // the last group's upload out of host_obs
    // cycle graph: captured on the second consecutive call with the same key
    struct Key {
        uint64_t epoch, generation;
        double dt, inflation, radius;
        size_t steps, nobs;
        bool operator==(const Key &o) const {
            return epoch == o.epoch && generation == o.generation && dt == o.dt &&
                   inflation == o.inflation && radius == o.radius && steps == o.steps &&
                   nobs == o.nobs;
        }
    };
    hipGraphExec_t exec = nullptr;
    Key exec_key{}, warm_key{};
    bool warm = false;
    hipStream_t stream = nullptr;

    EnkfState(int mode_, uint64_t seed, hipStream_t s) : mode(mode_), stream(s) {
        // std::mt19937_64(seed): the standard seeding recurrence, position 312
        std::vector<uint64_t> w(MT_N + 1);
        w[0] = seed;
        for (int i = 1; i < MT_N; ++i)
            w[i] = 6364136223846793005ULL * (w[i - 1] ^ (w[i - 1] >> 62)) + (uint64_t)i;
        w[MT_N] = MT_N;
        const double sp[2] = {0, 0};
        mt.alloc(w.size());
        spare.alloc(2);
        hcheck(hipMemcpyAsync(mt.p, w.data(), w.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                              s),
               "EnKF stream");
        hcheck(hipMemcpyAsync(spare.p, sp, sizeof sp, hipMemcpyHostToDevice, s), "EnKF stream");
        hcheck(hipStreamSynchronize(s), "EnKF stream");
        hcheck(hipHostMalloc(reinterpret_cast<void **>(&host_rec), 4 * sizeof(double)),
               "hipHostMalloc");
    }
    EnkfState(int mode_, hipStream_t s) : mode(mode_), stream(s) {} // mean / spread only
    ~EnkfState() override {
        if (stream)
            (void)hipStreamSynchronize(stream);
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (uploaded)
            (void)hipEventDestroy(uploaded);
        if (host_obs)
            (void)hipHostFree(host_obs);
        if (host_rec)
            (void)hipHostFree(host_rec);
    }
};

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

    // the staged group: values, sds, then x and y as ints
    const double *value_d() const { return st.obs.p; }
    const double *sd_d(size_t nobs) const { return st.obs.p + nobs; }
    const int *x_d(size_t nobs) const { return reinterpret_cast<const int *>(st.obs.p + 2 * nobs); }
    const int *y_d(size_t nobs) const { return x_d(nobs) + nobs; }

    void mean() {
        st.mean.alloc((size_t)spec);
        st.partial.alloc((size_t)CHUNKS * spec);
        hipLaunchKernelGGL(k_enkf_mean_partial<R>, dim3((unsigned)((spec + 255) / 256), CHUNKS),
                           dim3(256), 0, ev.stream, w, ev.B, spec, CHUNKS, st.partial.p);
        hipLaunchKernelGGL(k_enkf_cov_final, grid1d(spec), dim3(256), 0, ev.stream, st.partial.p,
                           CHUNKS, spec, st.mean.p);
    }

    // posterior spread into *out (device), after mean() (src/filter.cpp:343-347)
    void spread_into(double *out) {
        st.spart.alloc((size_t)CHUNKS * spec);
        hipLaunchKernelGGL(k_enkf_spread_partial<R>, dim3((unsigned)((spec + 255) / 256), CHUNKS),
                           dim3(256), 0, ev.stream, w, ev.B, spec, K, st.mean.p, CHUNKS,
                           st.spart.p);
        hipLaunchKernelGGL(k_enkf_spread_final, dim3(1), dim3(256), 0, ev.stream, st.spart.p,
                           (long)CHUNKS * spec, ev.B, out);
    }

    // posterior spread to the host (after mean())
    double spread() {
        st.rec.alloc(4);
        spread_into(st.rec.p + 1);
        double v = 0;
        hcheck(hipMemcpyAsync(&v, st.rec.p + 1, sizeof v, hipMemcpyDeviceToHost, ev.stream),
               "spread download");
        hcheck(hipStreamSynchronize(ev.stream), "spread");
        return v;
    }

    void mean_field(double *field) {
        std::vector<cx<double>> h((size_t)spec);
        hcheck(hipMemcpyAsync(h.data(), st.mean.p, spec * sizeof(cx<double>),
                              hipMemcpyDeviceToHost, ev.stream),
               "mean download");
        hcheck(hipStreamSynchronize(ev.stream), "mean");
        half_to_full<double>(ev.n, h.data(), field);
    }

    // Every device buffer a cycle of `nobs` observations touches (allocations
    // happen here, never inside a launch sequence).
    void reserve(size_t nobs) {
        const size_t B = (size_t)ev.B, n = (size_t)ev.n;
        st.obs.alloc(3 * nobs + 1);
        st.ex.alloc(nobs * (K + 1));
        st.ey.alloc(nobs * J);
        st.eps.alloc(nobs * B);
        st.pred.alloc(std::max(nobs, B));
        st.rec.alloc(4);
        st.mean.alloc((size_t)spec);
        st.partial.alloc((size_t)CHUNKS * spec);
        st.spart.alloc((size_t)CHUNKS * spec);
        if (!st.cs.p) {
            std::vector<double> c(n), sv(n);
            for (size_t j = 0; j < n; ++j) {
                c[j] = std::cos(2 * PI * (double)j / (double)n);
                sv[j] = std::sin(2 * PI * (double)j / (double)n);
            }
            st.cs.alloc(n);
            st.sn.alloc(n);
            hcheck(hipMemcpyAsync(st.cs.p, c.data(), n * sizeof(double), hipMemcpyHostToDevice,
                                  ev.stream),
                   "EnKF tables");
            hcheck(hipMemcpyAsync(st.sn.p, sv.data(), n * sizeof(double), hipMemcpyHostToDevice,
                                  ev.stream),
                   "EnKF tables");
            hcheck(hipStreamSynchronize(ev.stream), "EnKF tables");
        }
        if (st.mode == HZL_ENKF_SERIAL) {
            st.pred2.alloc(B);
            st.delta.alloc(B);
            st.gain.alloc(B);
            st.cov.alloc((size_t)spec);
            st.mixed.alloc((size_t)(K + 1) * n);
            st.phys.alloc(n * n);
            if (ev.ou_F) {
                st.fcov.alloc((size_t)ev.npairs);
                st.ffield.alloc((size_t)spec);
            }
        } else {
            const size_t m = nobs, D = 2 * (size_t)spec, KC = 256;
            const size_t slices = D / KC + (D % KC ? 1 : 0);
            st.H.alloc(m * D);
            st.HX.alloc(B * m);
            st.HA.alloc(B * m);
            st.Dm.alloc(m * B);
            st.P.alloc(D * m);
            st.S.alloc(m * m);
            st.Z.alloc(m * B);
            st.HXpart.alloc(slices * B * m);
            st.L.alloc(m * m);
            st.zw.alloc(m * B);
        }
        if (st.host_obs_cap < 3 * nobs + 1) {
            hcheck(hipStreamSynchronize(ev.stream), "staging growth");
            if (st.host_obs)
                (void)hipHostFree(st.host_obs);
            st.host_obs = nullptr;
            hcheck(hipHostMalloc(reinterpret_cast<void **>(&st.host_obs),
                                 (3 * nobs + 1) * sizeof(double)),
                   "hipHostMalloc");
            st.host_obs_cap = 3 * nobs + 1;
        }
    }

    // One pinned upload of the group: [values | sds | x, y as ints].
    void stage(size_t nobs, const int *x, const int *y, const double *value, const double *sd) {
        if (!st.uploaded)
            hcheck(hipEventCreateWithFlags(&st.uploaded, hipEventDisableTiming), "event");
        else
            hcheck(hipEventSynchronize(st.uploaded), "staging reuse"); // done long ago
        std::memcpy(st.host_obs, value, nobs * sizeof(double));
        std::memcpy(st.host_obs + nobs, sd, nobs * sizeof(double));
        int *xy = reinterpret_cast<int *>(st.host_obs + 2 * nobs);
        std::memcpy(xy, x, nobs * sizeof(int));
        std::memcpy(xy + nobs, y, nobs * sizeof(int));
        hcheck(hipMemcpyAsync(st.obs.p, st.host_obs, 3 * nobs * sizeof(double),
                              hipMemcpyHostToDevice, ev.stream),
               "observation upload");
        hcheck(hipEventRecord(st.uploaded, ev.stream), "event");
    }

    // The launches of one cycle (capturable: device buffers only).
    void launches(size_t nobs, double inflation, double radius, double dt, size_t steps) {
        const int B = ev.B;
        if (steps)
            eng.enqueue_steps(dt, steps);
        const long tab = (long)nobs * (K + 1 + J);
        hipLaunchKernelGGL(k_enkf_phase, grid1d(tab), dim3(256), 0, ev.stream, x_d(nobs),
                           y_d(nobs), (int)nobs, K, ev.n, st.cs.p, st.sn.p, st.ex.p, st.ey.p);
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
        hipLaunchKernelGGL(k_enkf_predict<double>, dim3(1, (unsigned)nobs), dim3(256), 0,
                           ev.stream, st.mean.p, spec, K, st.ex.p, st.ey.p, st.pred.p, 1L);
        hipLaunchKernelGGL(k_enkf_rmse, dim3(1), dim3(256), 0, ev.stream, st.pred.p, value_d(),
                           (int)nobs, st.rec.p);
        // perturbations, observation-major then member
        hipLaunchKernelGGL(k_normal_stream, dim3(1), dim3(256), 0, ev.stream, st.mt.p,
                           st.spare.p, (long)(nobs * (size_t)B), st.eps.p);
        if (st.mode == HZL_ENKF_SERIAL)
            serial(nobs, radius);
        else
            batch(nobs);
        mean();
        spread_into(st.rec.p + 1);
    }

    void assimilate(size_t nobs, const int *x, const int *y, const double *value,
                    const double *sd, double inflation, double radius, double dt, size_t steps,
                    double *record, double *mean_out) {
        reserve(nobs);
        stage(nobs, x, y, value, sd);
        if (steps)
            eng.prepare_steps(dt);
        const EnkfState::Key key{eng.launch_epoch(), alloc_generation(), dt,   inflation,
                                 radius,             steps,              nobs};
        const bool graphs = eng.graphs_enabled();
        if (graphs && st.exec && st.exec_key == key) {
            hcheck(hipGraphLaunch(st.exec, ev.stream), "EnKF graph launch");
        } else if (graphs && st.warm && st.warm_key == key) {
            if (st.exec) {
                (void)hipGraphExecDestroy(st.exec);
                st.exec = nullptr;
            }
            hipGraph_t g = nullptr;
            hcheck(hipStreamBeginCapture(ev.stream, hipStreamCaptureModeThreadLocal), "capture");
            try {
                launches(nobs, inflation, radius, dt, steps);
            } catch (...) {
                (void)hipStreamEndCapture(ev.stream, &g);
                if (g)
                    (void)hipGraphDestroy(g);
                throw;
            }
            hcheck(hipStreamEndCapture(ev.stream, &g), "capture end");
            const hipError_t e = hipGraphInstantiate(&st.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hcheck(e, "EnKF graph instantiate");
            st.exec_key = key;
            st.warm = false;
            hcheck(hipGraphLaunch(st.exec, ev.stream), "EnKF graph launch");
        } else {
            launches(nobs, inflation, radius, dt, steps);
            st.warm = graphs;
            // the key after this run's allocations (none happen in a replay)
            st.warm_key = EnkfState::Key{eng.launch_epoch(), alloc_generation(), dt, inflation,
                                         radius, steps, nobs};
        }
        hcheck(hipGetLastError(), "EnKF kernels");
        if (steps)
            eng.commit_steps(dt, steps);
        eng.invalidate_mixed();
        hcheck(hipMemcpyAsync(st.host_rec, st.rec.p, 4 * sizeof(double), hipMemcpyDeviceToHost,
                              ev.stream),
               "record download");
        std::vector<cx<double>> h;
        if (mean_out) {
            h.resize((size_t)spec);
            hcheck(hipMemcpyAsync(h.data(), st.mean.p, spec * sizeof(cx<double>),
                                  hipMemcpyDeviceToHost, ev.stream),
                   "mean download");
        }
        hcheck(hipStreamSynchronize(ev.stream), "EnKF cycle"); // the cycle's one wait
        if (st.host_rec[2] != 0)
            throw std::runtime_error("EnKF innovation covariance is not positive definite");
        record[0] = st.host_rec[0];
        record[1] = st.host_rec[1];
        if (mean_out)
            half_to_full<double>(ev.n, h.data(), mean_out);
    }

    // taper the half-plane FP64 field cov in place around observation o
    void localize(cx<double> *cov, size_t nobs, size_t o, double radius) {
        const int n = ev.n;
        const long mix = (long)(K + 1) * n;
        hipLaunchKernelGGL(k_loc_inv_y, grid1d(mix), dim3(256), 0, ev.stream, cov, K, n,
                           st.cs.p, st.sn.p, st.mixed.p);
        hipLaunchKernelGGL(k_loc_inv_x_taper, grid1d((long)n * n), dim3(256), 0, ev.stream,
                           st.mixed.p, K, n, st.cs.p, st.sn.p, x_d(nobs) + o, y_d(nobs) + o,
                           radius, st.phys.p);
        hipLaunchKernelGGL(k_loc_fwd_x, grid1d(mix), dim3(256), 0, ev.stream, st.phys.p, K, n,
                           st.cs.p, st.sn.p, st.mixed.p);
        hipLaunchKernelGGL(k_loc_fwd_y, grid1d(spec), dim3(256), 0, ev.stream, st.mixed.p, K, n,
                           st.cs.p, st.sn.p, cov);
    }

    void serial(size_t nobs, double radius) {
        const int B = ev.B, chunks = CHUNKS;
        double *cur = st.pred.p, *next = st.pred2.p;
        hipLaunchKernelGGL(k_enkf_predict<R>, dim3((unsigned)B, 1), dim3(256), 0, ev.stream, w,
                           spec, K, st.ex.p, st.ey.p, cur, 0L);
        for (size_t o = 0; o < nobs; ++o) {
            hipLaunchKernelGGL(k_enkf_innov, dim3(1), dim3(256), 0, ev.stream, cur, B,
                               value_d() + o, sd_d(nobs) + o, st.eps.p + o * (size_t)B,
                               st.delta.p, st.gain.p);
            hipLaunchKernelGGL(k_enkf_cov_partial<R>, dim3((unsigned)((spec + 255) / 256), chunks),
                               dim3(256), 0, ev.stream, w, B, spec, st.delta.p, chunks,
                               st.partial.p);
            hipLaunchKernelGGL(k_enkf_cov_final, grid1d(spec), dim3(256), 0, ev.stream,
                               st.partial.p, chunks, spec, st.cov.p);
            if (radius > 0)
                localize(st.cov.p, nobs, o, radius);
            if (ev.ou_F) {
                // the forcing state: its covariance with the prediction (all
                // members before any update), tapered like the state's
                const int P = ev.npairs;
                hipLaunchKernelGGL(k_ou_cov, grid1d(P), dim3(256), 0, ev.stream, ev.ou_F, B, P,
                                   st.delta.p, st.fcov.p);
                if (radius > 0) {
                    hipLaunchKernelGGL(k_zero_cx, grid1d(spec), dim3(256), 0, ev.stream,
                                       st.ffield.p, spec);
                    hipLaunchKernelGGL(k_ou_scatter, grid1d(P), dim3(256), 0, ev.stream,
                                       st.fcov.p, P, ev.ou_off, K, st.ffield.p);
                    localize(st.ffield.p, nobs, o, radius);
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

    void batch(size_t nobs);
};

template <> void Enkf<double>::batch(size_t) {
    throw std::invalid_argument("batch EnKF analysis needs an FP32 ensemble");
}

template <> void Enkf<float>::batch(size_t nobs) {
    const int B = ev.B, m = (int)nobs;
    const long D = 2 * spec; // real components per member
    hipLaunchKernelGGL(k_enkf_build_H, grid1d((long)m * spec), dim3(256), 0, ev.stream, st.ex.p,
                       st.ey.p, m, K, st.H.p);
    const Span<const float> X{reinterpret_cast<const float *>(w), (size_t)B * D};
    const Span<void> Xw{reinterpret_cast<float *>(w), (size_t)B * D};
    // HX = X H^T, split along the long K = 2 spec axis: full slices of KC, a
    // remainder slice, then a sum over slices
    {
        const long KC = 256, full = D / KC, rest = D - full * KC, slices = full + (rest ? 1 : 0);
        const long count = (long)B * m;
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
                       st.HX.p, B, m, value_d(), sd_d(nobs), st.eps.p, st.HA.p, st.Dm.p);
    // P = X^T HA / (M - 1)
    Gemm<float> g2((int)D, m, B, X, 1, D, cs(st.HA), 1, m, vs(st.P), m, 1, OUT_F32);
    g2.a.alpha = 1.0f / (float)(B - 1);
    g2.run(ev.stream);
    // S = HA^T HA / (M - 1)
    Gemm<float> g3(m, m, B, cs(st.HA), 1, m, cs(st.HA), 1, m, vs(st.S), m, 1, OUT_F32);
    g3.a.alpha = 1.0f / (float)(B - 1);
    g3.run(ev.stream);
    // Z = S^-1 Dm with S + diag(sd^2) = L L^T in FP64
    hipLaunchKernelGGL(k_enkf_cholesky, dim3(1), dim3(256), 0, ev.stream, st.S.p, sd_d(nobs), m,
                       st.L.p, st.rec.p + 2);
    hipLaunchKernelGGL(k_enkf_chol_solve, grid1d(B), dim3(256), 0, ev.stream, st.L.p, m,
                       st.Dm.p, B, st.zw.p, st.Z.p);
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
        e.enkf.reset(); // a running filter's graph and stream state go first
        e.enkf = std::make_unique<hzl_amd::EnkfState>(
            mode, hzl_amd::stream_seed(seed, (uint64_t)v.B), v.stream);
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
        if (v.precision == HZL_FP32)
            hzl_amd::Enkf<float>(e, st).assimilate(nobs, x, y, value, sd, inflation,
                                                   localization_radius, dt, steps, record,
                                                   mean_field);
        else
            hzl_amd::Enkf<double>(e, st).assimilate(nobs, x, y, value, sd, inflation,
                                                    localization_radius, dt, steps, record,
                                                    mean_field);
    });
}

int hzl_ensemble_mean_spread(hzl_ensemble_t h, double *spread, double *mean_field) {
    return guarded([&] {
        hzl_amd::EngineBase &e = engine_of(h);
        const hzl_amd::EngineViews v = e.views();
        if (v.B < 2)
            throw std::invalid_argument("spread needs at least two members");
        hzl_amd::EnkfState scratch(HZL_ENKF_SERIAL, v.stream);
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
