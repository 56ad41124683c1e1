// surrogate_kernels.hpp -- elementwise and reduction bodies of the learned
// forecast surrogate (spectral neural operator).  GEMM work lives in
// gemm_mfma.hpp; these are the glue kernels around it.
#pragma once
#include "gemm_mfma.hpp"
#include "kernels.hpp"

namespace hzl_amd {

// Input features of pixel p of sample s (layout [s][p][CIN]):
//   0: w(x, y) * in_scale (already written by the state->grid GEMM)
//   1: (log nu_s - log 1e-3)
//   2: (amp_s / FNO_AMP_REF) sin(kf y), the sample's Kolmogorov forcing
//      (amp null: FNO_AMP_REF for every sample)
constexpr double FNO_AMP_REF = 0.1;

// Features 1 and 2 of sample s, row y, before their rounding to the working type.
HZL_HD float feature_viscosity(double nu) {
    return (float)(log(nu) - log(1e-3));
}
HZL_HD float feature_forcing(const double *amp, long s, int kf, int y, int G) {
    const double a = amp ? amp[s] / FNO_AMP_REF : 1.0;
    return (float)(a * sin(2.0 * PI * kf * y / G));
}

template <typename T>
HZL_HD void feature_body(T *feat, const double *nu, const double *amp, int S, int G, int CIN,
                         int kf, long i) {
    const long P = (long)G * G;
    if (i >= S * P)
        return;
    const long s = i / P, p = i % P;
    const int y = (int)(p / G);
    feat[i * CIN + 1] = (T)feature_viscosity(nu[s]);
    feat[i * CIN + 2] = (T)feature_forcing(amp, s, kf, y, G);
}

// g = g_out * gelu'(pre) elementwise
template <typename T>
HZL_HD void gelu_back_body(const T *gout, const T *pre, T *g, long n, long i) {
    if (i < n)
        g[i] = (T)((float)gout[i] * gelu_grad((float)pre[i]));
}

// Head: out[m] = sum_j h[m][j] q2[j] + b2 (h after GELU).
template <typename T>
HZL_HD void head_body(const T *h, const float *q2, float b2, float *out, long M, int H, long m) {
    if (m >= M)
        return;
    float s = b2;
    for (int j = 0; j < H; ++j)
        s += (float)h[m * H + j] * q2[j];
    out[m] = s;
}

// Head backward: dhpre[m][j] = dout[m] q2[j] gelu'(hpre[m][j]).
template <typename T>
HZL_HD void head_back_body(const float *dout, const float *q2, const T *hpre, T *dhpre, long M,
                           int H, long i) {
    if (i >= M * H)
        return;
    const long m = i / H;
    const int j = (int)(i % H);
    dhpre[i] = (T)(dout[m] * q2[j] * gelu_grad((float)hpre[i]));
}

// Mean squared error and its gradient: loss partials per block.
HZL_HD void mse_body(const float *pred, const float *target, float *grad, double *partial,
                     long n, double scale, long i) {
    if (i < n) {
        const float d = pred[i] - target[i];
        grad[i] = (float)(2.0 * scale) * d;
        partial[i] = (double)d * d;
    }
}

// Column sums of a [M][N] matrix (row stride ld) split over `chunks` row
// ranges: out[c][n] partial sums.  Reduced afterwards by sum_chunks_body.
template <typename T>
HZL_HD void colsum_body(const T *x, long M, int N, long ld, int chunks, float *out, long i) {
    if (i >= (long)chunks * N)
        return;
    const int c = (int)(i / N), n = (int)(i % N);
    const long per = (M + chunks - 1) / chunks, lo = c * per, hi = lo + per < M ? lo + per : M;
    float s = 0;
    for (long m = lo; m < hi; ++m)
        s += (float)x[m * ld + n];
    out[i] = s;
}

// FP64 partial sums of x over `chunks` contiguous ranges: out[c].
HZL_HD void dsum_body(const double *x, long n, int chunks, double *out, long c) {
    if (c >= chunks)
        return;
    const long per = (n + chunks - 1) / chunks, lo = c * per, hi = lo + per < n ? lo + per : n;
    double s = 0;
    for (long i = lo; i < hi; ++i)
        s += x[i];
    out[c] = s;
}

// out[e] = sum_c part[c][e] (+ accumulate into out when add != 0)
HZL_HD void sum_chunks_body(const float *part, int chunks, long count, float *out, int add,
                            long e) {
    if (e >= count)
        return;
    float s = add ? out[e] : 0.f;
    for (int c = 0; c < chunks; ++c)
        s += part[(long)c * count + e];
    out[e] = s;
}

// Real block form of the spectral weights of one mode:
//   Rcat[mode][ri_o*C + co][ri_i*C + ci] = [[Rr, -Ri], [Ri, Rr]].
template <typename T>
HZL_HD void rcat_body(const float *Rr, const float *Ri, T *Rcat, int modes, int C, long i) {
    const long per = 4L * C * C;
    if (i >= modes * per)
        return;
    const long mode = i / per, r = i % per;
    const int n = (int)(r / (2 * C)), k = (int)(r % (2 * C));
    const int ro = n / C, co = n % C, ri = k / C, ci = k % C;
    const long w = mode * (long)C * C + (long)co * C + ci;
    float v;
    if (ro == ri)
        v = Rr[w];
    else
        v = ro == 0 ? -Ri[w] : Ri[w];
    Rcat[i] = (T)v;
}

// Fold the gradient of Rcat back onto (Rr, Ri).
HZL_HD void rcat_back_body(const float *dRcat, float *dRr, float *dRi, int modes, int C, long i) {
    if (i >= (long)modes * C * C)
        return;
    const long mode = i / ((long)C * C), r = i % ((long)C * C);
    const int co = (int)(r / C), ci = (int)(r % C);
    const float *g = dRcat + mode * 4L * C * C;
    const long ld = 2 * C;
    const float g00 = g[(long)co * ld + ci], g11 = g[(long)(C + co) * ld + C + ci];
    const float g01 = g[(long)co * ld + C + ci], g10 = g[(long)(C + co) * ld + ci];
    dRr[i] = g00 + g11;
    dRi[i] = g10 - g01;
}

// Adam (Kingma & Ba 2015) on FP32 master weights, with bias correction.
HZL_HD void adam_body(float *w, const float *g, float *m, float *v, long n, float lr, float b1,
                      float b2, float eps, float c1, float c2, long i) {
    if (i >= n)
        return;
    const float gi = g[i];
    m[i] = b1 * m[i] + (1 - b1) * gi;
    v[i] = b2 * v[i] + (1 - b2) * gi * gi;
    const float mh = m[i] / c1, vh = v[i] / c2;
    w[i] -= lr * mh / (sqrtf(vh) + eps);
}

template <typename T> HZL_HD void cast_body(const float *src, T *dst, long n, long i) {
    if (i < n)
        dst[i] = (T)src[i];
}

// Gather member states src[idx[s]] (retained half planes) into a batch.
HZL_HD void gather_states_body(const cx<float> *src, const int *idx, long spec, cx<float> *dst,
                               int S, long i) {
    if (i >= (long)S * spec)
        return;
    const long s = i / spec, e = i % spec;
    dst[i] = src[(long)idx[s] * spec + e];
}

// dst = a - b (complex half planes)
HZL_HD void diff_states_body(const cx<float> *a, const cx<float> *b, cx<float> *dst, long n,
                             long i) {
    if (i < n)
        dst[i] = {a[i].x - b[i].x, a[i].y - b[i].y};
}

// Advance retained mode (c, ky) of member s by the surrogate horizon h:
//   inside the surrogate band (c <= KS, |ky| <= KS): w += scale * increment
//     (layout [s][c][ky + KS][re, im]) plus, on the forcing shell, the white
//     forcing accumulated over the horizon, sigma sqrt(h / 2) (g0 + i g1);
//   outside the band: w *= exp(-(nu k^2 + alpha) h) (linear decay).
HZL_HD void apply_increment_body(cx<float> *w0, const float *incr, int S, int K, int KS,
                                 float scale, const double *nu, const double *alpha,
                                 const double *sigma, const uint64_t *keys, const int *pair_of,
                                 int kmin, int kmax, double horizon, uint64_t counter, long s,
                                 int r) {
    const int J = 2 * K + 1, JS = 2 * KS + 1, per = (K + 1) * J;
    if (s >= S || r >= per)
        return;
    const int c = r / J, ky = r % J - K;
    if (c == 0 && ky == 0)
        return;
    cx<float> &w = w0[s * per + r];
    const int k2 = c * c + ky * ky;
    if (c > KS || ky > KS || ky < -KS) {
        const float d = (float)exp(-(nu[s] * k2 + alpha[s]) * horizon);
        w.x *= d;
        w.y *= d;
        return;
    }
    const float *v = incr + 2 * (s * (long)(KS + 1) * JS + (long)c * JS + ky + KS);
    float re = scale * v[0], im = scale * v[1];
    if (sigma[s] > 0 && k2 >= kmin * kmin && k2 <= kmax * kmax) {
        const bool flip = c == 0 && ky < 0;
        const int b = flip ? -ky : ky;
        const int p = pair_of[c * (2 * kmax + 1) + (b + kmax)];
        if (p >= 0) {
            double g0, g1;
            philox_normal2(keys[s], (uint32_t)p, counter, 0x53555252u /* "SURR" */, g0, g1);
            const double a = sigma[s] * sqrt(horizon / 2.0);
            re += (float)(a * g0);
            im += (float)(flip ? -a * g1 : a * g1);
        }
    }
    w.x += re;
    w.y += im;
}

// Ensemble spread of one retained mode e: (kx > 0 ? 2 : 1) x the member
// variance of w_e, so that the sum over modes is the grid mean square
// deviation from the ensemble mean (Parseval).
// Ensemble spread per retained mode in two launches.  The first sums, for
// mode e over the members [c per, (c + 1) per) of chunk c, the deviations
// d_m = w_m - w_0 from member 0 and their squares; the second combines the
// chunks: sum |w_m - mean|^2 = sum |d|^2 - |sum d|^2 / B, shifted by member 0
// so that a mode whose mean is large against its spread loses no digits to
// the cancellation of raw moments.  part holds 3 doubles per (chunk, mode).
HZL_HD void spread_partial_body(const cx<float> *w0, int B, long spec, int per, double *part,
                                long e, int c) {
    if (e >= spec)
        return;
    const int lo = c * per, hi = lo + per < B ? lo + per : B;
    const cx<float> ref = w0[e];
    double sr = 0, si = 0, ss = 0;
#pragma unroll 8
    for (int m = lo; m < hi; ++m) {
        const cx<float> v = w0[(long)m * spec + e];
        const double dr = (double)v.x - (double)ref.x, di = (double)v.y - (double)ref.y;
        sr += dr;
        si += di;
        ss += dr * dr + di * di;
    }
    double *p = part + 3 * ((long)c * spec + e);
    p[0] = sr;
    p[1] = si;
    p[2] = ss;
}

// Mode e's share of the squared spread: the conjugate-symmetric modes with
// kx > 0 count twice (the grid RMS over the full plane).
HZL_HD void spread_finish_body(int B, int K, int chunks, const double *part, double *out,
                               long e) {
    const int J = 2 * K + 1;
    const long spec = (long)(K + 1) * J;
    if (e >= spec)
        return;
    double sr = 0, si = 0, ss = 0;
#pragma unroll 8
    for (int c = 0; c < chunks; ++c) {
        const double *p = part + 3 * ((long)c * spec + e);
        sr += p[0];
        si += p[1];
        ss += p[2];
    }
    const double s = ss - (sr * sr + si * si) / B;
    out[e] = (e / J > 0 ? 2.0 : 1.0) * (s > 0 ? s : 0.0) / B;
}

// Weighted column sums: out[c][n] = sum over rows m of chunk c of w[m] x[m][n].
template <typename T>
HZL_HD void wcolsum_body(const T *x, const float *w, long M, int N, long ld, int chunks,
                         float *out, long i) {
    if (i >= (long)chunks * N)
        return;
    const int c = (int)(i / N), n = (int)(i % N);
    const long per = (M + chunks - 1) / chunks, lo = c * per, hi = lo + per < M ? lo + per : M;
    float s = 0;
    for (long m = lo; m < hi; ++m)
        s += w[m] * (float)x[m * ld + n];
    out[i] = s;
}

} // namespace hzl_amd
