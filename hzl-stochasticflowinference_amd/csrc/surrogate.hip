// surrogate.hip -- learned forecast surrogate: a spectral neural operator
// (Li et al., "Fourier Neural Operator for Parametric PDEs", ICLR 2021) that
// maps an ensemble member w(t) to w(t + h) for a fixed horizon h of many
// solver steps, trained and evaluated on gfx950 by this library's own kernels.
// The reference has no surrogate; its role here is to replace the particle
// forecast of the filter loops (src/filter.cpp:46-48, src/inference.cpp:555-590).
//
// Network on an internal G x G grid (G = 64 for a 128^2 state):
//   features [s][y][x][3] = (w * in_scale, log(nu / 1e-3), (amp / 0.1) sin(kf y))
//   lift   v0 = feat Lw^T + Lb                                    (3 -> C)
//   layer  v' = act(IDFT(R . DFT(v)) + v W^T + b)   act = GELU, none on the last
//          DFT keeps kx in [0, KX), ky in [-KY/2, KY/2); R[mode] is C x C complex
//   head   out = GELU(v Q1^T + q1) Q2 + q2                       (C -> H -> 1)
// `out` is the grid increment w(t + h) - w(t) scaled by out_scale.  State and
// grid are related by exact truncated DFTs: the state's modes |kx|, |ky| <= KS
// (KS = G/2 - 1) are sampled on the grid; the increment's grid DFT on the same
// band is added back; modes outside it decay linearly (exp(-(nu k^2+alpha) h)).
//
// Every product is a batched C = A B^T on MFMA (gemm_mfma.hpp): DFTs as dense
// cos/sin matrices, per-mode complex mixing in real 2C x 2C block form,
// pointwise channel mixing, and all products of the backward pass (transposes
// are stride swaps).  Training runs on FP32 operands with FP32 master weights
// (Adam); inference runs on BF16 or FP32 operands, FP32 accumulation.
#include "engine.hpp"
#include "gemm_launch.hpp"
#include "surrogate_kernels.hpp"
#include "fno_fused.hpp"
#include "tables.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace hzl_amd {

// ----------------------------------------------------------- kernels -------

// Fused projection + head (inference), 32 rows per wave.
template <typename T> __global__ void __launch_bounds__(256) k_proj_head(HeadArgs<T> a) {
    extern __shared__ unsigned char smem[];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    T *lds = reinterpret_cast<T *>(smem) + wave * GemmTile<T>::LDS_PER_WAVE;
    proj_head_wave<T>(DevCtx{}, a, (long)blockIdx.x * 4 + wave, lane, lds);
}

#define HZL_GID ((long)blockIdx.x * 256 + threadIdx.x)

template <typename T>
__global__ void k_feature(T *f, const double *nu, const double *amp, int S, int G, int CIN,
                          int kf) {
    feature_body<T>(f, nu, amp, S, G, CIN, kf, HZL_GID);
}
// Fused BF16 inference passes of a spectral layer (fno_fused.hpp): one
// workgroup per sample.
// LIFT: layer 0, whose input is lifted from the features inside the pass.
template <bool LIFT>
__global__ void __launch_bounds__(fno::THREADS) k_fno_forward(FnoForwardArgs a) {
    __shared__ __attribute__((aligned(16))) __bf16 lds[LIFT ? fno::FWD_LDS_LIFT : fno::FWD_LDS];
    fno_forward_body<LIFT>(DevCtx{}, a, lds, (long)blockIdx.x);
}
template <bool LIFT>
__global__ void __launch_bounds__(fno::THREADS) k_fno_inverse(FnoInverseArgs a) {
    __shared__ __attribute__((aligned(16))) __bf16 lds[LIFT ? fno::INV_LDS_LIFT : fno::INV_LDS];
    fno_inverse_body<LIFT>(DevCtx{}, a, lds, (long)blockIdx.x);
}
__global__ void __launch_bounds__(SG_THREADS) k_state_to_grid(StateGridArgs a) {
    __shared__ float lds[SG_LDS];
    fno_state_to_grid_body(DevCtx{}, a, lds, (long)blockIdx.x);
}
__global__ void __launch_bounds__(SG_THREADS) k_grid_to_increment(GridIncrementArgs a) {
    __shared__ float lds[GS_LDS];
    fno_grid_to_increment_body(DevCtx{}, a, lds, (long)blockIdx.x);
}
__global__ void __launch_bounds__(MIX_THREADS) k_fno_mix(FnoMixArgs a) {
    fno_mix_body(DevCtx{}, a, (long)blockIdx.x);
}
__global__ void __launch_bounds__(256) k_fno_head(FnoHeadArgs a) {
    __shared__ float hq[2 * FNO_H];
    fno_head_stage(DevCtx{}, a, hq);
    // four wavefronts of 32 points; one past the last point computes zeros, stores nothing
    fno_head_wave(DevCtx{}, a, (long)blockIdx.x * 4 + threadIdx.x / 64, (int)(threadIdx.x % 64),
                  hq);
}

template <typename T> __global__ void k_gelu_back(const T *go, const T *pre, T *g, long n) {
    gelu_back_body<T>(go, pre, g, n, HZL_GID);
}
template <typename T>
__global__ void k_head(const T *h, const float *q2, const float *b2, float *out, long M, int H) {
    head_body<T>(h, q2, b2[0], out, M, H, HZL_GID);
}
template <typename T>
__global__ void k_head_back(const float *dout, const float *q2, const T *hpre, T *dh, long M,
                            int H) {
    head_back_body<T>(dout, q2, hpre, dh, M, H, HZL_GID);
}
__global__ void k_mse(const float *p, const float *t, float *g, double *part, long n, double sc) {
    mse_body(p, t, g, part, n, sc, HZL_GID);
}
template <typename T>
__global__ void k_colsum(const T *x, long M, int N, long ld, int chunks, float *out) {
    colsum_body<T>(x, M, N, ld, chunks, out, HZL_GID);
}
template <typename T>
__global__ void k_wcolsum(const T *x, const float *w, long M, int N, long ld, int chunks,
                          float *out) {
    wcolsum_body<T>(x, w, M, N, ld, chunks, out, HZL_GID);
}
__global__ void k_sum_chunks(const float *part, int chunks, long count, float *out, int add) {
    sum_chunks_body(part, chunks, count, out, add, HZL_GID);
}
// Column sums of x (M rows, N columns with 256 % N == 0, leading dimension
// ld), each row weighted by w[m] when W: workgroup b sums rows
// [b per, (b + 1) per) into part[b][n] with 256 / N row lanes per column,
// combined in LDS in a fixed order (the result does not depend on timing).
template <typename T, bool W>
__global__ void __launch_bounds__(256)
    k_colsum_blocks(const T *x, const float *w, long M, int N, long ld, long per, float *part) {
    __shared__ float red[256];
    const int tid = threadIdx.x, n = tid % N, lane = tid / N, lanes = 256 / N;
    const long lo = (long)blockIdx.x * per, hi = lo + per < M ? lo + per : M;
    float s = 0.f;
    for (long m = lo + lane; m < hi; m += lanes)
        s += (W ? w[m] : 1.f) * (float)x[m * ld + n];
    red[tid] = s;
    __syncthreads();
    for (int half = lanes / 2; half >= 1; half /= 2) {
        if (lane < half)
            red[tid] += red[tid + half * N];
        __syncthreads();
    }
    if (lane == 0)
        part[(long)blockIdx.x * N + n] = red[tid];
}
// out[e] = sum over c < chunks of part[c][e] (+ out[e] when add): 32
// consecutive elements x 8 chunk groups per workgroup, the groups combined in
// a fixed order.
__global__ void __launch_bounds__(256)
    k_sum_chunks_blocks(const float *part, int chunks, long count, float *out, int add) {
    __shared__ float red[256];
    const int tid = threadIdx.x, el = tid % 32, g = tid / 32;
    const long e = (long)blockIdx.x * 32 + el;
    float s = 0.f;
    if (e < count)
        for (int c = g; c < chunks; c += 8)
            s += part[(long)c * count + e];
    red[tid] = s;
    __syncthreads();
    if (g == 0 && e < count) {
        float t = add ? out[e] : 0.f;
        for (int k = 0; k < 8; ++k)
            t += red[k * 32 + el];
        out[e] = t;
    }
}
template <typename T>
__global__ void k_rcat(const float *Rr, const float *Ri, T *Rcat, int modes, int C) {
    rcat_body<T>(Rr, Ri, Rcat, modes, C, HZL_GID);
}
__global__ void k_rcat_back(const float *g, float *dRr, float *dRi, int modes, int C) {
    rcat_back_body(g, dRr, dRi, modes, C, HZL_GID);
}
__global__ void k_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    adam_body(w, g, m, v, n, lr, b1, b2, eps, c1, c2, HZL_GID);
}
template <typename T> __global__ void k_cast(const float *src, T *dst, long n) {
    cast_body<T>(src, dst, n, HZL_GID);
}
__global__ void k_gather_states(const cx<float> *src, const int *idx, long spec, cx<float> *dst,
                                int S) {
    gather_states_body(src, idx, spec, dst, S, HZL_GID);
}
__global__ void k_diff_states(const cx<float> *a, const cx<float> *b, cx<float> *d, long n) {
    diff_states_body(a, b, d, n, HZL_GID);
}
// part[e] = sum over members of w_e |a_m[e] - b_m[e]|^2 (w = 2 for kx > 0):
// with the sum over e, the members' total squared grid-RMS difference.
__global__ void k_sqdiff_modes(const cx<float> *a, const cx<float> *b, int B, int K,
                               double *part) {
    const int J = 2 * K + 1;
    const long spec = (long)(K + 1) * J, e = HZL_GID;
    if (e >= spec)
        return;
    double s = 0;
    for (int m = 0; m < B; ++m) {
        const cx<float> u = a[(long)m * spec + e], v = b[(long)m * spec + e];
        const double dr = (double)u.x - (double)v.x, di = (double)u.y - (double)v.y;
        s += dr * dr + di * di;
    }
    part[e] = (e / J > 0 ? 2.0 : 1.0) * s;
}

__global__ void k_gather_nu(const double *nu, const int *idx, double *out, int S) {
    const long i = HZL_GID;
    if (i < S)
        out[i] = nu[idx[i]];
}
__global__ void k_dsum(const double *x, long n, int chunks, double *out) {
    dsum_body(x, n, chunks, out, HZL_GID);
}
__global__ void k_spread_partial(const cx<float> *w0, int B, long spec, int per, double *part) {
    spread_partial_body(w0, B, spec, per, part, (long)blockIdx.x * blockDim.x + threadIdx.x,
                        (int)blockIdx.y);
}
__global__ void k_spread_finish(int B, int K, int chunks, const double *part, double *out) {
    spread_finish_body(B, K, chunks, part, out, HZL_GID);
}
__global__ void k_fill_f32(float *p, long n, float v) {
    const long i = HZL_GID;
    if (i < n)
        p[i] = v;
}
// The noise counter is read from device memory so that a captured forecast
// graph can be replayed with a new counter.  Grid (samples, ceil(modes /
// 256)): the sample is the workgroup's x index.
__global__ void __launch_bounds__(256) k_apply_increment(cx<float> *w0, const float *incr, int S,
                                                         int K, int KS, float scale,
                                                         const double *nu, const double *alpha,
                                                         const double *sigma, const uint64_t *keys,
                                                         const int *pair_of, int kmin, int kmax,
                                                         double horizon, const uint64_t *counter) {
    apply_increment_body(w0, incr, S, K, KS, scale, nu, alpha, sigma, keys, pair_of, kmin, kmax,
                         horizon, counter[0], (long)blockIdx.x,
                         (int)(blockIdx.y * 256 + threadIdx.x));
}
__global__ void k_set_u64(uint64_t *p, uint64_t v) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        p[0] = v;
}


// ------------------------------------------------------------- model ------
struct FnoShape {
    int G = 64, C = 32, L = 4, KX = 16, KY = 32, H = 128, CIN = 3, kf = 4;
    int N = 128, K = 42; // engine grid and its retained band
    int KS = 31;         // band of the state sampled on the grid: |kx|, |ky| <= KS
    static constexpr int KXS = 32; // padded kx extent of the state <-> grid products
    int modes() const {
        return KX * KY;
    }
    long P() const {
        return (long)G * G;
    }
    long SA() const { // x-transformed activations per sample: [2KX][G][C]
        return 2L * KX * G * C;
    }
    long SV() const { // spectral activations per sample: [KX][KY][2][C]
        return (long)modes() * 2 * C;
    }
    long spec() const {
        return (long)(K + 1) * (2 * K + 1);
    }
};

// Offsets of the parameter groups inside the one FP32 parameter vector.
struct FnoParams {
    long Lw = 0, Lb = 0, Q1 = 0, q1 = 0, Q2 = 0, q2 = 0, total = 0;
    std::vector<long> W, b, Rr, Ri;
    explicit FnoParams(const FnoShape &s) {
        long o = 0;
        auto take = [&](long n) {
            const long at = o;
            o += n;
            return at;
        };
        Lw = take((long)s.C * s.CIN); // [C][CIN]
        Lb = take(s.C);
        for (int l = 0; l < s.L; ++l) {
            W.push_back(take((long)s.C * s.C)); // [C out][C in]
            b.push_back(take(s.C));
            Rr.push_back(take((long)s.modes() * s.C * s.C)); // [kx][kyi][C out][C in]
            Ri.push_back(take((long)s.modes() * s.C * s.C));
        }
        Q1 = take((long)s.H * s.C); // [H][C]
        q1 = take(s.H);
        Q2 = take(s.H); // [H]
        q2 = take(1);
        total = o;
    }
};

// Constant matrices of the truncated transforms, in operand type T.
//   Fx  [2kx+ri][x]         forward x:  (cos, -sin)(2 pi kx x / G)
//   Fy2 [2kyi+ro][ri*G+y]   forward y:  [[c, s], [-s, c]]
//   Gy2 [ro*G+y][2kyi+ri]   inverse y:  [[c, -s], [s, c]]
//   Gx  [x][2kx+ri]         inverse x:  w / G^2 (cos, -sin), w = 1 (kx = 0) else 2
// and their transposes for the backward products.
template <typename T> struct FnoConstants {
    Buf<T> Fx, FxT, Fy2, Fy2T, Gy2, Gy2T, Gx, GxT;
    void build(const FnoShape &s) {
        const int G = s.G, KX = s.KX, KY = s.KY;
        std::vector<double> fx(2 * KX * G), fy(2 * KY * 2 * G), gy(2 * G * 2 * KY), gx(G * 2 * KX);
        for (int kx = 0; kx < KX; ++kx)
            for (int x = 0; x < G; ++x) {
                const double ph = 2 * PI * kx * x / G, w = (kx == 0 ? 1.0 : 2.0) / ((double)G * G);
                fx[(2 * kx) * G + x] = std::cos(ph);
                fx[(2 * kx + 1) * G + x] = -std::sin(ph);
                gx[x * 2 * KX + 2 * kx] = w * std::cos(ph);
                gx[x * 2 * KX + 2 * kx + 1] = -w * std::sin(ph);
            }
        for (int kyi = 0; kyi < KY; ++kyi) {
            const int ky = kyi < KY / 2 ? kyi : kyi - KY;
            for (int y = 0; y < G; ++y) {
                const double th = 2 * PI * ky * y / G, c = std::cos(th), sn = std::sin(th);
                fy[(2 * kyi) * 2 * G + y] = c;
                fy[(2 * kyi) * 2 * G + G + y] = sn;
                fy[(2 * kyi + 1) * 2 * G + y] = -sn;
                fy[(2 * kyi + 1) * 2 * G + G + y] = c;
                gy[y * 2 * KY + 2 * kyi] = c;
                gy[y * 2 * KY + 2 * kyi + 1] = -sn;
                gy[(G + y) * 2 * KY + 2 * kyi] = sn;
                gy[(G + y) * 2 * KY + 2 * kyi + 1] = c;
            }
        }
        auto up = [&](Buf<T> &b, const std::vector<double> &v, int rows, int cols, bool tr) {
            std::vector<T> h(v.size());
            for (int r = 0; r < rows; ++r)
                for (int c = 0; c < cols; ++c)
                    h[tr ? (size_t)c * rows + r : (size_t)r * cols + c] =
                        (T)(float)v[(size_t)r * cols + c];
            b.alloc(h.size());
            hcheck(hipMemcpy(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice),
                   "transform constants");
        };
        up(Fx, fx, 2 * KX, G, false);
        up(FxT, fx, 2 * KX, G, true);
        up(Fy2, fy, 2 * KY, 2 * G, false);
        up(Fy2T, fy, 2 * KY, 2 * G, true);
        up(Gy2, gy, 2 * G, 2 * KY, false);
        up(Gy2T, gy, 2 * G, 2 * KY, true);
        up(Gx, gx, G, 2 * KX, false);
        up(GxT, gx, G, 2 * KX, true);
    }
};

// Activations of one batch of S samples in operand type T, plus the weights
// cast to T.  Layouts: pixel tensors [s][y][x][channel]; A/Bi [s][2kx+ri][y][c];
// V/U [s][kx][kyi][ri][c].
template <typename T> struct FnoWork {
    int S = 0;
    bool built = false;
    FnoConstants<T> cst;
    Buf<T> feat, h, hpre, dh, dv0, dv1, g, dBi, dU, dV, dA, w_T, rcat;
    std::vector<std::unique_ptr<Buf<T>>> v, pre, A, V, U, Bi;
    Buf<float> out, dout;
    Buf<double> loss_part;
    uint64_t version = ~0ULL; // parameter version w_T / rcat were cast from

    int S_train = 0;
    // Buffers for S samples; training (backward) buffers only when `train`.
    void ensure(const FnoShape &s, int S_, long nparams, bool train) {
        if (!built) {
            cst.build(s);
            for (int l = 0; l <= s.L; ++l)
                v.emplace_back(new Buf<T>);
            for (int l = 0; l < s.L; ++l) {
                pre.emplace_back(new Buf<T>);
                A.emplace_back(new Buf<T>);
                V.emplace_back(new Buf<T>);
                U.emplace_back(new Buf<T>);
                Bi.emplace_back(new Buf<T>);
            }
            w_T.alloc(nparams);
            rcat.alloc((size_t)s.L * s.modes() * 4 * s.C * s.C);
            built = true;
        }
        if (S_ > S) {
            S = S_;
            const size_t SP = (size_t)S * s.P();
            feat.alloc(SP * s.CIN);
            for (auto &b : v)
                b->alloc(SP * s.C);
            for (int l = 0; l < s.L; ++l) {
                A[l]->alloc(S * s.SA());
                V[l]->alloc(S * s.SV());
                U[l]->alloc(S * s.SV());
                Bi[l]->alloc(S * s.SA());
            }
            out.alloc(SP);
            dout.alloc(SP);
            loss_part.alloc(SP);
        }
        if (train && S_ > S_train) {
            S_train = S_;
            const size_t SP = (size_t)S_ * s.P();
            for (int l = 0; l < s.L; ++l)
                pre[l]->alloc(SP * s.C);
            h.alloc(SP * s.H);
            hpre.alloc(SP * s.H);
            dh.alloc(SP * s.H);
            dv0.alloc(SP * s.C);
            dv1.alloc(SP * s.C);
            g.alloc(SP * s.C);
            dBi.alloc(S_ * s.SA());
            dA.alloc(S_ * s.SA());
            dU.alloc(S_ * s.SV());
            dV.alloc(S_ * s.SV());
        }
    }
};

class Surrogate {
  public:
    Surrogate(int state_n, int grid, int width, int layers, int modes, uint64_t seed)
        : s_(make_shape(state_n, grid, width, layers, modes)), prm_(s_) {
        hcheck(hipStreamCreateWithFlags(&own_, hipStreamNonBlocking), "hipStreamCreate");
        st_ = own_;
        master_.alloc(prm_.total);
        grad_.alloc(prm_.total);
        m_.alloc(prm_.total);
        v_.alloc(prm_.total);
        init_params(seed);
        conv_build();
    }
    ~Surrogate() {
        if (!graphs_.empty()) {
            (void)hipDeviceSynchronize();
            for (auto &c : graphs_)
                (void)hipGraphExecDestroy(c.exec);
        }
        if (own_) {
            (void)hipStreamSynchronize(own_);
            (void)hipStreamDestroy(own_);
        }
        if (order_ev_)
            (void)hipEventDestroy(order_ev_);
        if (judge_pending_)
            (void)hipEventSynchronize(judge_done_ev_);
        for (hipEvent_t e : {judge_in_ev_, judge_copied_ev_, judge_done_ev_})
            if (e)
                (void)hipEventDestroy(e);
        if (judge_host_)
            (void)hipHostFree(judge_host_);
    }
    Surrogate(const Surrogate &) = delete;
    Surrogate &operator=(const Surrogate &) = delete;

    long param_count() const {
        return prm_.total;
    }
    void get_params(float *out) {
        hcheck(hipMemcpy(out, master_.p, prm_.total * sizeof(float), hipMemcpyDeviceToHost),
               "parameter download");
    }
    void set_params(const float *in) {
        for (long i = 0; i < prm_.total; ++i)
            if (!std::isfinite(in[i]))
                throw std::invalid_argument("nonfinite surrogate parameter");
        hcheck(hipMemcpy(master_.p, in, prm_.total * sizeof(float), hipMemcpyHostToDevice),
               "parameter upload");
        reset_adam();
    }
    void set_scales(double in_scale, double out_scale) {
        if (!(in_scale > 0) || !(out_scale > 0) || !std::isfinite(in_scale) ||
            !std::isfinite(out_scale))
            throw std::invalid_argument("surrogate scales must be positive and finite");
        in_scale_ = in_scale;
        out_scale_ = out_scale;
        scales_set_ = true;
    }
    void get_scales(double *s) const {
        s[0] = in_scale_;
        s[1] = out_scale_;
    }
    size_t dataset_size() const {
        return pairs_;
    }

    // ---- host-array entry points (reference Field layout) -----------------
    void forward_host(int precision, size_t S, const double *fields, const double *nu,
                      const double *amp, double *grid) {
        check_count(S);
        check_precision(precision);
        st_ = own_;
        upload_fields(S, fields, hx_);
        upload_conditions(S, nu, amp);
        if (precision == HZL_BF16) {
            forward(wb_, hx_.p, hnu_.p, hamp_.p, (int)S, false);
            download_out(wb_, S, grid);
        } else {
            forward(wf_, hx_.p, hnu_.p, hamp_.p, (int)S, false);
            download_out(wf_, S, grid);
        }
    }
    void loss_grad_host(size_t S, const double *fields, const double *next, const double *nu,
                        const double *amp, double *loss, float *grad) {
        check_count(S);
        st_ = own_;
        upload_fields(S, fields, hx_);
        upload_fields(S, next, hy_);
        upload_conditions(S, nu, amp);
        *loss = loss_and_grad(wf_, hx_.p, hy_.p, hnu_.p, hamp_.p, (int)S, true);
        hcheck(hipMemcpy(grad, grad_.p, prm_.total * sizeof(float), hipMemcpyDeviceToHost),
               "gradient download");
    }

    // ---- training data from an FP32 ensemble ------------------------------
    // windows x members pairs (w(t), w(t + steps dt)); the ensemble advances
    // by windows * steps solver steps with its own forcing.  With noise_free,
    // each target is the same start state advanced without the white shell
    // forcing (a scratch ensemble with the members' nu, alpha and amplitude):
    // the network then learns the deterministic propagator, the conditional
    // mean of the forced flow map, and the forecast adds the noise itself.
    void collect(EngineBase &eng, double dt, size_t steps, size_t windows, bool noise_free) {
        const EngineViews ev = check_engine(eng);
        if (steps == 0 || windows == 0 || !(dt > 0) || !std::isfinite(dt))
            throw std::invalid_argument("collect needs dt > 0, steps > 0 and windows > 0");
        if (horizon_ && (steps != horizon_ || dt != dt_))
            throw std::invalid_argument("collect horizon differs from the dataset's");
        const size_t spec = (size_t)s_.spec(), B = (size_t)ev.B;
        std::unique_ptr<hzl_ensemble, int (*)(hzl_ensemble_t)> scratch(nullptr,
                                                                       hzl_ensemble_destroy);
        if (noise_free) {
            std::vector<double> nu(B), alpha(B), amp(B);
            eng.synchronize();
            hcheck(hipMemcpy(nu.data(), ev.nu, B * sizeof(double), hipMemcpyDeviceToHost), "nu");
            hcheck(hipMemcpy(alpha.data(), ev.alpha, B * sizeof(double), hipMemcpyDeviceToHost),
                   "alpha");
            hcheck(hipMemcpy(amp.data(), ev.amp, B * sizeof(double), hipMemcpyDeviceToHost), "amp");
            hzl_parameters p{nu[0], alpha[0], 0.0, 1.0, ev.kmin, ev.kmax, amp[0], ev.kf};
            hzl_ensemble_t h = nullptr;
            if (hzl_ensemble_create(&h, (size_t)s_.N, B, HZL_FP32, &p, 0) != HZL_OK)
                throw std::runtime_error("scratch ensemble: " + error_slot());
            scratch.reset(h);
            EngineBase &se = engine_of(h);
            for (size_t m = 0; m < B; ++m)
                se.set_member_parameters(m, nu[m], alpha[m], 0.0, amp[m]);
        }
        horizon_ = steps;
        dt_ = dt;
        for (size_t wdw = 0; wdw < windows; ++wdw) {
            const EngineViews v = eng.views();
            grow_dataset(pairs_ + B);
            const size_t bytes = B * spec * sizeof(cx<float>);
            hcheck(hipMemcpyAsync(ds_x_.p + pairs_ * spec, v.w0, bytes, hipMemcpyDeviceToDevice,
                                  v.stream),
                   "dataset copy");
            hcheck(hipMemcpyAsync(ds_nu_.p + pairs_, v.nu, B * sizeof(double),
                                  hipMemcpyDeviceToDevice, v.stream),
                   "dataset copy");
            hcheck(hipMemcpyAsync(ds_amp_.p + pairs_, v.amp, B * sizeof(double),
                                  hipMemcpyDeviceToDevice, v.stream),
                   "dataset copy");
            if (scratch) {
                EngineBase &se = engine_of(scratch.get());
                const EngineViews sv = se.views();
                eng.synchronize();
                hcheck(hipMemcpyAsync(sv.w0, v.w0, bytes, hipMemcpyDeviceToDevice, sv.stream),
                       "scratch state");
                se.invalidate_mixed();
                se.step(dt, steps);
                hcheck(hipMemcpyAsync(ds_y_.p + pairs_ * spec, sv.w0, bytes,
                                      hipMemcpyDeviceToDevice, sv.stream),
                       "dataset copy");
                se.synchronize();
                eng.step(dt, steps);
            } else {
                eng.step(dt, steps);
                hcheck(hipMemcpyAsync(ds_y_.p + pairs_ * spec, v.w0, bytes,
                                      hipMemcpyDeviceToDevice, v.stream),
                       "dataset copy");
            }
            eng.synchronize();
            if (!scales_set_)
                auto_scales(pairs_, B);
            pairs_ += B;
        }
    }

    // Minibatch Adam over pairs [first, first + count) with a cosine learning
    // rate from lr down to lr / 20.
    void fit(size_t first, size_t count, size_t epochs, size_t batch, double lr, uint64_t seed,
             double *epoch_loss) {
        if (count == 0 || first + count > pairs_)
            throw std::invalid_argument("fit range outside the dataset");
        if (batch == 0 || batch > count || batch > 4096 || !(lr > 0) || epochs == 0)
            throw std::invalid_argument(
                "fit needs epochs > 0, 0 < batch <= min(count, 4096) and lr > 0");
        st_ = own_;
        const size_t nb = count / batch, total = nb * epochs;
        std::mt19937_64 rng(seed);
        std::vector<int> order(count), idx(batch);
        size_t t = 0;
        for (size_t e = 0; e < epochs; ++e) {
            for (size_t i = 0; i < count; ++i)
                order[i] = (int)(first + i);
            std::shuffle(order.begin(), order.end(), rng);
            double sum = 0;
            for (size_t b = 0; b < nb; ++b, ++t) {
                std::copy(order.begin() + b * batch, order.begin() + (b + 1) * batch, idx.begin());
                gather_batch(idx);
                sum += loss_and_grad(wf_, bx_.p, by_.p, bnu_.p, bamp_.p, (int)batch, true);
                const double f = 0.5 * (1 + std::cos(PI * (double)t / (double)total));
                adam((float)(lr * (0.05 + 0.95 * f)));
            }
            if (epoch_loss)
                epoch_loss[e] = sum / (double)nb;
        }
        hcheck(hipStreamSynchronize(st_), "fit");
    }

    // Mean loss over pairs [first, first + count) (forward only).
    double evaluate(int precision, size_t first, size_t count) {
        if (count == 0 || first + count > pairs_)
            throw std::invalid_argument("evaluate range outside the dataset");
        check_precision(precision);
        st_ = own_;
        const size_t batch = std::min<size_t>(count, 256);
        std::vector<int> idx;
        double sum = 0;
        for (size_t done = 0; done < count;) {
            const size_t S = std::min(batch, count - done);
            idx.resize(S);
            for (size_t i = 0; i < S; ++i)
                idx[i] = (int)(first + done + i);
            gather_batch(idx);
            const double l =
                precision == HZL_BF16
                    ? loss_and_grad(wb_, bx_.p, by_.p, bnu_.p, bamp_.p, (int)S, false)
                    : loss_and_grad(wf_, bx_.p, by_.p, bnu_.p, bamp_.p, (int)S, false);
            sum += l * (double)S;
            done += S;
        }
        const double loss = sum / (double)count;
        // grid RMS error of the predicted increment, in vorticity units
        error_rms_ = std::sqrt(loss) / out_scale_;
        return loss;
    }

    double error_estimate() const {
        return error_rms_;
    }
    void set_error_estimate(double e) {
        if (!(e >= 0) || !std::isfinite(e))
            throw std::invalid_argument("error estimate must be finite and nonnegative");
        error_rms_ = e;
    }

    // Grid RMS deviation of the ensemble members from their mean.
    double spread(EngineBase &eng) {
        const EngineViews ev = check_engine(eng);
        const long spec = s_.spec();
        // members in chunks of 32 (more beyond 2M members: the second grid
        // dimension holds at most 65535) over the second grid dimension:
        // enough workgroups to fill the device at a few hundred members
        const int per = std::max(32, (ev.B + 65534) / 65535);
        const int chunks = (ev.B + per - 1) / per;
        spread_.alloc((size_t)spec);
        spread_part_.alloc(3 * (size_t)chunks * (size_t)spec);
        hipLaunchKernelGGL(k_spread_partial, dim3((unsigned)((spec + 255) / 256), (unsigned)chunks),
                           dim3(256), 0, ev.stream, reinterpret_cast<const cx<float> *>(ev.w0),
                           ev.B, (long)spec, per, spread_part_.p);
        hipLaunchKernelGGL(k_spread_finish, grid1d(spec), dim3(256), 0, ev.stream, ev.B, s_.K,
                           chunks, (const double *)spread_part_.p, spread_.p);
        std::vector<double> h((size_t)spec);
        hcheck(hipMemcpyAsync(h.data(), spread_.p, spec * sizeof(double), hipMemcpyDeviceToHost,
                              ev.stream),
               "spread download");
        hcheck(hipStreamSynchronize(ev.stream), "spread");
        double sum = 0;
        for (double v : h)
            sum += v;
        return std::sqrt(sum);
    }

    // Live judge of the forecast error: the first J members of `eng` (J =
    // members of `judge`) are copied into the judge ensemble with their nu,
    // alpha and amplitude and no shell noise, forecast over the horizon by the
    // network and, from the same states, by the plain solver; the grid RMS
    // difference per member becomes the error estimate.  `eng` is unchanged.
    double judge(EngineBase &eng, EngineBase &jg, double dt, size_t steps, int precision) {
        judge_start(eng, jg, dt, steps, precision);
        judge_finish(true);
        return error_rms_;
    }

    // The same judge, enqueued: the judge ensemble's stream waits for the
    // ensemble's pending work, copies the judged states, and the ensemble's
    // stream waits only for that copy; the network forecast of the copies is
    // ordered with the ensemble's forecasts (shared activations), the solver
    // run and the comparison overlap whatever the ensemble does next.
    // judge_finish() folds the result into the error estimate.  One judge is
    // in flight at a time: a start first finishes the previous one.
    void judge_start(EngineBase &eng, EngineBase &jg, double dt, size_t steps, int precision) {
        judge_finish(true);
        const EngineViews ev = check_engine(eng);
        const size_t J = jg.members();
        if (J > (size_t)ev.B)
            throw std::invalid_argument("the judge ensemble has more members than the ensemble");
        if (jg.forcing_process() != HZL_FORCING_WHITE)
            throw std::invalid_argument("the judge ensemble runs the white forcing process");
        if (&jg == &eng)
            throw std::invalid_argument("the judge ensemble must differ from the ensemble");
        check_precision(precision);
        if (horizon_ && (steps != horizon_ || std::fabs(dt - dt_) > 1e-12 * dt_))
            throw std::invalid_argument("forecast horizon differs from the trained horizon");
        // parameters from the ensemble's host copies (what views() uploaded);
        // unchanged members are not marked for another upload
        for (size_t m = 0; m < J; ++m) {
            double nu, alpha, sigma, amp, jn, ja, js, jamp;
            eng.get_member_parameters(m, &nu, &alpha, &sigma, &amp);
            jg.get_member_parameters(m, &jn, &ja, &js, &jamp);
            if (nu != jn || alpha != ja || js != 0.0 || amp != jamp)
                jg.set_member_parameters(m, nu, alpha, 0.0, amp);
        }
        const EngineViews jv = check_engine(jg); // uploads changed parameters
        const size_t spec = (size_t)s_.spec();
        const size_t bytes = J * spec * sizeof(cx<float>);
        judge_x_.alloc(J * spec);
        judge_y_.alloc(J * spec);
        if (judge_host_n_ < spec) {
            if (judge_host_)
                hcheck(hipHostFree(judge_host_), "judge sums");
            judge_host_ = judge_host_dev_ = nullptr;
            judge_host_n_ = 0;
            hcheck(hipHostMalloc((void **)&judge_host_, spec * sizeof(double),
                                 hipHostMallocMapped | hipHostMallocCoherent),
                   "judge sums");
            judge_host_n_ = spec;
            hcheck(hipHostGetDevicePointer((void **)&judge_host_dev_, judge_host_, 0),
                   "judge sums");
        }
        for (hipEvent_t *e : {&judge_in_ev_, &judge_copied_ev_, &judge_done_ev_})
            if (!*e)
                hcheck(hipEventCreateWithFlags(e, hipEventDisableTiming), "event");
        hcheck(hipEventRecord(judge_in_ev_, ev.stream), "judge order");
        hcheck(hipStreamWaitEvent(jv.stream, judge_in_ev_, 0), "judge order");
        hcheck(hipMemcpyAsync(judge_x_.p, ev.w0, bytes, hipMemcpyDeviceToDevice, jv.stream),
               "judge states");
        hcheck(hipEventRecord(judge_copied_ev_, jv.stream), "judge order");
        hcheck(hipStreamWaitEvent(ev.stream, judge_copied_ev_, 0), "judge order");
        hcheck(hipMemcpyAsync(jv.w0, judge_x_.p, bytes, hipMemcpyDeviceToDevice, jv.stream),
               "judge states");
        jg.invalidate_mixed();
        forecast(jg, dt, steps, precision); // sigma 0: the network's forecast alone
        hcheck(hipMemcpyAsync(judge_y_.p, jv.w0, bytes, hipMemcpyDeviceToDevice, jv.stream),
               "judge forecast");
        hcheck(hipMemcpyAsync(jv.w0, judge_x_.p, bytes, hipMemcpyDeviceToDevice, jv.stream),
               "judge states");
        jg.rewind_counter(steps);
        jg.invalidate_mixed();
        jg.step(dt, steps);
        // the per-mode sums go straight to mapped host memory: a download
        // queued here behind the solver run would hold the device's other
        // host transfers (the ensemble's spread and statistics) until it ran
        hipLaunchKernelGGL(k_sqdiff_modes, grid1d((long)spec), dim3(256), 0, jv.stream,
                           reinterpret_cast<const cx<float> *>(jv.w0), judge_y_.p, (int)J, s_.K,
                           judge_host_dev_);
        hcheck(hipGetLastError(), "judge difference");
        hcheck(hipEventRecord(judge_done_ev_, jv.stream), "judge order");
        judge_members_ = J;
        judge_modes_ = spec;
        judge_pending_ = true;
    }

    // Fold a judge in flight into the error estimate: returns 1 when it did,
    // 0 when none is in flight or (wait false) it has not finished yet.
    int judge_finish(bool wait) {
        if (!judge_pending_)
            return 0;
        if (wait) {
            hcheck(hipEventSynchronize(judge_done_ev_), "judge");
        } else {
            const hipError_t e = hipEventQuery(judge_done_ev_);
            if (e == hipErrorNotReady)
                return 0;
            hcheck(e, "judge");
        }
        judge_pending_ = false;
        double sum = 0;
        for (size_t i = 0; i < judge_modes_; ++i)
            sum += judge_host_[i];
        error_rms_ = std::sqrt(sum / (double)judge_members_);
        return 1;
    }

    bool judge_in_flight() const {
        return judge_pending_;
    }

    // Forecast with the network, then compare its error estimate with the
    // spread of the forecast ensemble: when the error exceeds threshold x
    // spread, the states are restored and the plain solver advances the
    // ensemble instead (its own forcing, same noise counters as a plain step).
    // Returns 1 when the solver ran.
    int forecast_or_solve(EngineBase &eng, double dt, size_t steps, int precision,
                          double threshold, double *spread_out) {
        if (!(threshold > 0) || !std::isfinite(threshold))
            throw std::invalid_argument("fallback threshold must be positive");
        if (!(error_rms_ >= 0))
            throw std::invalid_argument(
                "no error estimate: evaluate the surrogate or set one first");
        const EngineViews ev = check_engine(eng);
        const size_t bytes = (size_t)ev.B * s_.spec() * sizeof(cx<float>);
        backup_.alloc((size_t)ev.B * s_.spec());
        hcheck(hipMemcpyAsync(backup_.p, ev.w0, bytes, hipMemcpyDeviceToDevice, ev.stream),
               "state backup");
        forecast(eng, dt, steps, precision);
        const double sp = spread(eng);
        if (spread_out)
            *spread_out = sp;
        if (error_rms_ <= threshold * sp)
            return 0;
        hcheck(hipMemcpyAsync(ev.w0, backup_.p, bytes, hipMemcpyDeviceToDevice, ev.stream),
               "state restore");
        eng.rewind_counter(steps);
        eng.invalidate_mixed();
        eng.step(dt, steps);
        return 1;
    }

    // Advance every member of an FP32 ensemble by steps * dt with the network.
    void forecast(EngineBase &eng, double dt, size_t steps, int precision) {
        const EngineViews ev = check_engine(eng);
        check_precision(precision);
        if (steps == 0 || !(dt > 0) || !std::isfinite(dt))
            throw std::invalid_argument("forecast needs dt > 0 and steps > 0");
        if (horizon_ && (steps != horizon_ || std::fabs(dt - dt_) > 1e-12 * dt_))
            throw std::invalid_argument("forecast horizon differs from the trained horizon");
        // run on the engine's stream, in order with its own work; the noise
        // counter is the ensemble's step index, so a forecast is a function of
        // the ensemble state alone
        st_ = ev.stream;
        // the forecast's scratch buffers (activations, counter) are the
        // surrogate's: a forecast on another ensemble's stream waits on the
        // device for the previous one (outside any capture)
        if (!order_ev_)
            hcheck(hipEventCreateWithFlags(&order_ev_, hipEventDisableTiming), "event");
        else if (order_stream_ != st_)
            hcheck(hipStreamWaitEvent(st_, order_ev_, 0), "forecast order");
        counter_dev_.alloc(1);
        hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, st_, counter_dev_.p,
                           eng.step_index());
        if (precision == HZL_BF16)
            prepare_weights(wb_);
        else
            prepare_weights(wf_);
        const GraphKey key{ev.w0,    ev.nu,       ev.amp,    ev.alpha,   ev.sigma,  ev.keys,
                           ev.pair_of, ev.stream, ev.B,      ev.kmin,    ev.kmax,
                           precision, in_scale_,  out_scale_, (double)steps * dt,
                           alloc_generation()};
        if (!use_graph_ || ev.B > 1024) {
            launch_forecast(ev, steps, dt, precision);
        } else if (CachedForecast *c = cached(key)) {
            c->used = ++graph_clock_;
            hcheck(hipGraphLaunch(c->exec, st_), "forecast graph launch");
        } else if (seen(key)) {
            // second call with this key: every buffer exists, capture the sequence
            hipGraph_t g = nullptr;
            hcheck(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal), "capture");
            try {
                launch_forecast(ev, steps, dt, precision);
            } catch (...) {
                (void)hipStreamEndCapture(st_, &g);
                if (g)
                    (void)hipGraphDestroy(g);
                throw;
            }
            hcheck(hipStreamEndCapture(st_, &g), "capture end");
            hipGraphExec_t exec = nullptr;
            const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hcheck(e, "forecast graph instantiate");
            keep_graph(key, exec);
            hcheck(hipGraphLaunch(exec, st_), "forecast graph launch");
        } else {
            launch_forecast(ev, steps, dt, precision);
            seen_.push_back(key);
            if (seen_.size() > kSeenKeys)
                seen_.erase(seen_.begin());
        }
        hcheck(hipGetLastError(), "surrogate forecast");
        hcheck(hipEventRecord(order_ev_, st_), "forecast order");
        order_stream_ = st_;
        st_ = own_;
        eng.invalidate_mixed();
        eng.advance_counter(steps);
    }

    void set_graph(bool on) {
        use_graph_ = on;
        reset_graph();
    }

    void set_fused(bool on) {
        fused_ = on;
        reset_graph(); // captured forecasts name the previous path's kernels
    }

  private:
    // The launch sequence of one forecast (no allocation or host sync once
    // the buffers exist, so it can be captured).
    void launch_forecast(const EngineViews &ev, size_t steps, double dt, int precision) {
        const int chunk = std::min(ev.B, 1024);
        for (int s0 = 0; s0 < ev.B; s0 += chunk) {
            const int S = std::min(chunk, ev.B - s0);
            cx<float> *w0 = reinterpret_cast<cx<float> *>(ev.w0) + (long)s0 * s_.spec();
            if (precision == HZL_BF16) {
                forward(wb_, w0, ev.nu + s0, ev.amp + s0, S, false);
                if (fused_ && fused_shape())
                    grid_to_increment_fused(wb_.out.p, S);
                else
                    grid_to_increment(wb_.out.p, S);
            } else {
                forward(wf_, w0, ev.nu + s0, ev.amp + s0, S, false);
                grid_to_increment(wf_.out.p, S);
            }
            const dim3 agrid((unsigned)S, (unsigned)((s_.spec() + 255) / 256));
            hipLaunchKernelGGL(k_apply_increment, agrid, dim3(256), 0, st_, w0, incr_.p, S,
                               s_.K, s_.KS, (float)(1.0 / out_scale_), ev.nu + s0, ev.alpha + s0,
                               ev.sigma + s0, ev.keys + s0, ev.pair_of, ev.kmin, ev.kmax,
                               (double)steps * dt, counter_dev_.p);
        }
    }

    // Everything a captured forecast holds: the ensemble's device pointers
    // and stream, its sizes, the scales and horizon, and (through the
    // allocation generation) the surrogate's own buffer addresses.
    struct GraphKey {
        const void *w0, *nu, *amp, *alpha, *sigma, *keys, *pair_of;
        hipStream_t stream;
        int B, kmin, kmax, precision;
        double in_scale, out_scale, horizon;
        uint64_t generation;
        bool operator==(const GraphKey &o) const {
            return w0 == o.w0 && nu == o.nu && amp == o.amp && alpha == o.alpha &&
                   sigma == o.sigma &&
                   keys == o.keys && pair_of == o.pair_of && stream == o.stream && B == o.B &&
                   kmin == o.kmin && kmax == o.kmax && precision == o.precision &&
                   in_scale == o.in_scale && out_scale == o.out_scale &&
                   horizon == o.horizon && generation == o.generation;
        }
    };

    struct CachedForecast {
        GraphKey key;
        hipGraphExec_t exec;
        uint64_t used;
    };

    // Captured forecasts: up to kForecastGraphs launch sequences (one per
    // ensemble / horizon / precision key, least recently used evicted), so a
    // filter that alternates between its particles and a judge ensemble
    // replays both.  A key is captured on its second call (the first runs
    // eagerly, so every lazy allocation happens outside capture).
    CachedForecast *cached(const GraphKey &key) {
        for (auto &c : graphs_)
            if (c.key == key)
                return &c;
        return nullptr;
    }
    bool seen(const GraphKey &key) const {
        for (const auto &k : seen_)
            if (k == key)
                return true;
        return false;
    }
    void keep_graph(const GraphKey &key, hipGraphExec_t exec) {
        if (graphs_.size() >= kForecastGraphs) {
            size_t lru = 0;
            for (size_t i = 1; i < graphs_.size(); ++i)
                if (graphs_[i].used < graphs_[lru].used)
                    lru = i;
            // the evicted sequence may still run on its ensemble's stream
            (void)hipDeviceSynchronize();
            (void)hipGraphExecDestroy(graphs_[lru].exec);
            graphs_.erase(graphs_.begin() + (long)lru);
        }
        graphs_.push_back(CachedForecast{key, exec, ++graph_clock_});
        for (size_t i = 0; i < seen_.size(); ++i)
            if (seen_[i] == key) {
                seen_.erase(seen_.begin() + (long)i);
                break;
            }
    }

    void reset_graph() {
        if (!graphs_.empty()) {
            (void)hipDeviceSynchronize();
            for (auto &c : graphs_)
                (void)hipGraphExecDestroy(c.exec);
        }
        graphs_.clear();
        seen_.clear();
    }

    static FnoShape make_shape(int state_n, int grid, int width, int layers, int modes) {
        FnoShape s;
        if (state_n != 64 && state_n != 128 && state_n != 256)
            throw std::invalid_argument("surrogate state grid must be 64, 128 or 256");
        if (grid != 32 && grid != 64)
            throw std::invalid_argument("surrogate grid must be 32 or 64");
        if (width < 8 || width > 64 || width % 8 || layers < 1 || layers > 8 || modes < 2 ||
            2 * modes > grid / 2)
            throw std::invalid_argument("surrogate needs width in [8, 64] (multiple of 8), "
                                        "1..8 layers and 2 <= modes <= grid / 4");
        s.N = state_n;
        s.K = (state_n - 1) / 3;
        s.G = grid;
        s.C = width;
        s.L = layers;
        s.KX = modes;
        s.KY = 2 * modes;
        s.KS = std::min(grid / 2 - 1, s.K);
        return s;
    }

    static void check_count(size_t S) {
        if (S == 0 || S > 4096)
            throw std::invalid_argument("sample count must be in [1, 4096]");
    }
    static void check_precision(int precision) {
        if (precision != HZL_FP32 && precision != HZL_BF16)
            throw std::invalid_argument("surrogate precision must be HZL_FP32 or HZL_BF16");
    }

    EngineViews check_engine(EngineBase &eng) {
        const EngineViews ev = eng.views();
        if (ev.precision != 0)
            throw std::invalid_argument("the surrogate works on FP32 ensembles");
        if (ev.n != s_.N)
            throw std::invalid_argument("ensemble grid does not match the surrogate");
        return ev;
    }

    void init_params(uint64_t seed) {
        std::mt19937_64 rng(seed);
        std::uniform_real_distribution<double> u(-1.0, 1.0);
        std::vector<float> h(prm_.total, 0.f);
        auto fill = [&](long off, long n, double a) {
            for (long i = 0; i < n; ++i)
                h[off + i] = (float)(a * u(rng));
        };
        fill(prm_.Lw, (long)s_.C * s_.CIN, 1.0 / std::sqrt((double)s_.CIN));
        for (int l = 0; l < s_.L; ++l) {
            fill(prm_.W[l], (long)s_.C * s_.C, 1.0 / std::sqrt((double)s_.C));
            fill(prm_.Rr[l], (long)s_.modes() * s_.C * s_.C, 1.0 / s_.C);
            fill(prm_.Ri[l], (long)s_.modes() * s_.C * s_.C, 1.0 / s_.C);
        }
        fill(prm_.Q1, (long)s_.H * s_.C, 1.0 / std::sqrt((double)s_.C));
        fill(prm_.Q2, s_.H, 1.0 / std::sqrt((double)s_.H));
        hcheck(hipMemcpy(master_.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice),
               "parameter init");
        reset_adam();
    }

    void reset_adam() {
        hipLaunchKernelGGL(k_fill_f32, grid1d(prm_.total), dim3(256), 0, own_, m_.p, prm_.total,
                           0.f);
        hipLaunchKernelGGL(k_fill_f32, grid1d(prm_.total), dim3(256), 0, own_, v_.p, prm_.total,
                           0.f);
        hcheck(hipStreamSynchronize(own_), "optimizer reset");
        t_ = 0;
        ++version_;
    }

    // State <-> grid matrices (FP32):
    //   g1 [2y+ro][2jj+ri]  inverse y of the band (ky = jj - KS): [[c, -s], [s, c]]
    //   g2 [x][ri*KXS+kx]   real inverse x: w (cos, -sin), w = 1 (kx = 0) else 2
    //   f1 [2kx+ri][x]      forward x: (cos, -sin)
    //   f2 [2jj+ro][ri*G+y] forward y / G^2: [[c, s], [-s, c]] / G^2
    void conv_build() {
        const int G = s_.G, KS = s_.KS, JS = 2 * KS + 1, KXS = FnoShape::KXS;
        std::vector<float> g1((size_t)2 * G * 2 * JS), g2((size_t)G * 2 * KXS, 0.f),
            f1((size_t)2 * KXS * G, 0.f), f2((size_t)2 * JS * 2 * G);
        const double z = 1.0 / ((double)G * G);
        for (int y = 0; y < G; ++y)
            for (int jj = 0; jj < JS; ++jj) {
                const double th = 2 * PI * (jj - KS) * y / G, c = std::cos(th), sn = std::sin(th);
                g1[(size_t)(2 * y) * 2 * JS + 2 * jj] = (float)c;
                g1[(size_t)(2 * y) * 2 * JS + 2 * jj + 1] = (float)-sn;
                g1[(size_t)(2 * y + 1) * 2 * JS + 2 * jj] = (float)sn;
                g1[(size_t)(2 * y + 1) * 2 * JS + 2 * jj + 1] = (float)c;
                f2[(size_t)(2 * jj) * 2 * G + y] = (float)(z * c);
                f2[(size_t)(2 * jj) * 2 * G + G + y] = (float)(z * sn);
                f2[(size_t)(2 * jj + 1) * 2 * G + y] = (float)(-z * sn);
                f2[(size_t)(2 * jj + 1) * 2 * G + G + y] = (float)(z * c);
            }
        for (int kx = 0; kx <= KS; ++kx)
            for (int x = 0; x < G; ++x) {
                const double ph = 2 * PI * kx * x / G, w = kx == 0 ? 1.0 : 2.0;
                g2[(size_t)x * 2 * KXS + kx] = (float)(w * std::cos(ph));
                g2[(size_t)x * 2 * KXS + KXS + kx] = (float)(-w * std::sin(ph));
                f1[(size_t)(2 * kx) * G + x] = (float)std::cos(ph);
                f1[(size_t)(2 * kx + 1) * G + x] = (float)-std::sin(ph);
            }
        auto up = [&](Buf<float> &b, const std::vector<float> &v) {
            b.alloc(v.size());
            hcheck(hipMemcpy(b.p, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice),
                   "conversion matrices");
        };
        up(g1_, g1);
        up(g2_, g2);
        up(f1_, f1);
        up(f2_, f2);
    }

    // Grid values of S states (FP32 half planes): out[s*cb1 + y*cm + x*cn] =
    // alpha * w_s(x, y) over the band |kx|, |ky| <= KS.
    void state_to_grid(const cx<float> *states, int S, Span<void> out, long cm, long cn,
                       long cb1, int out_type, float alpha) {
        const int G = s_.G, KS = s_.KS, JS = 2 * KS + 1, KXS = FnoShape::KXS, J = 2 * s_.K + 1;
        const size_t nbm = (size_t)S * 2 * G * KXS;
        bm_.alloc(nbm);
        hipLaunchKernelGGL(k_fill_f32, grid1d((long)nbm), dim3(256), 0, st_, bm_.p, (long)nbm,
                           0.f);
        const float *sf = reinterpret_cast<const float *>(states);
        const size_t skip = 2 * (size_t)(s_.K - KS);
        const Span<const float> st{sf + skip, (size_t)S * 2 * s_.spec() - skip};
        // (i) inverse y of columns kx <= KS: Bm[s][2y+ro][kx]
        Gemm<float> a(2 * G, KS + 1, 2 * JS, cs(g1_), 2 * JS, 1, st, 2L * J, 1, vs(bm_), KXS, 1,
                      OUT_F32);
        a.batch(S, 0, 2 * s_.spec(), 2L * G * KXS);
        a.run(st_);
        // (ii) real inverse x: out[y][x] = sum_k Bm[s][y][k] g2[x][k]
        Gemm<float> b(G, G, 2 * KXS, cs(bm_), 2 * KXS, 1, cs(g2_), 2 * KXS, 1, out, cm, cn,
                      out_type);
        b.batch(S, 2L * G * KXS, 0, cb1);
        b.a.alpha = alpha;
        b.run(st_);
    }

    // state_to_grid / grid_to_increment of the fused inference path: one
    // workgroup per sample (fno_fused.hpp), same matrices and output layouts.
    void state_to_grid_fused(const cx<float> *states, int S, __bf16 *out, long cm, long cn,
                             long cb1, float alpha) {
        StateGridArgs a{reinterpret_cast<const float *>(states), 2 * s_.spec(), s_.K,
                        2 * s_.K + 1, g1_.p, g2_.p, out, 1, cm, cn, cb1, alpha, (long)S};
        hipLaunchKernelGGL(k_state_to_grid, dim3((unsigned)S), dim3(SG_THREADS), 0, st_, a);
    }
    void grid_to_increment_fused(const float *grid, int S) {
        const int KS = s_.KS, JS = 2 * KS + 1;
        incr_.alloc((size_t)S * (KS + 1) * 2 * JS);
        GridIncrementArgs a{grid, f1_.p, f2_.p, incr_.p, (long)S};
        hipLaunchKernelGGL(k_grid_to_increment, dim3((unsigned)S), dim3(SG_THREADS), 0, st_, a);
    }

    // Grid increment (S, G, G) -> band modes into incr_ [s][kx][ky+KS][re, im].
    void grid_to_increment(const float *grid, int S) {
        const int G = s_.G, KS = s_.KS, JS = 2 * KS + 1, KXS = FnoShape::KXS;
        dgrid_.alloc((size_t)S * 2 * KXS * G);
        const Span<const float> gs{grid, (size_t)S * s_.P()};
        Gemm<float> a(2 * (KS + 1), G, G, cs(f1_), G, 1, gs, G, 1, vs(dgrid_), G, 1, OUT_F32);
        a.batch(S, 0, s_.P(), 2L * KXS * G);
        a.run(st_);
        incr_.alloc((size_t)S * (KS + 1) * 2 * JS);
        Gemm<float> b(KS + 1, 2 * JS, 2 * G, cs(dgrid_), 2 * G, 1, cs(f2_), 2 * G, 1, vs(incr_),
                      2 * JS, 1, OUT_F32);
        b.batch(S, 2L * KXS * G, 0, (long)(KS + 1) * 2 * JS);
        b.run(st_);
    }

    // Cast the FP32 master weights into w's operand type (after a change).
    template <typename T> void prepare_weights(FnoWork<T> &w) {
        if (!w.built)
            w.ensure(s_, 1, prm_.total, false); // allocates w_T and rcat
        if (w.version == version_)
            return;
        const size_t rcat = (size_t)s_.L * 4 * s_.C * s_.C * s_.modes();
        if (!w.w_T.p || w.w_T.n < (size_t)prm_.total || !w.rcat.p || w.rcat.n < rcat)
            throw std::logic_error("surrogate weight buffers missing");
        hipLaunchKernelGGL(k_cast<T>, grid1d(prm_.total), dim3(256), 0, st_, master_.p, w.w_T.p,
                           prm_.total);
        const long per = 4L * s_.C * s_.C * s_.modes();
        for (int l = 0; l < s_.L; ++l)
            hipLaunchKernelGGL(k_rcat<T>, grid1d(per), dim3(256), 0, st_, master_.p + prm_.Rr[l],
                               master_.p + prm_.Ri[l], w.rcat.p + l * per, s_.modes(), s_.C);
        w.version = version_;
    }

    // Network output for S FP32 half-plane states into w.out (S x G x G).
    // With `train` the pre-activations and hidden layer are kept for backward.
    template <typename T>
    void forward(FnoWork<T> &w, const cx<float> *states, const double *nu, const double *amp,
                 int S, bool train) {
        w.ensure(s_, S, prm_.total, train);
        prepare_weights(w);
        const FnoShape &s = s_;
        const int G = s.G, C = s.C, KX = s.KX, KY = s.KY, TY = type_of<T>();
        const long P = s.P(), SP = (long)S * P, SA = s.SA(), SV = s.SV();
        const long rper = 4L * C * C * s.modes();
        if constexpr (std::is_same_v<T, __bf16>) {
            if (!train && fused_ && fused_shape()) {
                if (w.feat.n < (size_t)SP * s.CIN)
                    throw std::logic_error("fused surrogate features out of bounds");
                // features 1 and 2 are computed where layer 0 lifts its input
                state_to_grid_fused(states, S, w.feat.p, (long)G * s.CIN, s.CIN, P * s.CIN,
                                    (float)in_scale_);
                layers_fused(w, S, nu, amp);
                return;
            }
        }
        state_to_grid(states, S, vs(w.feat), (long)G * s.CIN, s.CIN, P * s.CIN, TY,
                      (float)in_scale_);
        hipLaunchKernelGGL(k_feature<T>, grid1d(SP), dim3(256), 0, st_, w.feat.p, nu, amp, S, G,
                           s.CIN, s.kf);
        {
            Gemm<T> a((int)SP, C, s.CIN, cs(w.feat), s.CIN, 1, cs(w.w_T, prm_.Lw), s.CIN, 1,
                      vs(*w.v[0]), C, 1, TY);
            a.a.bias = master_.p + prm_.Lb;
            a.run(st_);
        }
        for (int l = 0; l < s.L; ++l) {
            Buf<T> &vl = *w.v[l];
            // x-DFT: A[s][kx'][y][c] = sum_x Fx[kx'][x] v[s][y][x][c]
            Gemm<T> x(2 * KX, C, G, cs(w.cst.Fx), G, 1, cs(vl), 1, C, vs(*w.A[l]), (long)G * C, 1,
                      TY);
            x.batch(S, 0, P * C, SA, G, 0, (long)G * C, C);
            x.run(st_);
            // y-DFT: V[s][kx][k][c] = sum_m Fy2[k][m] A[s][kx][m][c]
            Gemm<T> y(2 * KY, C, 2 * G, cs(w.cst.Fy2), 2 * G, 1, cs(*w.A[l]), 1, C, vs(*w.V[l]), C,
                      1, TY);
            y.batch(S, 0, SA, SV, KX, 0, 2L * G * C, (long)KY * 2 * C);
            y.run(st_);
            // mixing per mode: U[s][mode] = Rcat[mode] V[s][mode]
            Gemm<T> r(S, 2 * C, 2 * C, cs(*w.V[l]), SV, 1, cs(w.rcat, l * rper), 2 * C, 1,
                      vs(*w.U[l]), SV, 1, TY);
            r.batch(s.modes(), 2 * C, 4L * C * C, 2 * C);
            r.run(st_);
            // inverse y: Bi[s][kx][m][c] = sum_k Gy2[m][k] U[s][kx][k][c]
            Gemm<T> iy(2 * G, C, 2 * KY, cs(w.cst.Gy2), 2 * KY, 1, cs(*w.U[l]), 1, C, vs(*w.Bi[l]),
                       C, 1, TY);
            iy.batch(S, 0, SV, SA, KX, 0, (long)KY * 2 * C, 2L * G * C);
            iy.run(st_);
            // inverse x plus the pointwise term v W^T + b, activation:
            //   v'[s][y][x][c] = act(sum_k Gx[x][k] Bi[s][k][y][c] + sum_c' v[s][y][x][c'] W[c][c'] + b[c])
            Gemm<T> ix(G, C, 2 * KX, cs(w.cst.Gx), 2 * KX, 1, cs(*w.Bi[l]), 1, (long)G * C,
                       vs(*w.v[l + 1]), C, 1, TY);
            ix.batch(S, 0, SA, P * C, G, 0, C, (long)G * C);
            ix.second(C, cs(vl), C, 1, P * C, (long)G * C, cs(w.w_T, prm_.W[l]), C, 1, 0, 0);
            ix.a.bias = master_.p + prm_.b[l];
            if (l + 1 < s.L) {
                if (train)
                    ix.a.pre = w.pre[l]->p;
                ix.a.act = ACT_GELU;
            }
            ix.run(st_);
        }
        if (train) {
            Gemm<T> q((int)SP, s.H, C, cs(*w.v[s.L]), C, 1, cs(w.w_T, prm_.Q1), C, 1, vs(w.h), s.H,
                      1, TY);
            q.a.bias = master_.p + prm_.q1;
            q.a.pre = w.hpre.p;
            q.a.act = ACT_GELU;
            q.run(st_);
            hipLaunchKernelGGL(k_head<T>, grid1d(SP), dim3(256), 0, st_, w.h.p,
                               master_.p + prm_.Q2, master_.p + prm_.q2, w.out.p, SP, s.H);
        } else {
            head_inference(w, S);
        }
    }

    // Projection and head of an inference pass (k_proj_head) from the last
    // layer's activations into w.out.
    template <typename T> void head_inference(FnoWork<T> &w, int S) {
        const FnoShape &s = s_;
        const long SP = (long)S * s.P();
        const int C = s.C;
        if (w.v[s.L]->n < (size_t)SP * C || w.w_T.n < (size_t)(prm_.Q1 + s.H * C) ||
            w.out.n < (size_t)SP || s.H % 32 != 0 || C > 64)
            throw std::logic_error("projection operands out of bounds");
        HeadArgs<T> ha{w.v[s.L]->p,         w.w_T.p + prm_.Q1, master_.p + prm_.q1,
                       master_.p + prm_.Q2, master_.p + prm_.q2, w.out.p,
                       SP,                  C,                 s.H};
        const long waves = (SP + 31) / 32;
        hipLaunchKernelGGL(k_proj_head<T>, dim3((unsigned)((waves + 3) / 4)), dim3(256),
                           4 * GemmTile<T>::LDS_PER_WAVE * sizeof(T), st_, ha);
    }

    // The network's shape is the one the fused layer passes are compiled for.
    bool fused_shape() const {
        return s_.G == fno::G && s_.C == fno::C && s_.KX == fno::KX && s_.KY == fno::KY &&
               s_.CIN == fno::CIN && s_.H == FNO_H && s_.KS == fno::KS &&
               FnoShape::KXS == fno::KXS;
    }

    // Lift, spectral layers and head of a BF16 inference pass through the
    // fused passes (fno_fused.hpp): per layer one forward-transform pass, the
    // per-mode mixing pass and one inverse-transform pass with the
    // pointwise term, bias and activation; then the projection and head.
    // w.feat holds the grid feature (channel 0), nu and amp the samples'
    // viscosities and forcing amplitudes; the output lands in w.out.
    void layers_fused(FnoWork<__bf16> &w, int S, const double *nu, const double *amp) {
        using namespace fno;
        const FnoShape &s = s_;
        const long SP = (long)S * s.P(), SV = s.SV(), rper = 4L * C * C * s.modes();
        if (w.feat.n < (size_t)SP * CIN || w.v[0]->n < (size_t)SP * C ||
            w.w_T.n < (size_t)prm_.total || w.V[0]->n < (size_t)S * SV ||
            w.U[0]->n < (size_t)S * SV)
            throw std::logic_error("fused surrogate operands out of bounds");
        // layer 0 lifts its input from the features inside both of its passes
        const FnoLift lift{w.feat.p, w.w_T.p + prm_.Lw, master_.p + prm_.Lb, nu, amp, s.kf};
        for (int l = 0; l < s.L; ++l) {
            const __bf16 *vin = l ? w.v[l]->p : nullptr;
            FnoForwardArgs fa{vin, w.cst.Fx.p, w.cst.Fy2.p, w.V[l]->p, (long)S, lift};
            if (l == 0)
                hipLaunchKernelGGL(k_fno_forward<true>, dim3((unsigned)S), dim3(THREADS), 0, st_,
                                   fa);
            else
                hipLaunchKernelGGL(k_fno_forward<false>, dim3((unsigned)S), dim3(THREADS), 0, st_,
                                   fa);
            if (w.rcat.n < (size_t)((l + 1) * rper))
                throw std::logic_error("fused surrogate mixing weights out of bounds");
            FnoMixArgs ma{w.V[l]->p, w.rcat.p + l * rper, w.U[l]->p, (long)S, s.modes()};
            hipLaunchKernelGGL(k_fno_mix, dim3((unsigned)s.modes()), dim3(MIX_THREADS), 0, st_, ma);
            FnoInverseArgs ia{w.U[l]->p,
                              w.cst.Gy2.p,
                              w.cst.Gx.p,
                              vin,
                              w.w_T.p + prm_.W[l],
                              master_.p + prm_.b[l],
                              w.v[l + 1]->p,
                              (long)S,
                              l + 1 < s.L ? ACT_GELU : ACT_NONE,
                              lift};
            if (l == 0)
                hipLaunchKernelGGL(k_fno_inverse<true>, dim3((unsigned)(S * NBLKI)), dim3(THREADS),
                                   0, st_, ia);
            else
                hipLaunchKernelGGL(k_fno_inverse<false>, dim3((unsigned)(S * NBLKI)),
                                   dim3(THREADS), 0, st_, ia);
        }
        if (s.H != FNO_H || w.v[s.L]->n < (size_t)SP * C || w.out.n < (size_t)SP)
            throw std::logic_error("fused surrogate head operands out of bounds");
        FnoHeadArgs ha{w.v[s.L]->p,         w.w_T.p + prm_.Q1, master_.p + prm_.q1,
                       master_.p + prm_.Q2, master_.p + prm_.q2, w.out.p, SP};
        const long waves = (SP + 31) / 32;
        hipLaunchKernelGGL(k_fno_head, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st_, ha);
    }


    // Mean squared error against the scaled true grid increment; with
    // want_grad, every parameter gradient (FP32) into grad_.
    template <typename T>
    double loss_and_grad(FnoWork<T> &w, const cx<float> *x, const cx<float> *y, const double *nu,
                         const double *amp, int S, bool want_grad) {
        forward(w, x, nu, amp, S, want_grad);
        const long n = (long)S * s_.spec(), SP = (long)S * s_.P();
        diff_.alloc((size_t)n);
        hipLaunchKernelGGL(k_diff_states, grid1d(n), dim3(256), 0, st_, y, x, diff_.p, n);
        target_.alloc((size_t)SP);
        state_to_grid(diff_.p, S, vs(target_), s_.G, 1, s_.P(), OUT_F32, (float)out_scale_);
        hipLaunchKernelGGL(k_mse, grid1d(SP), dim3(256), 0, st_, w.out.p, target_.p, w.dout.p,
                           w.loss_part.p, SP, 1.0 / (double)SP);
        if (want_grad)
            backward(w, S);
        const int chunks = 256;
        lsum_.alloc(chunks);
        hipLaunchKernelGGL(k_dsum, grid1d(chunks), dim3(256), 0, st_, w.loss_part.p, SP, chunks,
                           lsum_.p);
        double h[chunks];
        hcheck(hipMemcpyAsync(h, lsum_.p, sizeof(h), hipMemcpyDeviceToHost, st_), "loss download");
        hcheck(hipStreamSynchronize(st_), "loss");
        double sum = 0;
        for (double v : h)
            sum += v;
        return sum / (double)SP;
    }

    // grad[off + m*N + n] = sum over `rows` rows r of X1[r*ld1 + m] X2[r*ld2 + n]
    // (split over row chunks of 1024, then summed).
    template <typename T>
    void weight_grad(Span<const T> X1, int M, long ld1, Span<const T> X2, int N, long ld2,
                     long rows, long off) {
        const long kc = 1024;
        if (rows % kc != 0)
            throw std::logic_error("weight gradient rows must be a multiple of 1024");
        const int chunks = (int)(rows / kc);
        part_.alloc((size_t)chunks * M * N);
        Gemm<T> g(M, N, (int)kc, X1, 1, ld1, X2, 1, ld2, vs(part_), N, 1, OUT_F32);
        g.batch(chunks, kc * ld1, kc * ld2, (long)M * N);
        g.run(st_);
        sum_chunks(chunks, (long)M * N, grad_.p + off);
    }

    // out[e] = sum over the chunks of part_[c][e]
    void sum_chunks(int chunks, long count, float *out) {
        hipLaunchKernelGGL(k_sum_chunks_blocks, dim3((unsigned)((count + 31) / 32)), dim3(256), 0,
                           st_, part_.p, chunks, count, out, 0);
    }

    // out[n] = sum over `rows` rows of (w[m] *) X[m ld + n]: block partial sums
    // over row ranges, then their sum
    template <typename T, bool W>
    void column_sums(const T *X, const float *wgt, long rows, int N, long ld, float *out) {
        if (256 % N != 0) {
            const int chunks = 256;
            part_.alloc((size_t)chunks * N);
            if constexpr (W)
                hipLaunchKernelGGL(k_wcolsum<T>, grid1d((long)chunks * N), dim3(256), 0, st_, X,
                                   wgt, rows, N, ld, chunks, part_.p);
            else
                hipLaunchKernelGGL(k_colsum<T>, grid1d((long)chunks * N), dim3(256), 0, st_, X,
                                   rows, N, ld, chunks, part_.p);
            hipLaunchKernelGGL(k_sum_chunks, grid1d(N), dim3(256), 0, st_, part_.p, chunks,
                               (long)N, out, 0);
            return;
        }
        const long per = 512;
        const int blocks = (int)((rows + per - 1) / per);
        part_.alloc((size_t)blocks * N);
        hipLaunchKernelGGL((k_colsum_blocks<T, W>), dim3((unsigned)blocks), dim3(256), 0, st_, X,
                           wgt, rows, N, ld, per, part_.p);
        sum_chunks(blocks, (long)N, out);
    }

    template <typename T> void bias_grad(const T *X, long rows, int N, long ld, long off) {
        column_sums<T, false>(X, nullptr, rows, N, ld, grad_.p + off);
    }

    template <typename T> void backward(FnoWork<T> &w, int S) {
        const FnoShape &s = s_;
        const int G = s.G, C = s.C, KX = s.KX, KY = s.KY, H = s.H, TY = type_of<T>();
        const long P = s.P(), SP = (long)S * P, SA = s.SA(), SV = s.SV();
        const long rper = 4L * C * C * s.modes();
        hipLaunchKernelGGL(k_fill_f32, grid1d(prm_.total), dim3(256), 0, st_, grad_.p, prm_.total,
                           0.f);
        // head: dq2 = sum dout; dQ2[j] = sum dout h[.][j]; dhpre = dout Q2 gelu'(hpre)
        {
            column_sums<float, false>(w.dout.p, nullptr, SP, 1, 1L, grad_.p + prm_.q2);
            column_sums<T, true>(w.h.p, w.dout.p, SP, H, (long)H, grad_.p + prm_.Q2);
            hipLaunchKernelGGL(k_head_back<T>, grid1d(SP * H), dim3(256), 0, st_, w.dout.p,
                               master_.p + prm_.Q2, w.hpre.p, w.dh.p, SP, H);
        }
        // projection: dv = dh Q1; dQ1 = dh^T v_L; dq1 = colsum dh
        Buf<T> *dv = &w.dv0, *dvn = &w.dv1;
        {
            Gemm<T> a((int)SP, C, H, cs(w.dh), H, 1, cs(w.w_T, prm_.Q1), 1, C, vs(*dv), C, 1, TY);
            a.run(st_);
            weight_grad<T>(cs(w.dh), H, H, cs(*w.v[s.L]), C, C, SP, prm_.Q1);
            bias_grad<T>(w.dh.p, SP, H, H, prm_.q1);
        }
        for (int l = s.L - 1; l >= 0; --l) {
            Buf<T> *dpre = dv;
            if (l + 1 < s.L) {
                hipLaunchKernelGGL(k_gelu_back<T>, grid1d(SP * C), dim3(256), 0, st_, dv->p,
                                   w.pre[l]->p, w.g.p, SP * C);
                dpre = &w.g;
            }
            // inverse x: dBi[s][kx'][y][c] = sum_x Gx[x][kx'] dpre[s][y][x][c]
            Gemm<T> a(2 * KX, C, G, cs(w.cst.GxT), G, 1, cs(*dpre), 1, C, vs(w.dBi), (long)G * C,
                      1, TY);
            a.batch(S, 0, P * C, SA, G, 0, (long)G * C, C);
            a.run(st_);
            // pointwise: dv_l = dpre W; dW = dpre^T v_l; db = colsum dpre
            Gemm<T> b((int)SP, C, C, cs(*dpre), C, 1, cs(w.w_T, prm_.W[l]), 1, C, vs(*dvn), C, 1,
                      TY);
            b.run(st_);
            weight_grad<T>(cs(*dpre), C, C, cs(*w.v[l]), C, C, SP, prm_.W[l]);
            bias_grad<T>(dpre->p, SP, C, C, prm_.b[l]);
            // inverse y: dU[s][kx][k][c] = sum_m Gy2[m][k] dBi[s][kx][m][c]
            Gemm<T> c(2 * KY, C, 2 * G, cs(w.cst.Gy2T), 2 * G, 1, cs(w.dBi), 1, C, vs(w.dU), C, 1,
                      TY);
            c.batch(S, 0, SA, SV, KX, 0, 2L * G * C, (long)KY * 2 * C);
            c.run(st_);
            // mixing: dV = dU Rcat; dRcat[mode] = sum_s dU[s]^T V[s]
            Gemm<T> d(S, 2 * C, 2 * C, cs(w.dU), SV, 1, cs(w.rcat, l * rper), 1, 2 * C, vs(w.dV),
                      SV, 1, TY);
            d.batch(s.modes(), 2 * C, 4L * C * C, 2 * C);
            d.run(st_);
            rgrad_.alloc((size_t)rper);
            Gemm<T> e(2 * C, 2 * C, S, cs(w.dU), 1, SV, cs(*w.V[l]), 1, SV, vs(rgrad_), 2 * C, 1,
                      OUT_F32);
            e.batch(s.modes(), 2 * C, 2 * C, 4L * C * C);
            e.run(st_);
            hipLaunchKernelGGL(k_rcat_back, grid1d((long)s.modes() * C * C), dim3(256), 0, st_,
                               rgrad_.p, grad_.p + prm_.Rr[l], grad_.p + prm_.Ri[l], s.modes(), C);
            // forward y: dA[s][kx][m][c] = sum_k Fy2[k][m] dV[s][kx][k][c]
            Gemm<T> f(2 * G, C, 2 * KY, cs(w.cst.Fy2T), 2 * KY, 1, cs(w.dV), 1, C, vs(w.dA), C, 1,
                      TY);
            f.batch(S, 0, SV, SA, KX, 0, (long)KY * 2 * C, 2L * G * C);
            f.run(st_);
            // forward x, accumulated onto the pointwise part
            Gemm<T> x(G, C, 2 * KX, cs(w.cst.FxT), 2 * KX, 1, cs(w.dA), 1, (long)G * C, vs(*dvn),
                      C, 1, TY);
            x.batch(S, 0, SA, P * C, G, 0, C, (long)G * C);
            x.a.beta = 1.f;
            x.run(st_);
            std::swap(dv, dvn);
        }
        // lift: dLw = dv^T feat; dLb = colsum dv
        weight_grad<T>(cs(*dv), C, C, cs(w.feat), s.CIN, s.CIN, SP, prm_.Lw);
        bias_grad<T>(dv->p, SP, C, C, prm_.Lb);
        hcheck(hipGetLastError(), "surrogate backward");
    }

    void adam(float lr) {
        ++t_;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1 - std::pow(b1, (float)t_), c2 = 1 - std::pow(b2, (float)t_);
        hipLaunchKernelGGL(k_adam, grid1d(prm_.total), dim3(256), 0, st_, master_.p, grad_.p, m_.p,
                           v_.p, prm_.total, lr, b1, b2, 1e-8f, c1, c2);
        ++version_;
    }

    // ---- data movement ----------------------------------------------------
    void upload_fields(size_t S, const double *fields, Buf<cx<float>> &dst) {
        const size_t nn = (size_t)s_.N * s_.N, spec = (size_t)s_.spec();
        for (size_t i = 0; i < 2 * nn * S; ++i)
            if (!std::isfinite(fields[i]))
                throw std::invalid_argument("nonfinite surrogate input");
        std::vector<cx<float>> half(spec * S);
        for (size_t k = 0; k < S; ++k)
            full_to_half<float>(s_.N, fields + 2 * nn * k, half.data() + spec * k);
        dst.alloc(half.size());
        hcheck(hipMemcpy(dst.p, half.data(), half.size() * sizeof(cx<float>),
                         hipMemcpyHostToDevice),
               "input upload");
    }
    // per-sample viscosity and Kolmogorov amplitude (amp null: the reference
    // amplitude FNO_AMP_REF for every sample)
    void upload_conditions(size_t S, const double *nu, const double *amp) {
        for (size_t i = 0; i < S; ++i)
            if (!(nu[i] > 0) || !std::isfinite(nu[i]) || (amp && !std::isfinite(amp[i])))
                throw std::invalid_argument(
                    "viscosities must be positive and amplitudes finite");
        hnu_.alloc(S);
        hamp_.alloc(S);
        std::vector<double> a(S, FNO_AMP_REF);
        if (amp)
            std::copy(amp, amp + S, a.begin());
        hcheck(hipMemcpy(hnu_.p, nu, S * sizeof(double), hipMemcpyHostToDevice), "nu upload");
        hcheck(hipMemcpy(hamp_.p, a.data(), S * sizeof(double), hipMemcpyHostToDevice),
               "amplitude upload");
    }
    template <typename T> void download_out(FnoWork<T> &w, size_t S, double *grid) {
        std::vector<float> h((size_t)S * s_.P());
        hcheck(hipStreamSynchronize(st_), "surrogate forward");
        hcheck(hipMemcpy(h.data(), w.out.p, h.size() * sizeof(float), hipMemcpyDeviceToHost),
               "output download");
        for (size_t i = 0; i < h.size(); ++i)
            grid[i] = h[i];
    }

    void grow_dataset(size_t pairs) {
        if (pairs <= cap_)
            return;
        const size_t cap = std::max(pairs, 2 * cap_), spec = (size_t)s_.spec();
        Buf<cx<float>> x, y;
        Buf<double> nu, amp;
        x.alloc(cap * spec);
        y.alloc(cap * spec);
        nu.alloc(cap);
        amp.alloc(cap);
        if (pairs_) {
            hcheck(hipDeviceSynchronize(), "dataset grow");
            hcheck(hipMemcpy(x.p, ds_x_.p, pairs_ * spec * sizeof(cx<float>),
                             hipMemcpyDeviceToDevice),
                   "dataset grow");
            hcheck(hipMemcpy(y.p, ds_y_.p, pairs_ * spec * sizeof(cx<float>),
                             hipMemcpyDeviceToDevice),
                   "dataset grow");
            hcheck(hipMemcpy(nu.p, ds_nu_.p, pairs_ * sizeof(double), hipMemcpyDeviceToDevice),
                   "dataset grow");
            hcheck(hipMemcpy(amp.p, ds_amp_.p, pairs_ * sizeof(double), hipMemcpyDeviceToDevice),
                   "dataset grow");
        }
        std::swap(x.p, ds_x_.p);
        std::swap(x.n, ds_x_.n);
        std::swap(y.p, ds_y_.p);
        std::swap(y.n, ds_y_.n);
        std::swap(nu.p, ds_nu_.p);
        std::swap(nu.n, ds_nu_.n);
        std::swap(amp.p, ds_amp_.p);
        std::swap(amp.n, ds_amp_.n);
        cap_ = cap;
    }

    // Input / output scales from the first collected window: the inverse grid
    // RMS of w and of its increment over the horizon (Parseval on the band).
    void auto_scales(size_t first, size_t count) {
        const size_t spec = (size_t)s_.spec();
        std::vector<cx<float>> x(count * spec), y(count * spec);
        hcheck(hipMemcpy(x.data(), ds_x_.p + first * spec, x.size() * sizeof(cx<float>),
                         hipMemcpyDeviceToHost),
               "dataset statistics");
        hcheck(hipMemcpy(y.data(), ds_y_.p + first * spec, y.size() * sizeof(cx<float>),
                         hipMemcpyDeviceToHost),
               "dataset statistics");
        const int K = s_.K, J = 2 * K + 1, KS = s_.KS;
        double sw = 0, sd = 0;
        for (size_t m = 0; m < count; ++m)
            for (int c = 0; c <= KS; ++c)
                for (int ky = -KS; ky <= KS; ++ky) {
                    const size_t e = m * spec + (size_t)c * J + ky + K;
                    const double wgt = c > 0 ? 2.0 : 1.0;
                    const double xr = x[e].x, xi = x[e].y, dr = y[e].x - xr, di = y[e].y - xi;
                    sw += wgt * (xr * xr + xi * xi);
                    sd += wgt * (dr * dr + di * di);
                }
        sw /= (double)count;
        sd /= (double)count;
        if (sw > 0 && sd > 0) {
            in_scale_ = 1.0 / std::sqrt(sw);
            out_scale_ = 1.0 / std::sqrt(sd);
            scales_set_ = true;
        }
    }

    void gather_batch(const std::vector<int> &idx) {
        const int S = (int)idx.size();
        const long spec = s_.spec();
        bidx_.alloc(S);
        hcheck(hipMemcpy(bidx_.p, idx.data(), S * sizeof(int), hipMemcpyHostToDevice),
               "batch indices");
        bx_.alloc((size_t)S * spec);
        by_.alloc((size_t)S * spec);
        bnu_.alloc(S);
        bamp_.alloc(S);
        hipLaunchKernelGGL(k_gather_states, grid1d((long)S * spec), dim3(256), 0, st_, ds_x_.p,
                           bidx_.p, spec, bx_.p, S);
        hipLaunchKernelGGL(k_gather_states, grid1d((long)S * spec), dim3(256), 0, st_, ds_y_.p,
                           bidx_.p, spec, by_.p, S);
        hipLaunchKernelGGL(k_gather_nu, grid1d(S), dim3(256), 0, st_, ds_nu_.p, bidx_.p, bnu_.p, S);
        hipLaunchKernelGGL(k_gather_nu, grid1d(S), dim3(256), 0, st_, ds_amp_.p, bidx_.p, bamp_.p,
                           S);
    }

    FnoShape s_;
    FnoParams prm_;
    hipStream_t own_ = nullptr, st_ = nullptr;
    Buf<float> master_, grad_, m_, v_;
    Buf<float> g1_, g2_, f1_, f2_, bm_, dgrid_, incr_, target_, part_, rgrad_;
    Buf<cx<float>> diff_, hx_, hy_, ds_x_, ds_y_, bx_, by_, backup_, judge_x_, judge_y_;
    Buf<double> hnu_, ds_nu_, bnu_, hamp_, ds_amp_, bamp_, lsum_, spread_, spread_part_;
    double error_rms_ = -1;
    Buf<uint64_t> counter_dev_;
    bool use_graph_ = graphs_by_default();
    bool fused_ = true; // BF16 inference through the fused layer passes
    static constexpr size_t kForecastGraphs = 4, kSeenKeys = 8;
    std::vector<CachedForecast> graphs_;
    hipEvent_t order_ev_ = nullptr; // end of the last forecast, on order_stream_
    hipStream_t order_stream_ = nullptr;
    // judge in flight: its events, per-mode sums in mapped host memory (host
    // and device addresses), member and mode counts
    hipEvent_t judge_in_ev_ = nullptr, judge_copied_ev_ = nullptr, judge_done_ev_ = nullptr;
    double *judge_host_ = nullptr, *judge_host_dev_ = nullptr;
    size_t judge_host_n_ = 0, judge_members_ = 0, judge_modes_ = 0;
    bool judge_pending_ = false;
    std::vector<GraphKey> seen_;
    uint64_t graph_clock_ = 0;
    Buf<int> bidx_;
    FnoWork<float> wf_;
    FnoWork<__bf16> wb_;
    size_t pairs_ = 0, cap_ = 0, horizon_ = 0;
    double dt_ = 0;
    long t_ = 0;
    uint64_t version_ = 0;
    double in_scale_ = 1.0, out_scale_ = 1.0;
    bool scales_set_ = false;
};

} // namespace hzl_amd

// ---------------------------------------------------------------- C-ABI --
struct hzl_surrogate {
    std::unique_ptr<hzl_amd::Surrogate> net;
};

using hzl_amd::guarded;

namespace {
hzl_amd::Surrogate &net_of(hzl_surrogate_t h);
} // namespace

namespace hzl_amd {
void surrogate_forecast(hzl_surrogate_t s, EngineBase &eng, double dt, size_t steps,
                        int precision) {
    if (eng.forcing_process() == HZL_FORCING_OU)
        throw std::invalid_argument("surrogate forecasts run with the white forcing process");
    net_of(s).forecast(eng, dt, steps, precision);
}
} // namespace hzl_amd

namespace {
hzl_amd::Surrogate &net_of(hzl_surrogate_t h) {
    if (!h || !h->net)
        throw std::invalid_argument("null surrogate handle");
    return *h->net;
}
void need(const void *p) {
    if (!p)
        throw std::invalid_argument("null argument");
}
} // namespace

extern "C" {

int hzl_surrogate_create(hzl_surrogate_t *out, size_t state_n, int grid, int width, int layers,
                         int modes, uint64_t seed) {
    return guarded([&] {
        need(out);
        *out = nullptr;
        if (state_n > 4096)
            throw std::invalid_argument("surrogate state grid must be 64, 128 or 256");
        auto h = std::make_unique<hzl_surrogate>();
        h->net = std::make_unique<hzl_amd::Surrogate>((int)state_n, grid, width, layers, modes,
                                                      seed);
        *out = h.release();
    });
}

int hzl_surrogate_destroy(hzl_surrogate_t h) {
    return guarded([&] { delete h; });
}

int hzl_surrogate_param_count(hzl_surrogate_t h, size_t *count) {
    return guarded([&] {
        need(count);
        *count = (size_t)net_of(h).param_count();
    });
}

int hzl_surrogate_get_params(hzl_surrogate_t h, float *params) {
    return guarded([&] {
        need(params);
        net_of(h).get_params(params);
    });
}

int hzl_surrogate_set_params(hzl_surrogate_t h, const float *params) {
    return guarded([&] {
        need(params);
        net_of(h).set_params(params);
    });
}

int hzl_surrogate_set_scales(hzl_surrogate_t h, double input_scale, double output_scale) {
    return guarded([&] { net_of(h).set_scales(input_scale, output_scale); });
}

int hzl_surrogate_get_scales(hzl_surrogate_t h, double *scales) {
    return guarded([&] {
        need(scales);
        net_of(h).get_scales(scales);
    });
}

int hzl_surrogate_forward(hzl_surrogate_t h, int precision, size_t count, const double *fields,
                          const double *viscosity, const double *amplitude, double *grid_out) {
    return guarded([&] {
        need(fields);
        need(viscosity);
        need(grid_out);
        net_of(h).forward_host(precision, count, fields, viscosity, amplitude, grid_out);
    });
}

int hzl_surrogate_loss_grad(hzl_surrogate_t h, size_t count, const double *fields,
                            const double *next_fields, const double *viscosity,
                            const double *amplitude, double *loss, float *grad) {
    return guarded([&] {
        need(fields);
        need(next_fields);
        need(viscosity);
        need(loss);
        need(grad);
        net_of(h).loss_grad_host(count, fields, next_fields, viscosity, amplitude, loss, grad);
    });
}

int hzl_surrogate_collect(hzl_surrogate_t h, hzl_ensemble_t ensemble, double dt,
                          size_t horizon_steps, size_t windows, int noise_free_targets) {
    return guarded([&] {
        net_of(h).collect(hzl_amd::engine_of(ensemble), dt, horizon_steps, windows,
                          noise_free_targets != 0);
    });
}

int hzl_surrogate_dataset_size(hzl_surrogate_t h, size_t *pairs) {
    return guarded([&] {
        need(pairs);
        *pairs = net_of(h).dataset_size();
    });
}

int hzl_surrogate_fit(hzl_surrogate_t h, size_t first, size_t count, size_t epochs, size_t batch,
                      double learning_rate, uint64_t seed, double *epoch_loss) {
    return guarded([&] {
        net_of(h).fit(first, count, epochs, batch, learning_rate, seed, epoch_loss);
    });
}

int hzl_surrogate_evaluate(hzl_surrogate_t h, int precision, size_t first, size_t count,
                           double *loss) {
    return guarded([&] {
        need(loss);
        *loss = net_of(h).evaluate(precision, first, count);
    });
}

int hzl_surrogate_error_estimate(hzl_surrogate_t h, double *rms) {
    return guarded([&] {
        need(rms);
        *rms = net_of(h).error_estimate();
    });
}

int hzl_surrogate_set_error_estimate(hzl_surrogate_t h, double rms) {
    return guarded([&] { net_of(h).set_error_estimate(rms); });
}

int hzl_surrogate_spread(hzl_surrogate_t h, hzl_ensemble_t ensemble, double *rms) {
    return guarded([&] {
        need(rms);
        *rms = net_of(h).spread(hzl_amd::engine_of(ensemble));
    });
}

namespace {
// The learned forecast adds the white shell noise's variance; an OU ensemble's
// forcing state would not advance through it.
void require_white(hzl_amd::EngineBase &e) {
    if (e.forcing_process() == HZL_FORCING_OU)
        throw std::invalid_argument("the surrogate forecast runs with the white forcing process");
}
} // namespace

int hzl_surrogate_forecast_or_solve(hzl_surrogate_t h, hzl_ensemble_t ensemble, double dt,
                                    size_t steps, int precision, double threshold,
                                    int *used_solver, double *spread) {
    return guarded([&] {
        require_white(hzl_amd::engine_of(ensemble));
        const int r = net_of(h).forecast_or_solve(hzl_amd::engine_of(ensemble), dt, steps,
                                                  precision, threshold, spread);
        if (used_solver)
            *used_solver = r;
    });
}

int hzl_surrogate_set_fused(hzl_surrogate_t h, int enabled) {
    return guarded([&] { net_of(h).set_fused(enabled != 0); });
}

int hzl_surrogate_set_graph(hzl_surrogate_t h, int enabled) {
    return guarded([&] { net_of(h).set_graph(enabled != 0); });
}

int hzl_surrogate_judge(hzl_surrogate_t h, hzl_ensemble_t ensemble, hzl_ensemble_t judge,
                        double dt, size_t steps, int precision, double *rms) {
    return guarded([&] {
        hzl_amd::EngineBase &e = hzl_amd::engine_of(ensemble);
        hzl_amd::EngineBase &j = hzl_amd::engine_of(judge);
        require_white(e);
        const double v = net_of(h).judge(e, j, dt, steps, precision);
        if (rms)
            *rms = v;
    });
}

int hzl_surrogate_judge_start(hzl_surrogate_t h, hzl_ensemble_t ensemble, hzl_ensemble_t judge,
                              double dt, size_t steps, int precision) {
    return guarded([&] {
        hzl_amd::EngineBase &e = hzl_amd::engine_of(ensemble);
        hzl_amd::EngineBase &j = hzl_amd::engine_of(judge);
        require_white(e);
        net_of(h).judge_start(e, j, dt, steps, precision);
    });
}

int hzl_surrogate_judge_poll(hzl_surrogate_t h, int wait, int *finished, double *rms) {
    return guarded([&] {
        hzl_amd::Surrogate &s = net_of(h);
        const int f = s.judge_finish(wait != 0);
        if (finished)
            *finished = f;
        if (rms)
            *rms = s.error_estimate();
    });
}

int hzl_surrogate_forecast(hzl_surrogate_t h, hzl_ensemble_t ensemble, double dt, size_t steps,
                           int precision) {
    return guarded([&] {
        require_white(hzl_amd::engine_of(ensemble));
        net_of(h).forecast(hzl_amd::engine_of(ensemble), dt, steps, precision);
    });
}

} // extern "C"
