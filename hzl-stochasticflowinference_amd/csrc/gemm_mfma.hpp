// gemm_mfma.hpp -- batched C = A * B^T on CDNA4 matrix cores (MFMA).
//
// One wavefront computes one 32x32 output tile with
//   v_mfma_f32_32x32x16_bf16  (T = __bf16, 16 k per instruction), or
//   v_mfma_f32_32x32x2_f32    (T = float,   2 k per instruction),
// accumulating in FP32.  Operand layouts were checked on gfx950 against a host
// product: lane l holds row (l % 32) and the k-block (l / 32); accumulator
// register i holds row 8(i/4) + 4(l/32) + (i%4), column l % 32.
//
// A[m][k] and B[n][k] are addressed through arbitrary element strides, so the
// same kernel serves forward and backward products of every surrogate layer
// (transposes are stride swaps).  Each K-chunk of 32 is staged through the
// wave's private LDS tile: operands whose k or m (n) index is contiguous are
// read with 16-byte vector loads by consecutive lanes and written to LDS in
// [row][k] order, which the fragment reads need.  Batches have two levels
// (b1, b2) with independent strides.  An optional second product A2 B2^T
// (its own strides and K2) accumulates into the same tile, which fuses two
// layer terms into one pass.  Epilogue (FP32): alpha * acc + beta * C_old +
// bias[n] + aux, optional pre-activation store, optional GELU.
#pragma once
#include "fft_lds.hpp"
#include <cstdint>

namespace hzl_amd {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

enum GemmOut : int { OUT_F32 = 0, OUT_BF16 = 1 };
enum GemmAct : int { ACT_NONE = 0, ACT_GELU = 1 };

template <typename T> struct GemmArgs {
    int M, N, K, nb1, nb2;
    const T *A;
    long a_m, a_k, a_b1, a_b2;
    const T *B;
    long b_n, b_k, b_b1, b_b2;
    void *C;
    long c_m, c_n, c_b1, c_b2;
    int c_type;
    float alpha, beta;
    const float *bias;  // [N] or null
    const void *aux;    // added before the activation, or null
    long x_m, x_n, x_b1, x_b2;
    int aux_type;
    void *pre;          // pre-activation (C strides and type), or null
    int act;
    // optional second product accumulated into the same tile (A2 null: none)
    const T *A2;
    long a2_m, a2_k, a2_b1, a2_b2;
    const T *B2;
    long b2_n, b2_k, b2_b1, b2_b2;
    int K2;
};

HZL_HD float gelu(float x) {
    return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
// d gelu / dx
HZL_HD float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

template <typename T> HZL_HD float to_f(T v) {
    return (float)v;
}

HZL_HD float load_any(const void *p, long i, int type) {
    return type == OUT_BF16 ? (float)reinterpret_cast<const __bf16 *>(p)[i]
                            : reinterpret_cast<const float *>(p)[i];
}
HZL_HD void store_any(void *p, long i, int type, float v) {
    if (type == OUT_BF16)
        reinterpret_cast<__bf16 *>(p)[i] = (__bf16)v;
    else
        reinterpret_cast<float *>(p)[i] = v;
}

template <typename T> struct GemmTile {
    static constexpr int KC = 32;                         // k per staged chunk
    static constexpr int PAD = sizeof(T) == 2 ? 8 : 2;    // row padding (elements)
    static constexpr int LD = KC + PAD;
    static constexpr int VEC = 16 / sizeof(T);            // elements per 16-byte load
    static constexpr int LDS_PER_WAVE = 2 * 32 * LD;      // A and B tiles
};

// Stage rows r0..r0+31 (bounded by R) and k0..k0+KC-1 (bounded by K) of a
// strided operand into dst[row][k] (row pitch LD).  `lane` in [0, 64).
template <typename T>
HZL_HD void stage_operand(const T *base, long s_row, long s_k, int r0, int R, int k0, int K,
                          T *dst, int lane) {
    using GT = GemmTile<T>;
    constexpr int KC = GT::KC, LD = GT::LD, VEC = GT::VEC;
    const bool kvec = s_k == 1 && (s_row % VEC) == 0 &&
                      (reinterpret_cast<uintptr_t>(base) % 16) == 0;
    const bool rvec = !kvec && s_row == 1 && (s_k % VEC) == 0 &&
                      (reinterpret_cast<uintptr_t>(base) % 16) == 0;
    if (kvec) {
        // 32 rows x KC contiguous k: chunks of VEC along k
        constexpr int CHUNKS = 32 * KC / VEC;
        for (int c = lane; c < CHUNKS; c += 64) {
            const int row = c / (KC / VEC), seg = (c % (KC / VEC)) * VEC;
            const int r = r0 + row, k = k0 + seg;
            T *d = dst + row * LD + seg;
            if (r < R && k + VEC <= K) {
                struct alignas(16) V {
                    T v[VEC];
                };
                const V x = *reinterpret_cast<const V *>(base + (long)r * s_row + k);
#pragma unroll
                for (int j = 0; j < VEC; ++j)
                    d[j] = x.v[j];
            } else {
#pragma unroll
                for (int j = 0; j < VEC; ++j)
                    d[j] = (r < R && k + j < K) ? base[(long)r * s_row + (k + j)] : T(0);
            }
        }
    } else if (rvec) {
        // KC k-slices of 32 contiguous rows: chunks of VEC along the row index
        constexpr int CHUNKS = KC * 32 / VEC;
        for (int c = lane; c < CHUNKS; c += 64) {
            const int kk = c / (32 / VEC), seg = (c % (32 / VEC)) * VEC;
            const int k = k0 + kk, r = r0 + seg;
            if (k < K && r + VEC <= R) {
                struct alignas(16) V {
                    T v[VEC];
                };
                const V x = *reinterpret_cast<const V *>(base + r + (long)k * s_k);
#pragma unroll
                for (int j = 0; j < VEC; ++j)
                    dst[(seg + j) * LD + kk] = x.v[j];
            } else {
#pragma unroll
                for (int j = 0; j < VEC; ++j)
                    dst[(seg + j) * LD + kk] =
                        (k < K && r + j < R) ? base[(long)(r + j) + (long)k * s_k] : T(0);
            }
        }
    } else {
        for (int e = lane; e < 32 * KC; e += 64) {
            const int row = e / KC, kk = e % KC;
            const int r = r0 + row, k = k0 + kk;
            dst[row * LD + kk] = (r < R && k < K) ? base[(long)r * s_row + (long)k * s_k] : T(0);
        }
    }
}

// acc += A[m0..m0+31][0..K) B[n0..n0+31][0..K)^T through the wave's LDS tiles.
template <typename T, class Ctx>
HZL_HD void mfma_accumulate(const Ctx &ctx, const T *A, long a_m, long a_k, int M, const T *B,
                            long b_n, long b_k, int N, int K, int m0, int n0, T *lds, int lane,
                            f32x16_t &acc) {
    using GT = GemmTile<T>;
    constexpr int KC = GT::KC, LD = GT::LD;
    T *sA = lds, *sB = lds + 32 * LD;
    for (int k0 = 0; k0 < K; k0 += KC) {
        stage_operand<T>(A, a_m, a_k, m0, M, k0, K, sA, lane);
        stage_operand<T>(B, b_n, b_k, n0, N, k0, K, sB, lane);
        ctx.wsync();
        if constexpr (sizeof(T) == 2) {
#pragma unroll
            for (int kk = 0; kk < KC; kk += 16) {
                bf16x8_t fa, fb;
                const T *pa = sA + (lane % 32) * LD + kk + (lane / 32) * 8;
                const T *pb = sB + (lane % 32) * LD + kk + (lane / 32) * 8;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    fa[j] = pa[j];
                    fb[j] = pb[j];
                }
#if __HIP_DEVICE_COMPILE__
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc, 0, 0, 0);
#else
                // host emulation of the instruction (same lane layout)
                ctx.mfma_bf16(fa, fb, acc, lane);
#endif
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < KC; kk += 2) {
                const float fa = sA[(lane % 32) * LD + kk + lane / 32];
                const float fb = sB[(lane % 32) * LD + kk + lane / 32];
#if __HIP_DEVICE_COMPILE__
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(fa, fb, acc, 0, 0, 0);
#else
                ctx.mfma_f32(fa, fb, acc, lane);
#endif
            }
        }
        ctx.wsync();
    }
}

// One wavefront: one 32x32 tile.  `wave` is the global wave index; `lds` the
// wave's private LDS area (GemmTile<T>::LDS_PER_WAVE elements).
template <typename T, class Ctx>
HZL_HD void gemm_wave(const Ctx &ctx, const GemmArgs<T> &a, long wave, int lane, T *lds) {
    const int tm_count = (a.M + 31) / 32, tn_count = (a.N + 31) / 32;
    const long tiles = (long)a.nb1 * a.nb2 * tm_count * tn_count;
    if (wave >= tiles)
        return;
    const int tn = (int)(wave % tn_count);
    const int tm = (int)((wave / tn_count) % tm_count);
    const long b = wave / ((long)tn_count * tm_count);
    const long b1 = b / a.nb2, b2 = b % a.nb2;
    const int m0 = tm * 32, n0 = tn * 32;
    f32x16_t acc;
#pragma unroll
    for (int i = 0; i < 16; ++i)
        acc[i] = 0.f;
    mfma_accumulate<T>(ctx, a.A + b1 * a.a_b1 + b2 * a.a_b2, a.a_m, a.a_k, a.M,
                       a.B + b1 * a.b_b1 + b2 * a.b_b2, a.b_n, a.b_k, a.N, a.K, m0, n0, lds, lane,
                       acc);
    if (a.A2)
        mfma_accumulate<T>(ctx, a.A2 + b1 * a.a2_b1 + b2 * a.a2_b2, a.a2_m, a.a2_k, a.M,
                           a.B2 + b1 * a.b2_b1 + b2 * a.b2_b2, a.b2_n, a.b2_k, a.N, a.K2, m0, n0,
                           lds, lane, acc);
    // epilogue
    char *C = reinterpret_cast<char *>(a.C);
    const int csize = a.c_type == OUT_BF16 ? 2 : 4;
    void *Cb = C + (b1 * a.c_b1 + b2 * a.c_b2) * csize;
    void *Pb = a.pre ? reinterpret_cast<char *>(a.pre) + (b1 * a.c_b1 + b2 * a.c_b2) * csize
                     : nullptr;
    const int asize = a.aux_type == OUT_BF16 ? 2 : 4;
    const void *Xb = a.aux ? reinterpret_cast<const char *>(a.aux) +
                                 (b1 * a.x_b1 + b2 * a.x_b2) * asize
                           : nullptr;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int m = m0 + 8 * (i / 4) + 4 * (lane / 32) + (i % 4), n = n0 + lane % 32;
        if (m < a.M && n < a.N) {
            const long ci = (long)m * a.c_m + (long)n * a.c_n;
            float v = a.alpha * acc[i];
            if (a.beta != 0.f)
                v += a.beta * load_any(Cb, ci, a.c_type);
            if (a.bias)
                v += a.bias[n];
            if (Xb)
                v += load_any(Xb, (long)m * a.x_m + (long)n * a.x_n, a.aux_type);
            if (Pb)
                store_any(Pb, ci, a.c_type, v);
            if (a.act == ACT_GELU)
                v = gelu(v);
            store_any(Cb, ci, a.c_type, v);
        }
    }
}

// Projection and head of the surrogate in one pass (inference):
//   out[m] = q2 + sum_j Q2[j] gelu(q1[j] + sum_c X[m][c] Q1[j][c])
// for the 32 rows m0..m0+31 of one wavefront; X rows have C contiguous
// channels, Q1 is [H][C] (H a multiple of 32, C <= 64).  The hidden
// activations stay in registers: each 32-column block of them is reduced
// into per-row partial sums, then summed across the 32 lanes of a half-wave.
template <typename T> struct HeadArgs {
    const T *X;       // [M][C]
    const T *Q1;      // [H][C]
    const float *q1;  // [H]
    const float *Q2;  // [H]
    const float *q2;  // [1]
    float *out;       // [M]
    long M;
    int C, H;
};

template <typename T, class Ctx>
HZL_HD void proj_head_wave(const Ctx &ctx, const HeadArgs<T> &a, long wave, int lane, T *lds) {
    const long m0 = wave * 32;
    if (m0 >= a.M)
        return;
    const int rows = a.M - m0 < 32 ? (int)(a.M - m0) : 32;
    float rs[16];
#pragma unroll
    for (int i = 0; i < 16; ++i)
        rs[i] = 0.f;
    for (int n0 = 0; n0 < a.H; n0 += 32) {
        f32x16_t acc;
#pragma unroll
        for (int i = 0; i < 16; ++i)
            acc[i] = 0.f;
        mfma_accumulate<T>(ctx, a.X + m0 * a.C, a.C, 1, rows, a.Q1, a.C, 1, a.H, a.C, 0, n0, lds,
                           lane, acc);
        const int n = n0 + lane % 32;
        const float bias = a.q1[n], w = a.Q2[n];
#pragma unroll
        for (int i = 0; i < 16; ++i)
            rs[i] += w * gelu(acc[i] + bias);
    }
    // sum over the 32 columns held by lanes with the same lane / 32
#pragma unroll
    for (int i = 0; i < 16; ++i)
        for (int off = 1; off < 32; off *= 2)
            rs[i] += ctx.shfl_xor(rs[i], off, lane);
    if (lane % 32 == 0) {
        const float q2 = a.q2[0];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int r = 8 * (i / 4) + 4 * (lane / 32) + (i % 4);
            if (r < rows)
                a.out[m0 + r] = rs[i] + q2;
        }
    }
}

} // namespace hzl_amd
