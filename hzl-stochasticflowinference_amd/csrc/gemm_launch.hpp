// gemm_launch.hpp -- host side of the MFMA product (gemm_mfma.hpp): device
// buffers, operand spans with a host bounds check of every launch, and the
// launcher.  Shared by the translation units that run products.
#pragma once
#include "engine.hpp"
#include "gemm_mfma.hpp"

#include <stdexcept>
#include <string>

namespace hzl_amd {

// Four waves per block, each an independent 32 x 32 output tile.
template <typename T> __global__ void __launch_bounds__(256) k_gemm(GemmArgs<T> a) {
    extern __shared__ unsigned char smem[];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    T *lds = reinterpret_cast<T *>(smem) + wave * GemmTile<T>::LDS_PER_WAVE;
    gemm_wave<T>(DevCtx{}, a, (long)blockIdx.x * 4 + wave, lane, lds);
}

namespace gemm_host {

inline void hcheck(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline dim3 grid1d(long count) {
    return dim3((unsigned)((count + 255) / 256));
}

// Number of device allocations made by Buf so far: a captured graph that
// holds buffer addresses is valid while this has not changed.
inline uint64_t &alloc_generation() {
    static uint64_t g = 0;
    return g;
}

// Device buffer that only grows.  Growing waits for the device first: work
// still queued on any stream may read the old allocation.
template <typename T> struct Buf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        if (count <= n)
            return;
        if (p) {
            hcheck(hipDeviceSynchronize(), "buffer growth");
            (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
        hcheck(hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T)), "hipMalloc");
        n = count;
        ++alloc_generation();
    }
    Buf() = default;
    Buf(const Buf &) = delete;
    Buf &operator=(const Buf &) = delete;
    ~Buf() {
        if (p)
            (void)hipFree(p);
    }
};

template <typename T> constexpr int type_of() {
    return sizeof(T) == 2 ? OUT_BF16 : OUT_F32;
}

// Operand with its addressable extent (elements), for the host-side bounds
// check of every product before launch.
template <typename P> struct Span {
    P *p;
    size_t n;
};
template <typename T> Span<const T> cs(Buf<T> &b, size_t off = 0) {
    if (off > b.n)
        throw std::logic_error("operand offset beyond buffer");
    return {b.p + off, b.n - off};
}
template <typename T> Span<void> vs(Buf<T> &b, size_t off = 0) {
    if (off > b.n)
        throw std::logic_error("operand offset beyond buffer");
    return {b.p + off, b.n - off};
}

template <typename T> struct Gemm {
    GemmArgs<T> a{};
    size_t na = 0, nb = 0, nc = 0, nx = 0, na2 = 0, nb2 = 0;
    Gemm(int M, int N, int K, Span<const T> A, long am, long ak, Span<const T> B, long bn,
         long bk, Span<void> C, long cm, long cn, int ctype) {
        a.M = M;
        a.N = N;
        a.K = K;
        a.nb1 = a.nb2 = 1;
        a.A = A.p;
        a.a_m = am;
        a.a_k = ak;
        a.B = B.p;
        a.b_n = bn;
        a.b_k = bk;
        a.C = C.p;
        a.c_m = cm;
        a.c_n = cn;
        a.c_type = ctype;
        a.alpha = 1.f;
        a.beta = 0.f;
        a.aux_type = type_of<T>();
        a.act = ACT_NONE;
        na = A.n;
        nb = B.n;
        nc = C.n;
    }
    Gemm &batch(int nb1, long ab1, long bb1, long cb1, int nb2 = 1, long ab2 = 0, long bb2 = 0,
                long cb2 = 0) {
        a.nb1 = nb1;
        a.a_b1 = ab1;
        a.b_b1 = bb1;
        a.c_b1 = cb1;
        a.nb2 = nb2;
        a.a_b2 = ab2;
        a.b_b2 = bb2;
        a.c_b2 = cb2;
        return *this;
    }
    Gemm &aux(Span<const T> X, long xm, long xn, long xb1, long xb2) {
        a.aux = X.p;
        nx = X.n;
        a.x_m = xm;
        a.x_n = xn;
        a.x_b1 = xb1;
        a.x_b2 = xb2;
        return *this;
    }
    Gemm &second(int K2, Span<const T> A2, long am, long ak, long ab1, long ab2, Span<const T> B2,
                 long bn, long bk, long bb1, long bb2) {
        a.K2 = K2;
        a.A2 = A2.p;
        a.a2_m = am;
        a.a2_k = ak;
        a.a2_b1 = ab1;
        a.a2_b2 = ab2;
        a.B2 = B2.p;
        a.b2_n = bn;
        a.b2_k = bk;
        a.b2_b1 = bb1;
        a.b2_b2 = bb2;
        na2 = A2.n;
        nb2 = B2.n;
        return *this;
    }
    // Largest element index an operand touches (all strides are >= 0).
    static size_t extent(long b1, long s1, long b2, long s2, long r, long sr, long k, long sk) {
        return (size_t)((b1 - 1) * s1 + (b2 - 1) * s2 + (r - 1) * sr + (k - 1) * sk);
    }
    void check() const {
        const long strides[] = {a.a_m,  a.a_k,  a.a_b1, a.a_b2, a.b_n,  a.b_k,  a.b_b1,
                                a.b_b2, a.c_m,  a.c_n,  a.c_b1, a.c_b2, a.x_m,  a.x_n,
                                a.x_b1, a.x_b2, a.a2_m, a.a2_k, a.a2_b1, a.a2_b2, a.b2_n,
                                a.b2_k, a.b2_b1, a.b2_b2};
        for (long s : strides)
            if (s < 0)
                throw std::logic_error("negative GEMM stride");
        if (a.M <= 0 || a.N <= 0 || a.K <= 0 || a.nb1 <= 0 || a.nb2 <= 0)
            throw std::logic_error("empty GEMM");
        if (extent(a.nb1, a.a_b1, a.nb2, a.a_b2, a.M, a.a_m, a.K, a.a_k) >= na ||
            extent(a.nb1, a.b_b1, a.nb2, a.b_b2, a.N, a.b_n, a.K, a.b_k) >= nb ||
            extent(a.nb1, a.c_b1, a.nb2, a.c_b2, a.M, a.c_m, a.N, a.c_n) >= nc ||
            (a.aux && extent(a.nb1, a.x_b1, a.nb2, a.x_b2, a.M, a.x_m, a.N, a.x_n) >= nx))
            throw std::logic_error("GEMM operand out of bounds");
        if (a.A2 && (a.K2 <= 0 || !a.B2 ||
                     extent(a.nb1, a.a2_b1, a.nb2, a.a2_b2, a.M, a.a2_m, a.K2, a.a2_k) >= na2 ||
                     extent(a.nb1, a.b2_b1, a.nb2, a.b2_b2, a.N, a.b2_n, a.K2, a.b2_k) >= nb2))
            throw std::logic_error("second GEMM operand out of bounds");
    }
    void run(hipStream_t s) const {
        check();
        const long tiles = (long)a.nb1 * a.nb2 * ((a.M + 31) / 32) * ((a.N + 31) / 32);
        const size_t lds = 4 * GemmTile<T>::LDS_PER_WAVE * sizeof(T);
        hipLaunchKernelGGL(k_gemm<T>, dim3((unsigned)((tiles + 3) / 4)), dim3(256), lds, s, a);
    }
};

} // namespace gemm_host
using namespace gemm_host;

} // namespace hzl_amd
