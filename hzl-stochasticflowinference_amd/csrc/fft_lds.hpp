// fft_lds.hpp -- register + LDS FFT building blocks for gfx950 (CDNA4).
//
// A length-N transform is owned by a group of T = N/E threads that each hold E
// complex values in registers.  When T <= E it runs as two passes (the
// "four-step" factorisation N = T * E, Van Loan 1992 sec. 3.3; Stockham-style
// autosort so no bit reversal is ever materialised); when T > E (few values
// per lane, many lanes per line) as three passes, GroupFFT3 below.  Two passes:
//
//   forward  in : v[j]          = x[t + T*j]                (j < E)
//            pass 1: E-point DFT in registers over j, twiddle W_N^{t*k2}
//            LDS exchange (one write, one read, padded stride T+1)
//            pass 2: E/T T-point DFTs in registers
//            out: v[r*T + k1]  = X[(t + T*r) + E*k1]       (r < E/T, k1 < T)
//   inverse  takes the forward output ordering and returns the forward input
//            ordering, so spectral work between a forward and an inverse
//            transform never touches LDS.
//
// Small DFTs (E, T <= 32) are unrolled radix-2/4 with compile-time twiddles;
// the large twiddles W_N^{t*k2} come from a table computed on the host in
// double precision and rounded once to the working precision.
//
// Everything here is __host__ __device__ and synchronises through a context
// object, so the same code also runs in the CPU lane emulator used by the
// tests.  A group's T lanes are contiguous and never straddle a wavefront, so
// its LDS exchange needs only wave-local ordering (`ctx.wsync()`: LDS
// instructions of one wavefront complete in issue order), not a workgroup
// barrier.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

#define HZL_HD __host__ __device__ __forceinline__

namespace hzl_amd {

template <typename R> struct alignas(2 * sizeof(R)) cx {
    R x, y;
};

template <typename R> HZL_HD cx<R> cadd(cx<R> a, cx<R> b) {
    return {a.x + b.x, a.y + b.y};
}
template <typename R> HZL_HD cx<R> csub(cx<R> a, cx<R> b) {
    return {a.x - b.x, a.y - b.y};
}
template <typename R> HZL_HD cx<R> cmul(cx<R> a, cx<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename R> HZL_HD cx<R> cmulc(cx<R> a, cx<R> b) {
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename R> HZL_HD cx<R> cscale(cx<R> a, R s) {
    return {a.x * s, a.y * s};
}
// i * a and -i * a
template <typename R> HZL_HD cx<R> mul_i(cx<R> a) {
    return {-a.y, a.x};
}
template <typename R> HZL_HD cx<R> mul_mi(cx<R> a) {
    return {a.y, -a.x};
}

// cos(2 pi m / 64) for any integer m (quarter-wave table + symmetry).
HZL_HD constexpr double cos64(int m) {
    constexpr double q[17] = {1.0,
                              0.995184726672196886244837,
                              0.9807852804032304491261822,
                              0.9569403357322088649357979,
                              0.9238795325112867561281832,
                              0.8819212643483550297127569,
                              0.8314696123025452370787884,
                              0.7730104533627369608109066,
                              0.7071067811865475244008444,
                              0.6343932841636454982151716,
                              0.5555702330196022247428308,
                              0.4713967368259976485563876,
                              0.38268343236508977172846,
                              0.2902846772544623676361924,
                              0.1950903220161282678482849,
                              0.09801714032956060199419556,
                              0.0};
    m = ((m % 64) + 64) % 64;
    if (m <= 16)
        return q[m];
    if (m <= 32)
        return -q[32 - m];
    if (m <= 48)
        return -q[m - 32];
    return q[64 - m];
}

// a * W_L^k with W_L = exp(-2 pi i / L) (forward) or its conjugate (inverse).
// k is a compile-time constant after unrolling, so the branches fold away.
template <int L, bool INV, typename R> HZL_HD cx<R> twiddle_const(int k, cx<R> a) {
    static_assert(L <= 64 && (64 % L) == 0, "twiddle_const supports L | 64");
    if (k == 0)
        return a;
    if (4 * k == L)
        return INV ? mul_i(a) : mul_mi(a);
    const int m = k * (64 / L);
    const R c = static_cast<R>(cos64(m));
    const R s = static_cast<R>(cos64(m - 16)); // sin(2 pi m / 64)
    if (8 * k == L || 8 * k == 3 * L) {
        // |c| == |s| == 1/sqrt(2): sign flips fold, two multiplies instead of four
        const R h = static_cast<R>(0.7071067811865475244008444);
        const R sc = c > 0 ? R(1) : R(-1), ss = s > 0 ? R(1) : R(-1);
        if (INV)
            return {(sc * a.x - ss * a.y) * h, (sc * a.y + ss * a.x) * h};
        return {(sc * a.x + ss * a.y) * h, (sc * a.y - ss * a.x) * h};
    }
    if (INV)
        return {a.x * c - a.y * s, a.y * c + a.x * s};
    return {a.x * c + a.y * s, a.y * c - a.x * s};
}

// hi = e + a W_L^k, lo = e - a W_L^k.  A general twiddle is factored as
// f (1, g) with f = cos, g = tan (or f = sin, g = cot when |sin| > |cos|), so
// the twiddle and the butterfly take six fused multiply-adds, u = a (1 -/+ i g)
// then e +/- f u, instead of a complex product and four additions.
template <int L, bool INV, typename R>
HZL_HD void butterfly_const(int k, cx<R> e, cx<R> a, cx<R> &hi, cx<R> &lo) {
    const int m = k * (64 / L);
    const bool general = k != 0 && 4 * k != L && 8 * k != L && 8 * k != 3 * L;
    if (!general) {
        const cx<R> t = twiddle_const<L, INV>(k, a);
        hi = cadd(e, t);
        lo = csub(e, t);
        return;
    }
    const double c = cos64(m), s = cos64(m - 16); // W = c -/+ i s
    cx<R> u;
    R f;
    if (c * c >= s * s) {
        // a W = c a (1 -/+ i tan): forward (a.x + t a.y, a.y - t a.x)
        const R t = static_cast<R>(s / c);
        f = static_cast<R>(c);
        u = INV ? cx<R>{a.x - t * a.y, a.y + t * a.x} : cx<R>{a.x + t * a.y, a.y - t * a.x};
    } else {
        // a W = s a (cot -/+ i): forward (cot a.x + a.y, cot a.y - a.x)
        const R t = static_cast<R>(c / s);
        f = static_cast<R>(s);
        u = INV ? cx<R>{t * a.x - a.y, t * a.y + a.x} : cx<R>{t * a.x + a.y, t * a.y - a.x};
    }
    hi = {e.x + f * u.x, e.y + f * u.y};
    lo = {e.x - f * u.x, e.y - f * u.y};
}

// In-register DFT of L values, natural order in and out.
template <int L, bool INV, typename R> HZL_HD void dft(cx<R> *v) {
    if constexpr (L == 1) {
    } else if constexpr (L == 2) {
        const cx<R> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (L == 4) {
        const cx<R> s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
        const cx<R> s1 = cadd(v[1], v[3]), d1 = csub(v[1], v[3]);
        const cx<R> t = INV ? mul_i(d1) : mul_mi(d1);
        v[0] = cadd(s0, s1);
        v[2] = csub(s0, s1);
        v[1] = cadd(d0, t);
        v[3] = csub(d0, t);
    } else {
        cx<R> e[L / 2], o[L / 2];
#pragma unroll
        for (int i = 0; i < L / 2; ++i) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        dft<L / 2, INV>(e);
        dft<L / 2, INV>(o);
#pragma unroll
        for (int k = 0; k < L / 2; ++k)
            butterfly_const<L, INV>(k, e[k], o[k], v[k], v[k + L / 2]);
    }
}

// Two-pass transform of length N by a group of T = N/E threads (T | E).
// `lds` is the group's private region of at least LDS_ELEMS entries.
// `tw[t*E + k2]` = exp(-2 pi i t k2 / N) for t < T, k2 < E.
// ctx.wsync() orders the exchange within the wavefront; every lane of the
// wavefront must call forward/inverse the same number of times (idle groups
// included).
template <int N, int E, typename R> struct GroupFFT2 {
    static constexpr int T = N / E;
    static constexpr int Q = E / T;
    static constexpr int LD = T + 1;
    static constexpr int LDS_ELEMS = E * LD;
    static_assert(T * E == N && T >= 1 && (E % T) == 0, "need N = T*E with T | E");
    static_assert(LDS_ELEMS >= N, "LDS region must hold one natural-order line");
    static_assert(64 % T == 0, "a group must not straddle a wavefront");

    // Spectral index held in register slot idx by thread t (forward output order).
    HZL_HD static constexpr int index(int t, int idx) {
        return t + T * (idx / T) + E * (idx % T);
    }

    // True when slot idx holds, for every thread t < T, an index k with
    // K < k < N - K (outside the retained band |k| <= K).  idx is a
    // compile-time constant after unrolling, so such slots fold to zero.
    HZL_HD static constexpr bool slot_outside(int idx, int K) {
        const int lo = T * (idx / T) + E * (idx % T);
        return lo > K && lo + T - 1 < N - K;
    }

    // Twiddles of lane t held in registers: w[j] = W_N^{t j}, j < E.  With
    // E == T (one column of the T x E split per lane) the same E values serve
    // the forward and the inverse transform, since W^{t j} is symmetric in
    // (t, j); a group that transforms several lines loads them once.
    struct Twiddles {
        cx<R> w[E];
    };
    HZL_HD static Twiddles load_twiddles(const cx<R> *tw, int t) {
        Twiddles r;
#pragma unroll
        for (int j = 0; j < E; ++j)
            r.w[j] = tw[t * E + j];
        return r;
    }

    template <class Ctx>
    HZL_HD static void forward(const Ctx &ctx, cx<R> (&v)[E], cx<R> *lds, int t,
                               const Twiddles &tr) {
        dft<E, false>(v);
#pragma unroll
        for (int k2 = 1; k2 < E; ++k2)
            v[k2] = cmul(v[k2], tr.w[k2]);
        exchange_forward(ctx, v, lds, t);
    }

    template <class Ctx>
    HZL_HD static void inverse(const Ctx &ctx, cx<R> (&v)[E], cx<R> *lds, int t,
                               const Twiddles &tr) {
        static_assert(Q == 1, "register twiddles of the inverse need E == T");
        cx<R> u[T];
#pragma unroll
        for (int k1 = 0; k1 < T; ++k1)
            u[k1] = v[k1];
        dft<T, true>(u);
#pragma unroll
        for (int n1 = 0; n1 < T; ++n1)
            lds[t * LD + n1] = n1 == 0 ? u[n1] : cmulc(u[n1], tr.w[n1]);
        ctx.wsync();
#pragma unroll
        for (int k2 = 0; k2 < E; ++k2)
            v[k2] = lds[k2 * LD + t];
        dft<E, true>(v);
        ctx.wsync();
    }

    template <class Ctx>
    HZL_HD static void forward(const Ctx &ctx, cx<R> (&v)[E], cx<R> *lds, int t,
                               const cx<R> *tw) {
        dft<E, false>(v);
#pragma unroll
        for (int k2 = 1; k2 < E; ++k2)
            v[k2] = cmul(v[k2], tw[t * E + k2]);
        exchange_forward(ctx, v, lds, t);
    }

    // LDS exchange and second pass of the forward transform.
    template <class Ctx>
    HZL_HD static void exchange_forward(const Ctx &ctx, cx<R> (&v)[E], cx<R> *lds, int t) {
#pragma unroll
        for (int k2 = 0; k2 < E; ++k2)
            lds[k2 * LD + t] = v[k2];
        ctx.wsync();
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            const int k2 = t + T * r;
            cx<R> u[T];
#pragma unroll
            for (int n1 = 0; n1 < T; ++n1)
                u[n1] = lds[k2 * LD + n1];
            dft<T, false>(u);
#pragma unroll
            for (int k1 = 0; k1 < T; ++k1)
                v[r * T + k1] = u[k1];
        }
        ctx.wsync();
    }

    template <class Ctx>
    HZL_HD static void inverse(const Ctx &ctx, cx<R> (&v)[E], cx<R> *lds, int t,
                               const cx<R> *tw) {
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            const int k2 = t + T * r;
            cx<R> u[T];
#pragma unroll
            for (int k1 = 0; k1 < T; ++k1)
                u[k1] = v[r * T + k1];
            dft<T, true>(u);
#pragma unroll
            for (int n1 = 0; n1 < T; ++n1)
                lds[k2 * LD + n1] = n1 == 0 ? u[n1] : cmulc(u[n1], tw[n1 * E + k2]);
        }
        ctx.wsync();
#pragma unroll
        for (int k2 = 0; k2 < E; ++k2)
            v[k2] = lds[k2 * LD + t];
        dft<E, true>(v);
        ctx.wsync();
    }
};

// Three-pass transform of length N = E * T by a group of T = E * F lanes
// (F | E) that each hold E values -- the split for lanes that hold few values
// (E = 8: N = 512 over 64 lanes, 256 over 32, 128 over 16).  Natural order in
// and out:
//
//   n = t + T j              (lane t, slot j),  t = t1 + E t2, t1 < E, t2 < F
//   k = k0 + E k1 + T k2     (k0 < E, k1 < F, k2 < E)
//   pass 1, lane t:          E-point DFT over j -> k0, twiddle W_N^{t k0}
//   exchange 1:              lane t takes the pairs (t1, k0) = (p % E, p / E),
//                            p = t + T r for r < E / F, each with its F values
//                            over t2
//   pass 2:                  F-point DFTs over t2 -> k1, twiddle W_T^{t1 k1}
//   exchange 2:              lane L = k0 + E k1 takes its E values over t1
//   pass 3:                  E-point DFT over t1 -> k2
//
// so lane L ends with X[L + T k2] in slot k2.  As input and output orders
// coincide, the inverse is the same three passes with conjugate twiddles.
// Twiddle table (host): tw[t E + k0] = W_N^{t k0} (t < T, k0 < E), then
// tw[N + t1 F + k1] = W_T^{t1 k1} (t1 < E, k1 < F).
template <int N, int E, typename R> struct GroupFFT3 {
    static constexpr int T = N / E;
    static constexpr int F = T / E;
    static constexpr int RP = E / F; // (t1, k0) pairs per lane in pass 2
    static constexpr int Q = 1;      // a lane's twiddles fit in registers
    // exchange pitches: rows of T (pass-1 output, one row per k0) padded so
    // the pass-2 reads of E consecutive lanes per row spread over the banks;
    // rows of E ((k0, k1) groups) padded to an odd pitch
    static constexpr int P1 = T + 8;
    static constexpr int P2 = E + 1;
    static constexpr int LDS_ELEMS = E * P1 > T * P2 ? E * P1 : T * P2;
    static constexpr int TABLE = N + T; // twiddle table entries
    static_assert(T * E == N && F * E == T && F >= 1 && F <= E && (E % F) == 0,
                  "need N = E * E * F with F | E");
    static_assert(LDS_ELEMS >= N, "LDS region must hold one natural-order line");
    static_assert(64 % T == 0, "a group must not straddle a wavefront");

    HZL_HD static constexpr int index(int t, int idx) {
        return t + T * idx;
    }
    HZL_HD static constexpr bool slot_outside(int idx, int K) {
        const int lo = T * idx;
        return lo > K && lo + T - 1 < N - K;
    }

    struct Twiddles {
        cx<R> w1[E]; // W_N^{t k0}
        cx<R> w2[F]; // W_T^{t1 k1}, t1 = t % E
    };
    HZL_HD static Twiddles load_twiddles(const cx<R> *tw, int t) {
        Twiddles r;
#pragma unroll
        for (int k = 0; k < E; ++k)
            r.w1[k] = tw[t * E + k];
#pragma unroll
        for (int k = 0; k < F; ++k)
            r.w2[k] = tw[N + (t % E) * F + k];
        return r;
    }

    template <class Ctx>
    HZL_HD static void forward(const Ctx &ctx, cx<R> (&v)[E], cx<R> *lds, int t,
                               const Twiddles &tr) {
        run<false>(ctx, v, lds, t, tr.w1, tr.w2);
    }
    template <class Ctx>
    HZL_HD static void inverse(const Ctx &ctx, cx<R> (&v)[E], cx<R> *lds, int t,
                               const Twiddles &tr) {
        run<true>(ctx, v, lds, t, tr.w1, tr.w2);
    }
    template <class Ctx>
    HZL_HD static void forward(const Ctx &ctx, cx<R> (&v)[E], cx<R> *lds, int t,
                               const cx<R> *tw) {
        run<false>(ctx, v, lds, t, tw + t * E, tw + N + (t % E) * F);
    }
    template <class Ctx>
    HZL_HD static void inverse(const Ctx &ctx, cx<R> (&v)[E], cx<R> *lds, int t,
                               const cx<R> *tw) {
        run<true>(ctx, v, lds, t, tw + t * E, tw + N + (t % E) * F);
    }

    // w1, w2: the lane's pass-1 and pass-2 twiddles (registers or table)
    template <bool INV, class Ctx, class W1, class W2>
    HZL_HD static void run(const Ctx &ctx, cx<R> (&v)[E], cx<R> *lds, int t, const W1 &w1,
                           const W2 &w2) {
        dft<E, INV>(v);
#pragma unroll
        for (int k0 = 0; k0 < E; ++k0)
            lds[k0 * P1 + t] = k0 == 0 ? v[k0] : (INV ? cmulc(v[k0], w1[k0]) : cmul(v[k0], w1[k0]));
        ctx.wsync();
        const int t1 = t % E, kb = t / E;
        cx<R> u[RP][F];
#pragma unroll
        for (int r = 0; r < RP; ++r)
#pragma unroll
            for (int t2 = 0; t2 < F; ++t2)
                u[r][t2] = lds[(kb + F * r) * P1 + t1 + E * t2];
        ctx.wsync(); // every read of exchange 1 precedes the writes of exchange 2
#pragma unroll
        for (int r = 0; r < RP; ++r) {
            dft<F, INV>(u[r]);
#pragma unroll
            for (int k1 = 0; k1 < F; ++k1) {
                const cx<R> x = k1 == 0 ? u[r][k1]
                                        : (INV ? cmulc(u[r][k1], w2[k1]) : cmul(u[r][k1], w2[k1]));
                lds[((kb + F * r) * F + k1) * P2 + t1] = x;
            }
        }
        ctx.wsync();
        const int k0 = t % E, k1 = t / E;
#pragma unroll
        for (int j = 0; j < E; ++j)
            v[j] = lds[(k0 * F + k1) * P2 + j];
        dft<E, INV>(v);
        ctx.wsync();
    }
};

// The transform of a T x E split: two passes when each lane holds at least
// T values, else three.
template <int N, int E, typename R>
using GroupFFT = std::conditional_t<(E >= N / E), GroupFFT2<N, E, R>, GroupFFT3<N, E, R>>;

} // namespace hzl_amd
