// tables.hpp -- host-side tables and layout conversions shared by the engine
// and its tests.
#pragma once
#include "kernels.hpp"
#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace hzl_amd {

inline int wave_number(long i, long n) {
    return i < n / 2 ? (int)i : (int)(i - n);
}

inline int elems_per_thread(int n) {
    switch (n) {
    case 16:
        return Shape<16>::E;
    case 32:
        return Shape<32>::E;
    case 64:
        return Shape<64>::E;
    case 128:
        return Shape<128>::E;
    case 256:
        return Shape<256>::E;
    case 512:
        return Shape<512>::E;
    }
    throw std::invalid_argument("grid must be a power of two between 16 and 512");
}

// Values per lane of the default split of precision R (DefaultShape).
template <typename R> int default_split(int n) {
    switch (n) {
    case 16:
        return DefaultShape<16, R>::E;
    case 32:
        return DefaultShape<32, R>::E;
    case 64:
        return DefaultShape<64, R>::E;
    case 128:
        return DefaultShape<128, R>::E;
    case 256:
        return DefaultShape<256, R>::E;
    case 512:
        return DefaultShape<512, R>::E;
    }
    throw std::invalid_argument("grid must be a power of two between 16 and 512");
}

// tw[t*E + k2] = exp(-2 pi i t k2 / n), evaluated in double, rounded once,
// for the T x E split with E values per lane (default: the default split of
// precision R).
// A three-pass split (T > E, T = E F) appends its pass-2 twiddles
// tw[n + t1*F + k1] = exp(-2 pi i t1 k1 / T), t1 < E, k1 < F.
template <typename R> std::vector<cx<R>> twiddles(int n, int E = 0) {
    if (E == 0)
        E = default_split<R>(n);
    const int T = n / E;
    const int F = T > E ? T / E : 0;
    std::vector<cx<R>> tw((size_t)n + (size_t)(E * F));
    auto root = [](long p, long m) -> cx<R> {
        const double a = -2.0 * PI * (double)(p % m) / (double)m;
        return {(R)std::cos(a), (R)std::sin(a)};
    };
    for (int t = 0; t < T; ++t)
        for (int k2 = 0; k2 < E; ++k2)
            tw[(size_t)(t * E + k2)] = root((long)t * k2, n);
    for (int t1 = 0; t1 < E && F; ++t1)
        for (int k1 = 0; k1 < F; ++k1)
            tw[(size_t)(n + t1 * F + k1)] = root((long)t1 * k1, T);
    return tw;
}

// Canonical shell pairs kmin^2 <= |k|^2 <= kmax^2 of the half plane (kx > 0,
// or kx == 0 and ky > 0), numbered kx-major then ky; -1 elsewhere.  The pair
// set is the reference's forcing_pairs (src/solver.cpp:67-83).
inline std::vector<int> pair_table(int kmin, int kmax, int *count = nullptr) {
    const int w = 2 * kmax + 1;
    std::vector<int> table((size_t)((kmax + 1) * w), -1);
    int p = 0;
    for (int a = 0; a <= kmax; ++a)
        for (int b = -kmax; b <= kmax; ++b) {
            const int k2 = a * a + b * b;
            if ((a > 0 || b > 0) && k2 >= kmin * kmin && k2 <= kmax * kmax)
                table[(size_t)(a * w + b + kmax)] = p++;
        }
    if (count)
        *count = p;
    return table;
}

// Full reference Field (n*n complex, y*n + x) -> retained half plane [c][j].
template <typename R>
void full_to_half(int n, const double *full, cx<R> *half) {
    const int K = (n - 1) / 3, J = 2 * K + 1;
    for (int c = 0; c <= K; ++c)
        for (int ky = -K; ky <= K; ++ky) {
            const long y = ((ky % n) + n) % n;
            const long i = y * n + c;
            cx<R> v = {(R)full[2 * i], (R)full[2 * i + 1]};
            if (c == 0 && ky == 0)
                v = {0, 0};
            half[c * J + ky + K] = v;
        }
}

// Retained half plane -> full reference Field (conjugate mirror for kx > 0,
// zero outside the mask).
template <typename R>
void half_to_full(int n, const cx<R> *half, double *full) {
    const int K = (n - 1) / 3, J = 2 * K + 1;
    for (long i = 0; i < 2L * n * n; ++i)
        full[i] = 0;
    for (int c = 0; c <= K; ++c)
        for (int ky = -K; ky <= K; ++ky) {
            const cx<R> v = half[c * J + ky + K];
            const long y = ((ky % n) + n) % n;
            const long i = y * n + c;
            full[2 * i] = (double)v.x;
            full[2 * i + 1] = (double)v.y;
            if (c > 0) {
                const long my = ((-ky % n) + n) % n;
                const long j = my * n + (n - c);
                full[2 * j] = (double)v.x;
                full[2 * j + 1] = -(double)v.y;
            }
        }
}

} // namespace hzl_amd
