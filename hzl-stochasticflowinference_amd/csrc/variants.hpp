// variants.hpp -- launch-shape variants of the FP32 RK4 kernels that the
// autotuner times against the default shape (variant 0, Shape<N> with the
// register budgets of engine.hip).
//
// A column variant is (values per lane E, i.e. the T x E split of the y
// transform -- two passes when E >= T, three when E = 8 < T; lanes per
// workgroup; waves-per-EU hint; non-temporal streams).
// A row variant is (E of the x transforms; row pairs per workgroup;
// waves-per-EU hint; staged or direct spectra; non-temporal streams).  Every
// variant computes the same stage; only the FFT factorisation (and so the
// FP32 rounding) and the scheduling differ.
#pragma once
#include "kernels.hpp"
#include <cstdint>
#include <hip/hip_runtime.h>

// X(N, index, E, lanes per workgroup, waves per EU, NT flags of Geometry: 1 streams the
// mixed and nonlinear fields, 2 reads the twiddles of RK4 stages 3-4 from the table)
#define HZL_COL_VARIANTS(X)                                                                    \
    X(128, 1, 16, 128, 1, 0)                                                                    \
    X(128, 2, 32, 64, 1, 0)                                                                     \
    X(128, 3, 16, 64, 2, 0)                                                                     \
    X(128, 4, 16, 64, 1, 1)                                                                     \
    X(256, 1, 16, 128, 1, 0)                                                                    \
    X(256, 2, 32, 64, 1, 0)                                                                     \
    X(256, 3, 16, 64, 2, 0)                                                                     \
    X(256, 4, 16, 64, 1, 1)                                                                     \
    X(256, 5, 16, 128, 1, 1)                                                                    \
    X(256, 6, 16, 256, 1, 1)                                                                    \
    X(256, 7, 16, 256, 1, 3)                                                                    \
    X(256, 8, 16, 128, 1, 3)                                                                    \
    X(512, 1, 32, 128, 1, 0)                                                                    \
    X(512, 2, 32, 64, 2, 0)                                                                     \
    X(512, 3, 32, 64, 1, 1)                                                                     \
    X(512, 4, 8, 64, 2, 0)                                                                      \
    X(512, 5, 8, 128, 2, 0)                                                                     \
    X(512, 6, 8, 256, 4, 0)                                                                     \
    X(512, 7, 8, 256, 4, 1)                                                                     \
    X(512, 8, 8, 128, 4, 0)                                                                     \
    X(256, 9, 8, 128, 4, 0)                                                                     \
    X(256, 10, 8, 256, 2, 1)                                                                    \
    X(128, 5, 8, 128, 4, 0)                                                                     \
    X(128, 6, 8, 256, 2, 1)                                                                     \
    X(128, 7, 8, 64, 4, 1)

// X(N, index, E, row pairs per workgroup, waves per EU, direct, NT flags of Geometry (1:
// non-temporal loads of the mixed fields, 2: twiddles from the table)): direct = 1 builds
// the row spectra straight from HBM (row_body_direct) instead of staging the tiles
// through LDS.
#define HZL_ROW_VARIANTS(X)                                                                    \
    X(128, 1, 16, 16, 2, 0, 0)                                                                  \
    X(128, 2, 32, 16, 2, 0, 0)                                                                  \
    X(128, 3, 16, 8, 3, 0, 0)                                                                   \
    X(128, 4, 16, 8, 1, 1, 0)                                                                   \
    X(128, 5, 16, 8, 2, 0, 1)                                                                   \
    X(256, 1, 16, 8, 3, 0, 0)                                                                   \
    X(256, 2, 16, 8, 4, 0, 0)                                                                   \
    X(256, 3, 16, 16, 2, 0, 0)                                                                  \
    X(256, 4, 32, 8, 2, 0, 0)                                                                   \
    X(256, 5, 16, 8, 1, 1, 0)                                                                   \
    X(256, 6, 16, 8, 2, 1, 0)                                                                   \
    X(256, 7, 16, 8, 2, 0, 1)                                                                   \
    X(256, 8, 16, 16, 2, 0, 1)                                                                  \
    X(256, 9, 16, 8, 3, 0, 2)                                                                   \
    X(256, 10, 16, 8, 3, 0, 3)                                                                  \
    X(256, 11, 16, 16, 3, 0, 3)                                                                 \
    X(512, 1, 32, 8, 3, 0, 0)                                                                   \
    X(512, 2, 32, 8, 1, 1, 0)                                                                   \
    X(512, 3, 32, 8, 2, 0, 1)                                                                   \
    X(512, 4, 8, 8, 4, 0, 0)                                                                    \
    X(512, 5, 8, 8, 2, 1, 0)                                                                    \
    X(512, 6, 8, 4, 4, 1, 0)                                                                    \
    X(512, 7, 8, 8, 4, 0, 1)                                                                    \
    X(256, 12, 8, 8, 4, 0, 0)                                                                   \
    X(256, 13, 8, 8, 2, 1, 0)                                                                   \
    X(256, 14, 8, 4, 4, 1, 0)                                                                   \
    X(128, 6, 8, 8, 4, 0, 0)                                                                    \
    X(128, 7, 8, 16, 4, 0, 1)                                                                   \
    X(128, 8, 8, 8, 4, 1, 0)

namespace hzl_amd {

// FNV-1a hash of the variant lists above (keys the tuning cache, so a cached
// choice never names a variant of another build's table).
uint64_t variant_table_id();
// Number of column / row variants at grid n for FP32, default included.
int col_variant_count(int n);
int row_variant_count(int n);
// Values per lane of a variant (selects the twiddle table).
int col_variant_elems(int n, int v);
int row_variant_elems(int n, int v);
// Launch variant v >= 1 (FP32); false when (n, v) names none.
bool launch_col_variant(int n, int v, int mode, const StageArgs<float> &a, hipStream_t s);
bool launch_row_variant(int n, int v, const RowArgs<float> &a, hipStream_t s);

} // namespace hzl_amd
