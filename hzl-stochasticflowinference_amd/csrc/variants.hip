// variants.hip -- instantiations of the FP32 launch-shape variants listed in
// variants.hpp (kept apart from engine.hip so the two compile in parallel).
#include "engine.hpp"
#include "variants.hpp"

namespace hzl_amd {

template <int N, int MODE, class G, int WPE>
__global__ void __launch_bounds__(G::COL_THREADS)
    __attribute__((amdgpu_waves_per_eu(WPE))) k_col_v(StageArgs<float> a) {
    __shared__ cx<float> lds[G::COL_G * GroupFFT<N, G::E, float>::LDS_ELEMS];
    col_body<N, MODE, float, G>(DevCtx{}, a, lds, (long)blockIdx.x);
}

template <int N, bool CFL, class G, int WPE>
__global__ void __launch_bounds__(G::ROW_THREADS)
    __attribute__((amdgpu_waves_per_eu(WPE))) k_row_v(RowArgs<float> a) {
    __shared__ cx<float> lds[RowShape<G>::template lds_elems<float>()];
    row_body<N, float, CFL, G>(DevCtx{}, a, lds, (long)blockIdx.x);
}

template <int N, bool CFL, class G, int WPE>
__global__ void __launch_bounds__(G::ROW_THREADS)
    __attribute__((amdgpu_waves_per_eu(WPE))) k_row_direct_v(RowArgs<float> a) {
    __shared__ cx<float> lds[G::ROW_PPW * GroupFFT<N, G::E, float>::LDS_ELEMS];
    row_body_direct<N, float, CFL, G>(DevCtx{}, a, lds, (long)blockIdx.x);
}

namespace {

template <int N, class G, int WPE>
void col_v(int mode, const StageArgs<float> &a, hipStream_t s) {
    const long cols = (long)a.B * G::KP;
    const dim3 grid((unsigned)((cols + G::COL_G - 1) / G::COL_G)), block(G::COL_THREADS);
    switch (mode) {
    case 0:
        hipLaunchKernelGGL((k_col_v<N, 0, G, WPE>), grid, block, 0, s, a);
        return;
    case 1:
        hipLaunchKernelGGL((k_col_v<N, 1, G, WPE>), grid, block, 0, s, a);
        return;
    case 2:
        hipLaunchKernelGGL((k_col_v<N, 2, G, WPE>), grid, block, 0, s, a);
        return;
    case 3:
        hipLaunchKernelGGL((k_col_v<N, 3, G, WPE>), grid, block, 0, s, a);
        return;
    case 4:
        hipLaunchKernelGGL((k_col_v<N, 4, G, WPE>), grid, block, 0, s, a);
        return;
    }
    throw std::invalid_argument("bad stage");
}

template <int N, class G, int WPE, int DIRECT>
void row_v(const RowArgs<float> &a, hipStream_t s) {
    const long pairs = (long)a.B * (N / 2);
    const dim3 grid((unsigned)((pairs + G::ROW_PPW - 1) / G::ROW_PPW)), block(G::ROW_THREADS);
    if constexpr (DIRECT) {
        if (a.cfl)
            hipLaunchKernelGGL((k_row_direct_v<N, true, G, WPE>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((k_row_direct_v<N, false, G, WPE>), grid, block, 0, s, a);
    } else {
        if (a.cfl)
            hipLaunchKernelGGL((k_row_v<N, true, G, WPE>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((k_row_v<N, false, G, WPE>), grid, block, 0, s, a);
    }
}

int default_elems(int n) {
    switch (n) {
    case 128:
        return Shape<128>::E;
    case 256:
        return Shape<256>::E;
    case 512:
        return Shape<512>::E;
    }
    return 0;
}

} // namespace

uint64_t variant_table_id() {
#define HZL_X(NN, IDX, E, A, W, S) #NN "," #IDX "," #E "," #A "," #W "," #S ";"
#define HZL_Y(NN, IDX, E, A, W, D, S) #NN "," #IDX "," #E "," #A "," #W "," #D "," #S ";"
    static const char table[] = "col:" HZL_COL_VARIANTS(HZL_X) "row:" HZL_ROW_VARIANTS(HZL_Y);
#undef HZL_X
#undef HZL_Y
    uint64_t h = 0xcbf29ce484222325ULL;
    for (const char *c = table; *c; ++c)
        h = (h ^ (uint64_t)(unsigned char)*c) * 0x100000001b3ULL;
    return h;
}

int col_variant_count(int n) {
    int c = 1;
#define HZL_X(NN, IDX, E, CT, W, S) c += n == NN;
    HZL_COL_VARIANTS(HZL_X)
#undef HZL_X
    return c;
}

int row_variant_count(int n) {
    int c = 1;
#define HZL_X(NN, IDX, E, PPW, W, D, S) c += n == NN;
    HZL_ROW_VARIANTS(HZL_X)
#undef HZL_X
    return c;
}

int col_variant_elems(int n, int v) {
    if (v == 0)
        return default_elems(n);
#define HZL_X(NN, IDX, E, CT, W, S)                                                            \
    if (n == NN && v == IDX)                                                                   \
        return E;
    HZL_COL_VARIANTS(HZL_X)
#undef HZL_X
    return 0;
}

int row_variant_elems(int n, int v) {
    if (v == 0)
        return default_elems(n);
#define HZL_X(NN, IDX, E, PPW, W, D, S)                                                        \
    if (n == NN && v == IDX)                                                                   \
        return E;
    HZL_ROW_VARIANTS(HZL_X)
#undef HZL_X
    return 0;
}

bool launch_col_variant(int n, int v, int mode, const StageArgs<float> &a, hipStream_t s) {
#define HZL_X(NN, IDX, E, CT, W, S)                                                            \
    if (n == NN && v == IDX) {                                                                 \
        col_v<NN, Geometry<NN, E, CT, 0, S>, W>(mode, a, s);                                   \
        return true;                                                                           \
    }
    HZL_COL_VARIANTS(HZL_X)
#undef HZL_X
    return false;
}

bool launch_row_variant(int n, int v, const RowArgs<float> &a, hipStream_t s) {
#define HZL_X(NN, IDX, E, PPW, W, D, S)                                                        \
    if (n == NN && v == IDX) {                                                                 \
        row_v<NN, Geometry<NN, E, 64, PPW, S>, W, D>(a, s);                                    \
        return true;                                                                           \
    }
    HZL_ROW_VARIANTS(HZL_X)
#undef HZL_X
    return false;
}

} // namespace hzl_amd
