// engine.hip -- gfx950 kernels, the batched ensemble engine and its C-ABI.
//
// Behavioural spec: hzl::CudaSolver (src/cuda_solver.cu:86-235) and
// hzl::Solver::step / rhs (src/solver.cpp:186-254), batched over members with
// device-generated forcing.  See kernels.hpp for the data layout and the
// split of one RHS into a column kernel and a row kernel.
#include "hzl_amd.h"
#include "engine.hpp"
#include "tables.hpp"
#include "variants.hpp"
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <sys/stat.h>
#include <tuple>
#include <limits>
#include <random>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace hzl_amd {

// ------------------------------------------------------------ kernels ----
// Register budgets (waves per SIMD) chosen by measurement on MI355X.
#ifndef HZL_COL_WPE
#define HZL_COL_WPE 1
#endif
#ifndef HZL_ROW_WPE
#define HZL_ROW_WPE 2
#endif

template <int N, int MODE, typename R>
__global__ void __launch_bounds__((DefaultGeometry<N, R>::COL_THREADS))
    __attribute__((amdgpu_waves_per_eu(HZL_COL_WPE))) k_col(StageArgs<R> a) {
    using G = DefaultGeometry<N, R>;
    __shared__ cx<R> lds[G::COL_G * GroupFFT<N, G::E, R>::LDS_ELEMS];
    col_body<N, MODE, R, G>(DevCtx{}, a, lds, (long)blockIdx.x);
}

template <typename R> __global__ void __launch_bounds__(256) k_noise(NoiseArgs<R> a) {
    noise_body<R>(a, (long)blockIdx.x * 256 + threadIdx.x);
}

// One 64-lane workgroup per member (grid = members).
// One workgroup per member: its n spectrum bins (spectrum_body).
template <typename R>
__global__ void __launch_bounds__(256) k_spectrum(int K, int n, long spec, const cx<R> *w,
                                                  double *out) {
    spectrum_body<R>(DevCtx{}, K, n, w + (long)blockIdx.x * spec, out + (long)blockIdx.x * n,
                     256);
}

template <typename R> __global__ void __launch_bounds__(OU_LANES) k_ou(OuArgs<R> a) {
    __shared__ uint64_t s[MT_N];
    __shared__ int pos;
    ou_body<R>(DevCtx{}, a, s, &pos, (int)blockIdx.x);
}

template <int N, typename R, bool CFL>
__global__ void __launch_bounds__((DefaultGeometry<N, R>::ROW_THREADS))
    __attribute__((amdgpu_waves_per_eu(HZL_ROW_WPE))) k_row(RowArgs<R> a) {
    using G = DefaultGeometry<N, R>;
    __shared__ cx<R> lds[RowShape<G>::template lds_elems<R>()];
    row_body<N, R, CFL, G>(DevCtx{}, a, lds, (long)blockIdx.x);
}

template <int N, typename R>
__global__ void __launch_bounds__((DefaultGeometry<N, R>::ROW_THREADS))
    k_row_products(ProductArgs<R> a) {
    using G = DefaultGeometry<N, R>;
    __shared__ cx<R> lds[G::ROW_PPW * GroupFFT<N, G::E, R>::LDS_ELEMS];
    row_body_products<N, R, G>(DevCtx{}, a, lds, (long)blockIdx.x);
}

template <typename R> __global__ void __launch_bounds__(256) k_grf(GrfArgs<R> a) {
    __shared__ double red[256];
    grf_body<R>(DevCtx{}, a, red, 256, (long)blockIdx.x);
}

__global__ void __launch_bounds__(256) k_fill_f64(long count, double value, double *y) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count)
        y[i] = value;
}

template <typename R> __global__ void __launch_bounds__(256) k_diag(DiagArgs<R> a) {
    __shared__ double red[3 * 256];
    diag_body<R>(DevCtx{}, a, red, 256, (long)blockIdx.x);
}

template <typename R> __global__ void __launch_bounds__(256) k_pf_loglik(PfArgs<R> a) {
    __shared__ double red[256];
    pf_loglik_body<R>(DevCtx{}, a, red, 256, (long)blockIdx.x);
}

// one workgroup per filter
__global__ void __launch_bounds__(256) k_pf_normalize(PfControl a) {
    __shared__ double red[2 * 256];
    __shared__ double cw[kPfLocalWeights];
    pf_normalize_body(DevCtx{}, a, red, 256, (int)blockIdx.x, cw, kPfLocalWeights);
}

template <typename R>
__global__ void __launch_bounds__(256) k_pf_mean_partial(int B, long spec, int per,
                                                         const cx<R> *w0, const double *wn,
                                                         double *part) {
    pf_mean_partial_body<R>(B, spec, per, w0, wn, part, (long)blockIdx.x * 256 + threadIdx.x,
                            (int)blockIdx.y);
}
__global__ void __launch_bounds__(256) k_pf_mean_finish(long spec, int chunks,
                                                        const double *part, double *mean) {
    pf_mean_finish_body(spec, chunks, part, mean, (long)blockIdx.x * 256 + threadIdx.x);
}

// Resampling copies, a device-side no-op for every filter of P particles that
// kept its particles this group (stats[8 f + 3] == 0): gather the ancestors'
// states into dst, then commit dst back to the live states.  Grid (members,
// ceil(spec / 256)): member m = blockIdx.x, so the filter's resampling flag
// is one value per workgroup.
template <typename R>
__global__ void __launch_bounds__(256) k_pf_gather(int B, long spec, const cx<R> *src, cx<R> *dst,
                                                   const int *anc, const double *stats, int P) {
    const long m = blockIdx.x, e = (long)blockIdx.y * 256 + threadIdx.x;
    if (m >= B || e >= spec || stats[8 * (m / P) + 3] == 0)
        return;
    dst[m * spec + e] = src[(long)anc[m] * spec + e];
}

template <typename R>
__global__ void __launch_bounds__(256) k_pf_commit(int B, long spec, const cx<R> *src, cx<R> *dst,
                                                   const double *stats, int P) {
    const long m = blockIdx.x, e = (long)blockIdx.y * 256 + threadIdx.x;
    if (m >= B || e >= spec || stats[8 * (m / P) + 3] == 0)
        return;
    dst[m * spec + e] = src[m * spec + e];
}

// ----------------------------------------------- adjoint likelihood ----
// y = a x + b z over `count` complex entries (z may be null; y may alias x or z).
template <typename R>
__global__ void __launch_bounds__(256) k_axpby(long count, cx<R> *y, R a, const cx<R> *x, R b,
                                               const cx<R> *z) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= count)
        return;
    const cx<R> xv = x[i];
    cx<R> r = {a * xv.x, a * xv.y};
    if (z) {
        const cx<R> zv = z[i];
        r = {r.x + b * zv.x, r.y + b * zv.y};
    }
    y[i] = r;
}

// y = 0 over `count` complex entries (a fill, so stale NaNs are cleared too).
template <typename R> __global__ void __launch_bounds__(256) k_zero(long count, cx<R> *y) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count)
        y[i] = {R(0), R(0)};
}

// pred[m * nobs + o]: member m's physical vorticity at observation o, from
// the mixed w~ field (one thread per member and observation).
template <typename R> __global__ void __launch_bounds__(256) k_point_values(PfArgs<R> a) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)a.B * a.nobs)
        return;
    const long m = i / a.nobs;
    const int o = (int)(i % a.nobs);
    const cx<R> *w = a.mix + m * 4 * (long)(a.K + 1) * a.N + 2 * (long)(a.K + 1) * a.N;
    a.pred[i] = point_value(a, w, a.ox[o], a.oy[o]);
}

// cot[m] += sum_o weight[m][o] exp(-2 pi i (kx x_o + ky y_o) / n) on the
// retained half plane: the forward transform of point residuals injected at
// the observation sites (src/adjoint.cpp:96-114).
template <typename R>
__global__ void __launch_bounds__(256) k_inject(int B, int n, int K, cx<R> *cot, int nobs,
                                                const int *ox, const int *oy, const double *weight,
                                                const double *cs, const double *sn) {
    const int J = 2 * K + 1;
    const long spec = (long)(K + 1) * J;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)B * spec)
        return;
    const long m = i / spec;
    const int c = (int)((i % spec) / J), ky = (int)(i % J) - K;
    if (c == 0 && ky == 0)
        return;
    double re = 0, im = 0;
    for (int o = 0; o < nobs; ++o) {
        const long ph = (((long)c * ox[o] + (long)ky * oy[o]) % n + n) % n;
        const double wv = weight[m * nobs + o];
        re += wv * cs[ph];
        im -= wv * sn[ph];
    }
    cot[i] = {cot[i].x + (R)re, cot[i].y + (R)im};
}

// Parameter terms of the adjoint gradient (src/adjoint.cpp:115-123): with
// inner = Re(conj(W) S) summed over the full grid (kx > 0 entries twice),
// grad[2m] -= nu_m sum |k|^2 inner, grad[2m+1] -= alpha_m sum inner.
template <typename R>
__global__ void __launch_bounds__(256) k_param_terms(int K, const cx<R> *S, const cx<R> *W,
                                                     const double *nu, const double *alpha,
                                                     double *grad) {
    __shared__ double red[2 * 256];
    const int J = 2 * K + 1;
    const long spec = (long)(K + 1) * J, m = blockIdx.x;
    double a = 0, b = 0;
    for (long e = threadIdx.x; e < spec; e += 256) {
        const int c = (int)(e / J), ky = (int)(e % J) - K;
        const cx<R> s = S[m * spec + e], w = W[m * spec + e];
        const double inner = (c > 0 ? 2.0 : 1.0) * ((double)w.x * s.x + (double)w.y * s.y);
        a += (double)(c * c + ky * ky) * inner;
        b += inner;
    }
    red[threadIdx.x] = a;
    red[256 + threadIdx.x] = b;
    __syncthreads();
    for (int h = 128; h > 0; h /= 2) {
        if ((int)threadIdx.x < h) {
            red[threadIdx.x] += red[threadIdx.x + h];
            red[256 + threadIdx.x] += red[256 + threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        grad[2 * m] -= nu[m] * red[0];
        grad[2 * m + 1] -= alpha[m] * red[256];
    }
}

// out -= c X on the retained half plane with c = nu_m |k|^2 (j = 0) or
// alpha_m (j = 1): the explicit parameter derivatives of the drift with
// respect to log nu and log alpha (src/solver.cpp:333-340).
template <typename R>
__global__ void __launch_bounds__(256) k_param_source(int B, int K, cx<R> *out, const cx<R> *X,
                                                      const double *nu, const double *alpha,
                                                      int j) {
    const int J = 2 * K + 1;
    const long spec = (long)(K + 1) * J;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)B * spec)
        return;
    const long m = i / spec;
    const int c = (int)((i % spec) / J), ky = (int)(i % J) - K;
    const R f = j == 0 ? (R)(nu[m] * (double)(c * c + ky * ky)) : (R)alpha[m];
    const cx<R> x = X[i], o = out[i];
    out[i] = {o.x - f * x.x, o.y - f * x.y};
}

// Device counters read by replayed graphs: [0] step index, [1] filter group,
// [2] last-group flag.
__global__ void k_set_counters(uint64_t *c, uint64_t step, uint64_t group, uint64_t is_last) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c[0] = step;
        c[1] = group;
        c[2] = is_last;
    }
}

// ------------------------------------------------------------ helpers ----
[[noreturn]] void hip_fail(hipError_t e, const char *what) {
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline void check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        hip_fail(e, what);
}

template <typename T> struct DeviceBuffer {
    T *p = nullptr;
    size_t count = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) : count(n) {
        if (n)
            check(hipMalloc(reinterpret_cast<void **>(&p), n * sizeof(T)), "hipMalloc");
    }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : p(o.p), count(o.count) {
        o.p = nullptr;
        o.count = 0;
    }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
        std::swap(p, o.p);
        std::swap(count, o.count);
        return *this;
    }
    ~DeviceBuffer() {
        if (p)
            (void)hipFree(p);
    }
};

template <int N, typename R> void launch_col(int mode, const StageArgs<R> &a, hipStream_t s) {
    using G = DefaultGeometry<N, R>;
    const long cols = (long)a.B * G::KP;
    const dim3 grid((unsigned)((cols + G::COL_G - 1) / G::COL_G)), block(G::COL_THREADS);
    switch (mode) {
    case 0:
        hipLaunchKernelGGL((k_col<N, 0, R>), grid, block, 0, s, a);
        return;
    case 1:
        hipLaunchKernelGGL((k_col<N, 1, R>), grid, block, 0, s, a);
        return;
    case 2:
        hipLaunchKernelGGL((k_col<N, 2, R>), grid, block, 0, s, a);
        return;
    case 3:
        hipLaunchKernelGGL((k_col<N, 3, R>), grid, block, 0, s, a);
        return;
    case 4:
        hipLaunchKernelGGL((k_col<N, 4, R>), grid, block, 0, s, a);
        return;
    case 5:
        hipLaunchKernelGGL((k_col<N, 5, R>), grid, block, 0, s, a);
        return;
    case 6:
        hipLaunchKernelGGL((k_col<N, 6, R>), grid, block, 0, s, a);
        return;
    case 7:
        hipLaunchKernelGGL((k_col<N, 7, R>), grid, block, 0, s, a);
        return;
    case 8:
        hipLaunchKernelGGL((k_col<N, 8, R>), grid, block, 0, s, a);
        return;
    case 9:
        hipLaunchKernelGGL((k_col<N, 9, R>), grid, block, 0, s, a);
        return;
    case 10:
        hipLaunchKernelGGL((k_col<N, 10, R>), grid, block, 0, s, a);
        return;
    }
    throw std::invalid_argument("bad stage");
}

template <int N, typename R> void launch_products(const ProductArgs<R> &a, hipStream_t s) {
    using G = DefaultGeometry<N, R>;
    const long pairs = (long)a.B * (N / 2);
    hipLaunchKernelGGL((k_row_products<N, R>),
                       dim3((unsigned)((pairs + G::ROW_PPW - 1) / G::ROW_PPW)),
                       dim3(G::ROW_THREADS), 0, s, a);
}

template <typename R> void dispatch_products(int n, const ProductArgs<R> &a, hipStream_t s) {
    switch (n) {
    case 16:
        return launch_products<16, R>(a, s);
    case 32:
        return launch_products<32, R>(a, s);
    case 64:
        return launch_products<64, R>(a, s);
    case 128:
        return launch_products<128, R>(a, s);
    case 256:
        return launch_products<256, R>(a, s);
    case 512:
        return launch_products<512, R>(a, s);
    }
    throw std::invalid_argument("unsupported grid");
}

template <int N, typename R> void launch_row(const RowArgs<R> &a, hipStream_t s) {
    using G = DefaultGeometry<N, R>;
    const long pairs = (long)a.B * (N / 2);
    const long blocks = (pairs + G::ROW_PPW - 1) / G::ROW_PPW;
    if (a.cfl)
        hipLaunchKernelGGL((k_row<N, R, true>), dim3((unsigned)blocks), dim3(G::ROW_THREADS), 0,
                           s, a);
    else
        hipLaunchKernelGGL((k_row<N, R, false>), dim3((unsigned)blocks), dim3(G::ROW_THREADS), 0,
                           s, a);
}

template <typename R> void dispatch_col(int n, int mode, const StageArgs<R> &a, hipStream_t s) {
    switch (n) {
    case 16:
        return launch_col<16, R>(mode, a, s);
    case 32:
        return launch_col<32, R>(mode, a, s);
    case 64:
        return launch_col<64, R>(mode, a, s);
    case 128:
        return launch_col<128, R>(mode, a, s);
    case 256:
        return launch_col<256, R>(mode, a, s);
    case 512:
        return launch_col<512, R>(mode, a, s);
    }
    throw std::invalid_argument("unsupported grid");
}

template <typename R> void dispatch_row(int n, const RowArgs<R> &a, hipStream_t s) {
    switch (n) {
    case 16:
        return launch_row<16, R>(a, s);
    case 32:
        return launch_row<32, R>(a, s);
    case 64:
        return launch_row<64, R>(a, s);
    case 128:
        return launch_row<128, R>(a, s);
    case 256:
        return launch_row<256, R>(a, s);
    case 512:
        return launch_row<512, R>(a, s);
    }
    throw std::invalid_argument("unsupported grid");
}

// ------------------------------------------------------ tuning cache ----
// Chosen launch shapes per (device architecture, variant table, grid,
// precision, member-count bucket): remembered for the process and in a text
// file ($HZL_TUNE_CACHE, else $HOME/.cache/hzl_amd/launch_shapes.txt), one
// "key col_variant row_variant" line per entry, the last line for a key wins.
std::mutex &tune_mutex() {
    static std::mutex m;
    return m;
}

// A remembered launch choice: column variant, row variant.
struct TuneChoice {
    int cv = 0, rv = 0;
};

std::map<std::string, TuneChoice> &tune_memo() {
    static std::map<std::string, TuneChoice> m;
    return m;
}

std::string tune_file() {
    if (const char *p = std::getenv("HZL_TUNE_CACHE"))
        return p;
    if (const char *h = std::getenv("HOME"))
        return std::string(h) + "/.cache/hzl_amd/launch_shapes.txt";
    return "";
}

std::string tune_key(int n, int precision, size_t members) {
    int dev = 0;
    hipDeviceProp_t prop;
    check(hipGetDevice(&dev), "hipGetDevice");
    check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    size_t bucket = 1;
    while (bucket < members)
        bucket *= 2;
    std::ostringstream k;
    std::string arch = prop.gcnArchName;
    for (char &c : arch)
        if (c == ' ')
            c = '_';
    k << arch << "/" << std::hex << variant_table_id() << std::dec << "/n" << n << "/"
      << (precision == HZL_FP32 ? "fp32" : "fp64") << "/b" << bucket;
    return k.str();
}

bool tune_lookup(const std::string &key, TuneChoice *choice) {
    std::lock_guard<std::mutex> lock(tune_mutex());
    auto it = tune_memo().find(key);
    if (it == tune_memo().end()) {
        const std::string path = tune_file();
        if (path.empty())
            return false;
        std::ifstream in(path);
        std::string line;
        bool found = false;
        TuneChoice v;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string k;
            TuneChoice t;
            if (ls >> k >> t.cv >> t.rv && k == key && t.cv >= 0 && t.rv >= 0) {
                v = t;
                found = true;
            }
        }
        if (!found)
            return false;
        it = tune_memo().emplace(key, v).first;
    }
    *choice = it->second;
    return true;
}

void tune_store(const std::string &key, const TuneChoice &choice) {
    std::lock_guard<std::mutex> lock(tune_mutex());
    tune_memo()[key] = choice;
    const std::string path = tune_file();
    if (path.empty())
        return;
    // best effort: create the directory chain, ignore a read-only home
    for (size_t i = path.find('/', 1); i != std::string::npos; i = path.find('/', i + 1))
        (void)::mkdir(path.substr(0, i).c_str(), 0755);
    std::ofstream out(path, std::ios::app);
    if (out)
        out << key << " " << choice.cv << " " << choice.rv << "\n";
}

// ------------------------------------------------------------- engine ----
template <typename R> class Engine final : public EngineBase {
  public:
    Engine(int n, size_t members, const hzl_parameters &p, uint64_t seed)
        : n_(n), B_(members), K_((n - 1) / 3), KP_(K_ + 1), J_(2 * K_ + 1), p_(p) {
        spec_ = (size_t)KP_ * J_;
        mixf_ = (size_t)KP_ * n;
        check(hipGetDevice(&device_), "hipGetDevice");
        check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
        w0_ = DeviceBuffer<cx<R>>(B_ * spec_);
        w1_ = DeviceBuffer<cx<R>>(B_ * spec_);
        acc_ = DeviceBuffer<cx<R>>(B_ * spec_);
        mix_ = DeviceBuffer<cx<R>>(B_ * 4 * mixf_);
        nl_ = DeviceBuffer<cx<R>>(B_ * mixf_);
        ext_ = DeviceBuffer<cx<R>>(B_ * spec_);
        cfl_ = DeviceBuffer<double>(B_ * (size_t)(n / 2));
        cfl_diag_ = DeviceBuffer<double>(B_ * (size_t)(n / 2));
        const auto tw = twiddles<R>(n);
        tw_ = DeviceBuffer<cx<R>>(tw.size());
        upload(tw_.p, tw.data(), tw.size());
        tw_col_ = tw_row_ = tw_.p;
        const auto pairs = pair_table(p.forcing_min, p.forcing_max, &npairs_);
        pairs_ = DeviceBuffer<int>(pairs.size());
        upload(pairs_.p, pairs.data(), pairs.size());
        noise_ = DeviceBuffer<cx<R>>(B_ * (size_t)npairs_);
        check(hipMemsetAsync(noise_.p, 0, noise_.count * sizeof(cx<R>), stream_), "clear noise");
        nu_.assign(B_, p.viscosity);
        alpha_.assign(B_, p.friction);
        sigma_.assign(B_, p.forcing_sigma);
        tau_.assign(B_, p.forcing_tau);
        amp_.assign(B_, p.kolmogorov_amplitude);
        keys_.resize(B_);
        for (size_t m = 0; m < B_; ++m)
            keys_[m] = stream_seed(seed, m);
        d_nu_ = DeviceBuffer<double>(B_);
        d_alpha_ = DeviceBuffer<double>(B_);
        d_sigma_ = DeviceBuffer<double>(B_);
        d_tau_ = DeviceBuffer<double>(B_);
        d_amp_ = DeviceBuffer<double>(B_);
        d_keys_ = DeviceBuffer<uint64_t>(B_);
        upload(d_keys_.p, keys_.data(), B_);
        push_parameters();
        check(hipMemsetAsync(w0_.p, 0, B_ * spec_ * sizeof(cx<R>), stream_), "clear state");
        check(hipMemsetAsync(cfl_.p, 0, cfl_.count * sizeof(double), stream_), "clear CFL");
        ctr_ = DeviceBuffer<uint64_t>(3);
        check(hipStreamSynchronize(stream_), "engine setup");
    }
    ~Engine() override {
        enkf.reset(); // module state may wait on stream_
        if (stream_) {
            (void)hipStreamSynchronize(stream_);
            clear_graphs();
            (void)hipStreamDestroy(stream_);
        }
        if (obs_event_)
            (void)hipEventDestroy(obs_event_);
        if (obs_host_)
            (void)hipHostFree(obs_host_);
        if (pf_host_)
            (void)hipHostFree(pf_host_);
    }

    // One observation group into device memory with a single asynchronous
    // copy from pinned staging: [values | sds | x, y as int pairs].
    void stage_observations(size_t nobs, const int *x, const int *y, const double *value,
                            const double *sd) {
        const size_t words = 3 * nobs + 1;
        if (!obs_event_)
            check(hipEventCreateWithFlags(&obs_event_, hipEventDisableTiming), "event");
        else
            check(hipEventSynchronize(obs_event_), "staging reuse"); // last copy done
        if (obs_cap_ < words) {
            check(hipStreamSynchronize(stream_), "staging growth"); // kernels may read it
            clear_graphs(); // captured cycles hold the old device address
            if (obs_host_)
                (void)hipHostFree(obs_host_);
            obs_host_ = nullptr;
            check(hipHostMalloc(reinterpret_cast<void **>(&obs_host_), words * sizeof(double)),
                  "hipHostMalloc");
            obs_dev_ = DeviceBuffer<double>(words);
            obs_cap_ = words;
        }
        std::memcpy(obs_host_, value, nobs * sizeof(double));
        std::memcpy(obs_host_ + nobs, sd, nobs * sizeof(double));
        int *xy = reinterpret_cast<int *>(obs_host_ + 2 * nobs);
        std::memcpy(xy, x, nobs * sizeof(int));
        std::memcpy(xy + nobs, y, nobs * sizeof(int));
        check(hipMemcpyAsync(obs_dev_.p, obs_host_, 3 * nobs * sizeof(double),
                             hipMemcpyHostToDevice, stream_),
              "observation upload");
        check(hipEventRecord(obs_event_, stream_), "event");
        obs_v_ = obs_dev_.p;
        obs_sd_ = obs_dev_.p + nobs;
        obs_x_ = reinterpret_cast<const int *>(obs_dev_.p + 2 * nobs);
        obs_y_ = obs_x_ + nobs;
    }

    void set_member_parameters(size_t m, double nu, double alpha, double sigma,
                               double amp) override {
        member_check(m);
        if (!std::isfinite(nu) || nu < 0 || !std::isfinite(alpha) || alpha < 0 ||
            !std::isfinite(sigma) || sigma < 0 || !std::isfinite(amp))
            throw std::invalid_argument("nu, alpha, sigma must be finite and nonnegative");
        nu_[m] = nu;
        alpha_[m] = alpha;
        sigma_[m] = sigma;
        amp_[m] = amp;
        params_dirty_ = true;
    }

    void get_member_parameters(size_t m, double *nu, double *alpha, double *sigma,
                               double *amp) const override {
        member_check(m);
        *nu = nu_[m];
        *alpha = alpha_[m];
        *sigma = sigma_[m];
        *amp = amp_[m];
    }

    void set_member_tau(size_t m, double tau) override {
        member_check(m);
        if (!std::isfinite(tau) || tau <= 0)
            throw std::invalid_argument("tau must be finite and positive");
        tau_[m] = tau;
        params_dirty_ = true;
    }

    void set_state(size_t m, const double *field) override {
        const size_t first = m == (size_t)-1 ? 0 : m, count = m == (size_t)-1 ? B_ : 1;
        if (m != (size_t)-1)
            member_check(m);
        const size_t nn = (size_t)n_ * n_;
        for (size_t i = 0; i < 2 * nn * count; ++i)
            if (!std::isfinite(field[i]))
                throw std::invalid_argument("nonfinite initial GPU state");
        std::vector<cx<R>> half(spec_ * count);
        for (size_t k = 0; k < count; ++k)
            full_to_half<R>(n_, field + 2 * nn * k, half.data() + spec_ * k);
        upload(w0_.p + first * spec_, half.data(), half.size());
        mix_valid_ = false;
    }

    void get_state(size_t m, double *field) override {
        const size_t first = m == (size_t)-1 ? 0 : m, count = m == (size_t)-1 ? B_ : 1;
        if (m != (size_t)-1)
            member_check(m);
        std::vector<cx<R>> half(spec_ * count);
        check(hipMemcpyAsync(half.data(), w0_.p + first * spec_, half.size() * sizeof(cx<R>),
                             hipMemcpyDeviceToHost, stream_),
              "state download");
        check(hipStreamSynchronize(stream_), "state download");
        for (const auto &v : half)
            if (!std::isfinite((double)v.x) || !std::isfinite((double)v.y))
                throw std::runtime_error("nonfinite GPU result");
        const size_t nn = (size_t)n_ * n_;
        for (size_t k = 0; k < count; ++k)
            half_to_full<R>(n_, half.data() + spec_ * k, field + 2 * nn * k);
    }

    void init_grf() override {
        GrfArgs<R> a{K_, w0_.p, d_keys_.p};
        hipLaunchKernelGGL(k_grf<R>, dim3((unsigned)B_), dim3(256), 0, stream_, a);
        check(hipGetLastError(), "GRF kernel");
        mix_valid_ = false;
    }

    void step(double dt, size_t steps) override {
        validate_dt(dt);
        if (!use_graph_ || steps < 2) {
            for (size_t s = 0; s < steps; ++s)
                advance(dt, nullptr, 0);
            return;
        }
        // Whole chunks of up to kGraphSteps steps replay one captured launch
        // sequence whose noise kernels read the step index from ctr_[0].
        flush_parameters();
        ensure_mixed(dt);
        const size_t chunk = std::min(steps, kGraphSteps);
        size_t done = 0;
        for (; steps - done >= chunk; done += chunk) {
            set_counters(0, 0);
            const GraphKey key{1, dt, chunk, 0, 0, nullptr, false};
            run_captured(key, [&]() {
                for (size_t s = 0; s < chunk; ++s)
                    launch_step(dt, nullptr, 0, ctr_.p, s, nullptr);
            });
            step_ += chunk;
        }
        last_dt_ = dt;
        max_dt_ = std::max(max_dt_, dt);
        for (; done < steps; ++done)
            advance(dt, nullptr, 0);
    }

    void set_graph(bool on) override {
        check(hipStreamSynchronize(stream_), "graph switch");
        use_graph_ = on;
        clear_graphs();
    }


    // ------------------------------------- SSPRK3 and calibration drifts ----
    // Solver::step(dt, TimeIntegrator::ssprk3) (src/solver.cpp:235-246) for
    // every member: device noise held through 3 x (row, col) launches.
    void step_integrator(double dt, size_t steps, int integrator) override {
        if (integrator == HZL_RK4)
            return step(dt, steps);
        if (integrator != HZL_SSPRK3)
            throw std::invalid_argument("integrator must be HZL_RK4 or HZL_SSPRK3");
        validate_dt(dt);
        flush_parameters();
        for (size_t s = 0; s < steps; ++s) {
            ensure_mixed(dt);
            if (ou_) {
                launch_ou(1, dt / 2, noise_.p);
            } else {
                NoiseArgs<R> na{(int)B_, npairs_, d_sigma_.p, d_keys_.p, dt, step_, noise_.p};
                const long count = (long)B_ * npairs_;
                hipLaunchKernelGGL(k_noise<R>, dim3((unsigned)((count + 255) / 256)), dim3(256),
                                   0, stream_, na);
            }
            RowArgs<R> r{(int)B_, mix_.p, nl_.p, tw_row_, nullptr};
            for (int stage = 1; stage <= 3; ++stage) {
                r.cfl = stage == 1 ? cfl_.p : nullptr;
                run_row(r);
                run_col(4 + stage, stage_args(dt, nullptr, 0));
            }
            if (ou_)
                launch_ou(1, dt / 2, nullptr);
            check(hipGetLastError(), "SSPRK3 kernels");
            ++step_;
            last_dt_ = dt;
            max_dt_ = std::max(max_dt_, dt);
        }
    }

    // Solver::drift (src/solver.cpp:255-263) of every member's state: the
    // right-hand side without forcing.  out: members reference Fields.
    void drift(double *out) override {
        flush_parameters();
        calib_buffers();
        drift_into(w0_.p, cout_.p);
        download_fields(cout_.p, out);
    }

    // Solver::linearized_drift (src/solver.cpp:264-290): the tangent of the
    // drift at each member's state along directions[m],
    // -(J(d, w) + J(w, d)) - (nu k^2 + alpha) d, with J(a, b) = u(a) . grad b
    // formed by the row pass from a two-state column pass.
    void linearized_drift(const double *directions, double *out) override {
        flush_parameters();
        calib_buffers();
        upload_fields(directions, caux_.p);
        jacobian(caux_.p, w0_.p, calA_.p, nlA_.p); // J(d, w)
        jacobian(w0_.p, caux_.p, calB_.p, nlB_.p); // J(w, d)
        StageArgs<R> a = stage_args(1.0, nullptr, 0);
        a.nl = nlA_.p;
        a.nl2 = nlB_.p;
        a.w1 = caux_.p;
        a.out = cout_.p;
        run_col(9, a);
        download_fields(cout_.p, out);
    }

    // Solver::adjoint_drift (src/solver.cpp:291-323) at each member's state
    // applied to cotangents[m]: J(w, lam) + i ky/k^2 F[lam wx]
    // - i kx/k^2 F[lam wy] - (nu k^2 + alpha) lam.
    void adjoint_drift(const double *cotangents, double *out) override {
        flush_parameters();
        calib_buffers();
        upload_fields(cotangents, caux_.p);
        adjoint_into(w0_.p, caux_.p, cout_.p);
        download_fields(cout_.p, out);
    }

    // likelihood_adjoint (src/adjoint.cpp:43-161) with member m's parameters
    // (log nu, log alpha) = log_params[2m, 2m+1]: forward RK4 from `initial`
    // without forcing, checkpoints every `stride` steps, Gaussian residuals at
    // the observations; then per checkpoint block (last first) a replay of
    // its states and the reverse sweep of each step's RK4 stages through the
    // batched adjoint drift.  counts (optional): checkpoint fields, peak
    // block fields, replayed steps (the reference's AdjointEvaluation).
    void likelihood_adjoint(const double *initial, double dt, size_t nobs, const size_t *ostep,
                            const int *ox, const int *oy, const double *oval, const double *osd,
                            const double *log_params, size_t stride, double *objective,
                            double *gradient, size_t *counts) override {
        if (nobs == 0 || stride == 0 || !std::isfinite(dt) || dt <= 0)
            throw std::invalid_argument("adjoint requires observations, dt > 0 and a positive "
                                        "checkpoint stride");
        for (size_t o = 0; o < nobs; ++o)
            if ((o > 0 && ostep[o] < ostep[o - 1]) || ox[o] < 0 || oy[o] < 0 || ox[o] >= n_ ||
                oy[o] >= n_ || !std::isfinite(oval[o]) || !std::isfinite(osd[o]) || osd[o] <= 0)
                throw std::invalid_argument("observations must be time ordered, finite, in "
                                            "bounds, and have positive uncertainty");
        // members outside the reference's domain get +inf and harmless parameters
        std::vector<char> valid(B_, 1);
        const double largest = (double)((n_ - 1) / 3);
        for (size_t m = 0; m < B_; ++m) {
            const double q0 = log_params[2 * m], q1 = log_params[2 * m + 1];
            valid[m] = std::isfinite(q0) && std::isfinite(q1) && q0 >= -12 && q0 <= 0 &&
                       q1 >= -12 && q1 <= 0;
            if (valid[m] &&
                dt * (2 * std::exp(q0) * largest * largest + std::exp(q1)) > 2.5)
                valid[m] = 0;
            nu_[m] = valid[m] ? std::exp(q0) : 0.0;
            alpha_[m] = valid[m] ? std::exp(q1) : 0.0;
            sigma_[m] = 0;
            amp_[m] = 0;
        }
        push_parameters();
        fill_state(initial);
        calib_buffers();
        if (pf_cs_.count != (size_t)n_) {
            check(hipStreamSynchronize(stream_), "table growth");
            clear_graphs();
            std::vector<double> cs(n_), sn(n_);
            for (int j = 0; j < n_; ++j) {
                cs[j] = std::cos(2 * PI * j / n_);
                sn[j] = std::sin(2 * PI * j / n_);
            }
            pf_cs_ = DeviceBuffer<double>(n_);
            pf_sn_ = DeviceBuffer<double>(n_);
            upload(pf_cs_.p, cs.data(), n_);
            upload(pf_sn_.p, sn.data(), n_);
        }
        const size_t last = ostep[nobs - 1];
        std::vector<size_t> cstep{0};
        for (size_t step = 1; step <= last; ++step)
            if (step % stride == 0 || step == last)
                cstep.push_back(step);
        const size_t nck = cstep.size(), SB = B_ * spec_;
        DeviceBuffer<cx<R>> ck(nck * SB);
        auto copy = [&](cx<R> *dst, const cx<R> *src) {
            check(hipMemcpyAsync(dst, src, SB * sizeof(cx<R>), hipMemcpyDeviceToDevice, stream_),
                  "state copy");
        };
        auto axpby = [&](cx<R> *y, double a, const cx<R> *x, double b, const cx<R> *z) {
            hipLaunchKernelGGL(k_axpby<R>, dim3((unsigned)((SB + 255) / 256)), dim3(256), 0,
                               stream_, (long)SB, y, (R)a, x, (R)b, z);
        };
        // observation groups: [first, end) rows sharing a step
        std::vector<size_t> gfirst;
        for (size_t o = 0; o < nobs; ++o)
            if (o == 0 || ostep[o] != ostep[o - 1])
                gfirst.push_back(o);
        gfirst.push_back(nobs);
        DeviceBuffer<int> d_ox(nobs), d_oy(nobs);
        upload(d_ox.p, ox, nobs);
        upload(d_oy.p, oy, nobs);
        std::vector<double> weight(B_ * nobs, 0.0), value(B_, 0.0);
        DeviceBuffer<double> d_pred(B_ * nobs), d_weight(B_ * nobs);
        // -- forward
        copy(ck.p, w0_.p);
        size_t g = 0, ci = 1;
        for (size_t step = 0; step <= last; ++step) {
            if (step > 0) {
                advance(dt, nullptr, 0);
                if (ci < nck && cstep[ci] == step)
                    copy(ck.p + (ci++) * SB, w0_.p);
            }
            if (g + 1 < gfirst.size() && ostep[gfirst[g]] == step) {
                const size_t o0 = gfirst[g], cnt = gfirst[g + 1] - o0;
                ensure_mixed(dt);
                PfArgs<R> pa{(int)B_, n_, K_, mix_.p, (int)cnt, d_ox.p + o0, d_oy.p + o0,
                             nullptr, nullptr, pf_cs_.p, pf_sn_.p, nullptr, nullptr, nullptr,
                             d_pred.p};
                const long total = (long)B_ * (long)cnt;
                hipLaunchKernelGGL(k_point_values<R>, dim3((unsigned)((total + 255) / 256)),
                                   dim3(256), 0, stream_, pa);
                std::vector<double> pred((size_t)total);
                check(hipMemcpyAsync(pred.data(), d_pred.p, pred.size() * sizeof(double),
                                     hipMemcpyDeviceToHost, stream_),
                      "predictions");
                check(hipStreamSynchronize(stream_), "predictions");
                for (size_t m = 0; m < B_; ++m)
                    for (size_t k = 0; k < cnt; ++k) {
                        const size_t o = o0 + k;
                        const double r = (pred[m * cnt + k] - oval[o]) / osd[o];
                        value[m] += 0.5 * r * r;
                        weight[m * nobs + o] = r / osd[o];
                    }
                ++g;
            }
        }
        upload(d_weight.p, weight.data(), weight.size());
        // -- reverse sweep
        std::vector<DeviceBuffer<cx<R>>> wk;
        for (int k = 0; k < 11; ++k)
            wk.emplace_back(SB);
        cx<R> *S2 = wk[0].p, *S3 = wk[1].p, *S4 = wk[2].p, *kb = wk[3].p, *cot = wk[4].p,
              *prev = wk[5].p, *K1 = wk[6].p, *K2 = wk[7].p, *K3 = wk[8].p, *K4 = wk[9].p,
              *bb = wk[10].p;
        hipLaunchKernelGGL(k_zero<R>, dim3((unsigned)((SB + 255) / 256)), dim3(256), 0, stream_,
                           (long)SB, cot);
        DeviceBuffer<double> d_grad(2 * B_);
        {
            std::vector<double> z(2 * B_, 0.0);
            upload(d_grad.p, z.data(), z.size());
        }
        size_t peak = 0, replayed = 0, obs_left = nobs;
        auto inject = [&](size_t step) {
            size_t hi = obs_left, lo = obs_left;
            while (lo > 0 && ostep[lo - 1] == step)
                --lo;
            if (lo == hi)
                return;
            // weights of rows [lo, hi): pass the member-major weights of those rows
            std::vector<double> w(B_ * (hi - lo));
            for (size_t m = 0; m < B_; ++m)
                for (size_t o = lo; o < hi; ++o)
                    w[m * (hi - lo) + (o - lo)] = weight[m * nobs + o];
            upload(d_pred.p, w.data(), w.size());
            hipLaunchKernelGGL(k_inject<R>, dim3((unsigned)((SB + 255) / 256)), dim3(256), 0,
                               stream_, (int)B_, n_, K_, cot, (int)(hi - lo), d_ox.p + lo,
                               d_oy.p + lo, (const double *)d_pred.p, (const double *)pf_cs_.p,
                               (const double *)pf_sn_.p);
            obs_left = lo;
        };
        auto terms = [&](const cx<R> *S, const cx<R> *W) {
            hipLaunchKernelGGL(k_param_terms<R>, dim3((unsigned)B_), dim3(256), 0, stream_, K_, S,
                               W, (const double *)d_nu_.p, (const double *)d_alpha_.p, d_grad.p);
        };
        for (size_t b = nck - 1; b >= 1; --b) {
            const size_t begin = cstep[b - 1], end = cstep[b], len = end - begin;
            DeviceBuffer<cx<R>> states((len + 1) * SB);
            copy(states.p, ck.p + (b - 1) * SB);
            for (size_t k = 0; k < len; ++k) {
                copy(w0_.p, states.p + k * SB);
                mix_valid_ = false;
                advance(dt, nullptr, 0);
                copy(states.p + (k + 1) * SB, w0_.p);
                ++replayed;
            }
            peak = std::max(peak, len + 1);
            for (size_t step = end; step > begin; --step) {
                inject(step);
                const cx<R> *S1 = states.p + (step - begin - 1) * SB;
                drift_into(S1, kb);
                axpby(S2, 1.0, S1, dt / 2, kb);
                drift_into(S2, kb);
                axpby(S3, 1.0, S1, dt / 2, kb);
                drift_into(S3, kb);
                axpby(S4, 1.0, S1, dt, kb);
                ++replayed;
                copy(prev, cot);
                axpby(K1, dt / 6, cot, 0.0, nullptr);
                axpby(K2, dt / 3, cot, 0.0, nullptr);
                axpby(K3, dt / 3, cot, 0.0, nullptr);
                axpby(K4, dt / 6, cot, 0.0, nullptr);
                terms(S4, K4);
                adjoint_into(S4, K4, bb);
                axpby(prev, 1.0, prev, 1.0, bb);
                axpby(K3, 1.0, K3, dt, bb);
                terms(S3, K3);
                adjoint_into(S3, K3, bb);
                axpby(prev, 1.0, prev, 1.0, bb);
                axpby(K2, 1.0, K2, dt / 2, bb);
                terms(S2, K2);
                adjoint_into(S2, K2, bb);
                axpby(prev, 1.0, prev, 1.0, bb);
                axpby(K1, 1.0, K1, dt / 2, bb);
                terms(S1, K1);
                adjoint_into(S1, K1, bb);
                axpby(prev, 1.0, prev, 1.0, bb);
                std::swap(cot, prev);
            }
            check(hipStreamSynchronize(stream_), "adjoint block");
        }
        check(hipGetLastError(), "adjoint kernels");
        std::vector<double> grad(2 * B_);
        check(hipMemcpy(grad.data(), d_grad.p, grad.size() * sizeof(double),
                        hipMemcpyDeviceToHost),
              "gradient download");
        for (size_t m = 0; m < B_; ++m) {
            objective[m] = valid[m] ? value[m] : std::numeric_limits<double>::infinity();
            gradient[2 * m] = valid[m] ? grad[2 * m] : 0.0;
            gradient[2 * m + 1] = valid[m] ? grad[2 * m + 1] : 0.0;
        }
        if (counts) {
            counts[0] = nck;
            counts[1] = peak;
            counts[2] = replayed;
        }
        mix_valid_ = false;
        params_dirty_ = false;
    }

    // Solver::step_tangent (src/solver.cpp:324-380), batched: `steps` RK4 steps
    // of every member together with its two sensitivities (with respect to
    // log nu and log alpha).  tangents: [member][2][n*n] reference Fields,
    // read (tangents_in) and written (tangents_out, may alias).
    void step_tangent(double dt, size_t steps, const double *tangents_in,
                      double *tangents_out) override {
        for (size_t m = 0; m < B_; ++m)
            if (sigma_[m] != 0 || amp_[m] != 0)
                throw std::invalid_argument(
                    "tangent implementation currently requires zero forcing");
        validate_dt(dt);
        flush_parameters();
        tangent_buffers();
        const size_t nn = (size_t)n_ * n_;
        std::vector<double> f(2 * nn * B_);
        for (int j = 0; j < 2; ++j) {
            for (size_t m = 0; m < B_; ++m)
                std::memcpy(f.data() + 2 * nn * m, tangents_in + 2 * nn * (2 * m + j),
                            2 * nn * sizeof(double));
            upload_fields(f.data(), tg_[j].p);
        }
        for (size_t s = 0; s < steps; ++s)
            tangent_step(dt);
        for (int j = 0; j < 2; ++j) {
            download_fields(tg_[j].p, f.data());
            for (size_t m = 0; m < B_; ++m)
                std::memcpy(tangents_out + 2 * nn * (2 * m + j), f.data() + 2 * nn * m,
                            2 * nn * sizeof(double));
        }
    }

    // likelihood_gradient (src/inference.cpp:340-382), batched over members'
    // (log nu, log alpha): the deterministic negative log-likelihood and its
    // gradient by forward sensitivities (step_tangent from zero tangents).
    void likelihood_gradient(const double *initial, double dt, size_t nobs, const size_t *ostep,
                             const int *ox, const int *oy, const double *oval,
                             const double *osd, const double *log_params, double *objective,
                             double *gradient) override {
        if (nobs == 0 || !std::isfinite(dt) || dt <= 0)
            throw std::invalid_argument("likelihood gradient requires observations and dt > 0");
        for (size_t o = 0; o < nobs; ++o)
            if ((o > 0 && ostep[o] < ostep[o - 1]) || ox[o] < 0 || oy[o] < 0 || ox[o] >= n_ ||
                oy[o] >= n_ || !std::isfinite(oval[o]) || !std::isfinite(osd[o]) || osd[o] <= 0)
                throw std::invalid_argument("observations must be time ordered, finite, in "
                                            "bounds, and have positive uncertainty");
        std::vector<char> valid(B_, 1);
        const double largest = (double)((n_ - 1) / 3);
        for (size_t m = 0; m < B_; ++m) {
            const double q0 = log_params[2 * m], q1 = log_params[2 * m + 1];
            valid[m] = std::isfinite(q0) && std::isfinite(q1) && q0 >= -12 && q0 <= 0 &&
                       q1 >= -12 && q1 <= 0;
            if (valid[m] &&
                dt * (2 * std::exp(q0) * largest * largest + std::exp(q1)) > 2.5)
                valid[m] = 0;
            nu_[m] = valid[m] ? std::exp(q0) : 0.0;
            alpha_[m] = valid[m] ? std::exp(q1) : 0.0;
            sigma_[m] = 0;
            amp_[m] = 0;
        }
        push_parameters();
        fill_state(initial);
        calib_buffers();
        tangent_buffers();
        point_tables();
        for (int j = 0; j < 2; ++j)
            hipLaunchKernelGGL(k_zero<R>, dim3((unsigned)((B_ * spec_ + 255) / 256)), dim3(256),
                               0, stream_, (long)(B_ * spec_), tg_[j].p);
        DeviceBuffer<int> d_ox(nobs), d_oy(nobs);
        upload(d_ox.p, ox, nobs);
        upload(d_oy.p, oy, nobs);
        DeviceBuffer<double> d_pred(3 * B_ * nobs);
        std::vector<double> value(B_, 0.0), g0(B_, 0.0), g1(B_, 0.0);
        size_t step = 0, o = 0;
        while (o < nobs) {
            while (step < ostep[o]) {
                tangent_step(dt);
                ++step;
            }
            size_t hi = o;
            while (hi < nobs && ostep[hi] == step)
                ++hi;
            const size_t cnt = hi - o;
            // point values of the state (mixed fields of w0) and of both tangents
            const cx<R> *mixes[3];
            ensure_mixed(dt);
            mixes[0] = mix_.p;
            for (int j = 0; j < 2; ++j) {
                StageArgs<R> m0 = stage_args(1.0, nullptr, 0);
                m0.w0 = tg_[j].p;
                m0.mix = j == 0 ? calA_.p : calB_.p;
                run_col(0, m0);
                mixes[1 + j] = m0.mix;
            }
            for (int f = 0; f < 3; ++f) {
                PfArgs<R> pa{(int)B_, n_, K_, mixes[f], (int)cnt, d_ox.p + o, d_oy.p + o,
                             nullptr, nullptr, pf_cs_.p, pf_sn_.p, nullptr, nullptr, nullptr,
                             d_pred.p + (size_t)f * B_ * cnt};
                const long total = (long)(B_ * cnt);
                hipLaunchKernelGGL(k_point_values<R>, dim3((unsigned)((total + 255) / 256)),
                                   dim3(256), 0, stream_, pa);
            }
            std::vector<double> pred(3 * B_ * cnt);
            check(hipMemcpyAsync(pred.data(), d_pred.p, pred.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, stream_),
                  "predictions");
            check(hipStreamSynchronize(stream_), "predictions");
            for (size_t m = 0; m < B_; ++m)
                for (size_t k = 0; k < cnt; ++k) {
                    const size_t r = o + k;
                    const double res = (pred[m * cnt + k] - oval[r]) / osd[r];
                    value[m] += 0.5 * res * res;
                    g0[m] += res * pred[(B_ + m) * cnt + k] / osd[r];
                    g1[m] += res * pred[(2 * B_ + m) * cnt + k] / osd[r];
                }
            o = hi;
        }
        check(hipGetLastError(), "tangent kernels");
        for (size_t m = 0; m < B_; ++m) {
            objective[m] = valid[m] ? value[m] : std::numeric_limits<double>::infinity();
            gradient[2 * m] = valid[m] ? g0[m] : 0.0;
            gradient[2 * m + 1] = valid[m] ? g1[m] : 0.0;
        }
        mix_valid_ = false;
    }

    // ------------------------------------------------ launch shapes ----
    int col_variants() const override {
        return std::is_same_v<R, float> ? col_variant_count(n_) : 1;
    }
    int row_variants() const override {
        return std::is_same_v<R, float> ? row_variant_count(n_) : 1;
    }

    void set_tuning(int cv, int rv) override {
        if (cv < 0 || rv < 0 || cv >= col_variants() || rv >= row_variants())
            throw std::invalid_argument("no such launch-shape variant for this grid and precision");
        check(hipStreamSynchronize(stream_), "tuning switch");
        clear_graphs(); // captured launches name the previous kernels
        const int ec = cv ? col_variant_elems(n_, cv) : default_split<R>(n_);
        const int er = rv ? row_variant_elems(n_, rv) : default_split<R>(n_);
        tw_col_ = twiddle_table(ec);
        tw_row_ = twiddle_table(er);
        col_var_ = cv;
        row_var_ = rv;
        // mixed and nonlinear fields are stored in natural y order, so they
        // stay valid across variants
    }

    void get_tuning(int *cv, int *rv) const override {
        *cv = col_var_;
        *rv = row_var_;
    }

    // Time every column variant (with the default row shape), keep the
    // fastest, then every row variant with it.  A variant replaces the
    // default only when it is more than 1% faster.  The states are restored
    // afterwards.  col_ms / row_ms (optional) receive the per-step kernel
    // times of each variant.  With use_cache, a remembered choice is applied
    // without timing (the outputs are then left unchanged).
    void tune(bool use_cache, double *col_ms, double *row_ms) override {
        const int nc = col_variants(), nr = row_variants();
        if (nc == 1 && nr == 1)
            return;
        const std::string key = tune_key(n_, sizeof(R) == 4 ? HZL_FP32 : HZL_FP64, B_);
        TuneChoice remembered;
        if (use_cache && tune_lookup(key, &remembered) && remembered.cv < nc &&
            remembered.rv < nr) {
            set_tuning(remembered.cv, remembered.rv);
            return;
        }
        flush_parameters();
        DeviceBuffer<cx<R>> backup(B_ * spec_);
        check(hipMemcpyAsync(backup.p, w0_.p, B_ * spec_ * sizeof(cx<R>),
                             hipMemcpyDeviceToDevice, stream_),
              "tuning backup");
        const bool mix_was_valid = mix_valid_;
        const double dt = 1e-6;
        // per-step (col, row) kernel milliseconds of one variant pair: the
        // minimum over 3 timed steps after one untimed step
        auto measure = [&](int c, int r) {
            set_tuning(c, r);
            run_col(0, stage_args(dt, nullptr, 0));
            hipEvent_t ev[9];
            for (auto &e : ev)
                check(hipEventCreate(&e), "hipEventCreate");
            double best_c = 1e300, best_r = 1e300;
            for (int rep = 0; rep < 4; ++rep) {
                launch_step(dt, nullptr, 0, nullptr, step_, ev);
                check(hipEventSynchronize(ev[8]), "tuning step");
                double tc = 0, tr = 0;
                for (int i = 0; i < 8; ++i) {
                    float ms = 0;
                    check(hipEventElapsedTime(&ms, ev[i], ev[i + 1]), "hipEventElapsedTime");
                    (i % 2 == 0 ? tr : tc) += ms;
                }
                if (rep > 0) {
                    best_c = std::min(best_c, tc);
                    best_r = std::min(best_r, tr);
                }
            }
            for (auto &e : ev)
                (void)hipEventDestroy(e);
            return std::make_pair(best_c, best_r);
        };
        // the fastest variant if it beats the default by more than 1%; among
        // variants within 0.5% of it, the lowest-numbered (simplest) one, so
        // near ties resolve the same way on every run
        auto pick = [](const std::vector<double> &t) {
            int best = 0;
            for (int i = 1; i < (int)t.size(); ++i)
                if (t[i] < 0.99 * t[0] && t[i] < t[best])
                    best = i;
            for (int i = 1; i < best; ++i)
                if (t[i] < 0.99 * t[0] && t[i] <= 1.005 * t[best])
                    return i;
            return best;
        };
        std::vector<double> ct(nc), rt(nr);
        for (int c = 0; c < nc; ++c)
            ct[c] = measure(c, 0).first;
        const int bc = pick(ct);
        for (int r = 0; r < nr; ++r)
            rt[r] = measure(bc, r).second;
        const int br = pick(rt);
        check(hipMemcpyAsync(w0_.p, backup.p, B_ * spec_ * sizeof(cx<R>),
                             hipMemcpyDeviceToDevice, stream_),
              "tuning restore");
        check(hipMemsetAsync(cfl_.p, 0, cfl_.count * sizeof(double), stream_), "reset CFL");
        check(hipStreamSynchronize(stream_), "tuning restore");
        mix_valid_ = false;
        if (mix_was_valid)
            ensure_mixed(1.0);
        set_tuning(bc, br);
        tune_store(key, TuneChoice{bc, br});
        if (col_ms)
            std::copy(ct.begin(), ct.end(), col_ms);
        if (row_ms)
            std::copy(rt.begin(), rt.end(), row_ms);
    }

    void step_forced(double dt, const double *forcing, bool shared) override {
        validate_dt(dt);
        const size_t nn = (size_t)n_ * n_, count = shared ? 1 : B_;
        for (size_t k = 0; k < count; ++k) {
            const double *F = forcing + 2 * nn * k;
            for (size_t i = 0; i < nn; ++i) {
                const double re = F[2 * i], im = F[2 * i + 1];
                if (!std::isfinite(re) || !std::isfinite(im))
                    throw std::invalid_argument("nonfinite GPU forcing");
                const size_t x = i % n_, y = i / n_;
                const size_t j = ((n_ - y) % n_) * n_ + (n_ - x) % n_;
                const double dr = re - F[2 * j], di = im + F[2 * j + 1];
                if (std::sqrt(dr * dr + di * di) > 1e-12 * (1 + std::hypot(re, im)))
                    throw std::invalid_argument("GPU forcing must preserve conjugate symmetry");
            }
        }
        std::vector<cx<R>> half(spec_ * count);
        for (size_t k = 0; k < count; ++k)
            full_to_half<R>(n_, forcing + 2 * nn * k, half.data() + spec_ * k);
        upload(ext_.p, half.data(), half.size());
        advance(dt, ext_.p, shared ? 0 : (long long)spec_);
    }

    void synchronize() override {
        check(hipStreamSynchronize(stream_), "synchronize GPU");
        check(hipGetLastError(), "GPU kernels");
        std::vector<double> c(cfl_.count);
        check(hipMemcpy(c.data(), cfl_.p, c.size() * sizeof(double), hipMemcpyDeviceToHost),
              "CFL download");
        double mx = 0;
        bool bad = false;
        for (double v : c) {
            if (!std::isfinite(v))
                bad = true;
            else
                mx = std::max(mx, v);
        }
        check(hipMemsetAsync(cfl_.p, 0, cfl_.count * sizeof(double), stream_), "reset CFL");
        check(hipStreamSynchronize(stream_), "reset CFL");
        const double rate = (double)mx * n_ / (2 * PI);
        last_cfl_rate_ = rate;
        const double dt = max_dt_;
        max_dt_ = 0;
        if (bad)
            throw std::runtime_error("nonfinite flow state");
        if (dt * rate > 0.5)
            throw std::runtime_error("GPU CFL limit exceeded");
    }

    void diagnostics(double *out) override {
        DeviceBuffer<double> staging(B_ * 6);
        diagnostics_with(out, staging.p);
    }

    // Diagnostics through `staging` (6 doubles per member, device memory the
    // caller owns); allocates nothing.  Every entry of the B x 6 block is
    // written before it is read.
    void diagnostics_with(double *out, double *staging) override {
        flush_parameters();
        // CFL rate of the current state (src/solver.cpp:392-396): one row pass
        // over the mixed fields of the current state into a private buffer.
        if (!mix_valid_) {
            run_col(0, stage_args(1.0, nullptr, 0));
            mix_valid_ = true;
        }
        const long cfl_count = (long)cfl_diag_.count;
        hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((cfl_count + 255) / 256)), dim3(256), 0,
                           stream_, cfl_count, 0.0, cfl_diag_.p);
        RowArgs<R> r{(int)B_, mix_.p, nl_.p, tw_row_, cfl_diag_.p};
        run_row(r);
        // white forcing: the noise buffer still holds the last step's shell
        // noise; OU: the current coefficients (Solver::forcing(), src/solver.cpp:389)
        if (ou_)
            launch_ou(3, 0.0, noise_.p);
        DiagArgs<R> a{K_, w0_.p, d_nu_.p, forcing_model(), (ou_ || step_ > 0) ? 1 : 0,
                      cfl_diag_.p, n_ / 2, n_, staging};
        hipLaunchKernelGGL(k_diag<R>, dim3((unsigned)B_), dim3(256), 0, stream_, a);
        check(hipGetLastError(), "diagnostics kernel");
        check(hipMemcpyAsync(out, staging, B_ * 6 * sizeof(double), hipMemcpyDeviceToHost,
                             stream_),
              "diagnostics download");
        check(hipStreamSynchronize(stream_), "diagnostics");
    }

    size_t members() const override { return B_; }
    int device() const override { return device_; }

    void energy_spectrum(double *out) override {
        DeviceBuffer<double> bins(B_ * (size_t)n_);
        hipLaunchKernelGGL(k_spectrum<R>, dim3((unsigned)B_), dim3(256), 0, stream_, K_, n_,
                           (long)spec_, (const cx<R> *)w0_.p, bins.p);
        check(hipGetLastError(), "spectrum kernel");
        check(hipMemcpyAsync(out, bins.p, bins.count * sizeof(double), hipMemcpyDeviceToHost,
                             stream_),
              "spectrum download");
        check(hipStreamSynchronize(stream_), "spectrum");
    }

    void profile_step(double dt, float *ms) override {
        validate_dt(dt);
        hipEvent_t ev[9];
        for (auto &e : ev)
            check(hipEventCreate(&e), "hipEventCreate");
        advance(dt, nullptr, 0, ev);
        check(hipEventSynchronize(ev[8]), "profile step");
        for (int i = 0; i < 8; ++i)
            check(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]), "hipEventElapsedTime");
        for (auto &e : ev)
            (void)hipEventDestroy(e);
    }

    // ------------------------------------------------ particle filter ----
    void pf_begin(uint64_t seed, size_t max_groups) override {
        pf_begin_filters(&seed, 1, max_groups, false);
    }

    // F independent filters over contiguous member segments of B / F
    // particles; filter f draws its resampling uniforms from
    // mt19937_64(stream_seed(seeds[f], UINT64_MAX)) and reseeds children with
    // stream_seed(seeds[f], (g+1) P + i) (src/filter.cpp:54,115).  With
    // reset_keys, particle i of filter f starts on key stream_seed(seeds[f], i)
    // (the reference's Solver(n, p, stream_seed(seed, i)) per particle).
    void pf_begin_filters(const uint64_t *seeds, size_t filters, size_t max_groups,
                          bool reset_keys) override {
        if (filters == 0 || B_ % filters != 0)
            throw std::invalid_argument("members must split evenly into the filters");
        const size_t P = B_ / filters;
        if (P < 2)
            throw std::invalid_argument("particle filter needs at least two particles");
        check(hipStreamSynchronize(stream_), "filter start");
        clear_graphs(); // captured cycles hold the previous filter's buffers
        pf_filters_ = filters;
        pf_group_ = 0;
        pf_wstate_ = DeviceBuffer<double>(B_);
        pf_wnorm_ = DeviceBuffer<double>(B_);
        pf_logw_ = DeviceBuffer<double>(B_);
        pf_stats_ = DeviceBuffer<double>(8 * filters);
        pf_anc_ = DeviceBuffer<int>(B_);
        pf_mean_ = DeviceBuffer<double>(2 * spec_);
        if (pf_host_)
            (void)hipHostFree(pf_host_);
        pf_host_ = nullptr;
        check(hipHostMalloc(reinterpret_cast<void **>(&pf_host_),
                            (8 * filters + 2 * spec_) * sizeof(double)),
              "hipHostMalloc");
        pf_mean_part_ = DeviceBuffer<double>(
            2 * spec_ * ((B_ + pf_mean_chunk((long)B_) - 1) / pf_mean_chunk((long)B_)));
        std::vector<double> w(B_, 1.0 / (double)P), st(8 * filters, 0.0), cs(n_), sn(n_);
        upload(pf_wstate_.p, w.data(), B_);
        upload(pf_stats_.p, st.data(), 8 * filters);
        // resampling uniforms, one per resampling event, (x >> 11) 2^-53
        // (src/filter.cpp:11-13,54)
        const size_t U = std::max<size_t>(max_groups, 1);
        std::vector<double> u(U * filters);
        for (size_t f = 0; f < filters; ++f) {
            std::mt19937_64 rng(stream_seed(seeds[f], std::numeric_limits<uint64_t>::max()));
            for (size_t k = 0; k < U; ++k)
                u[f * U + k] = (double)(rng() >> 11) * 0x1.0p-53;
        }
        pf_uniforms_ = DeviceBuffer<double>(u.size());
        upload(pf_uniforms_.p, u.data(), u.size());
        pf_max_uniforms_ = (int)U;
        pf_seeds_ = DeviceBuffer<uint64_t>(filters);
        upload(pf_seeds_.p, seeds, filters);
        if (reset_keys) {
            for (size_t f = 0; f < filters; ++f)
                for (size_t i = 0; i < P; ++i)
                    keys_[f * P + i] = stream_seed(seeds[f], i);
            upload(d_keys_.p, keys_.data(), B_);
            if (ou_) {
                // new particles: Solver(n, p, stream_seed(seed_f, i)) draws its
                // stationary OU start (src/solver.cpp:120-139)
                flush_parameters();
                launch_ou(0, 0.0, nullptr);
                check(hipGetLastError(), "OU start");
            }
        }
        for (int j = 0; j < n_; ++j) {
            cs[j] = std::cos(2 * PI * j / n_);
            sn[j] = std::sin(2 * PI * j / n_);
        }
        pf_cs_ = DeviceBuffer<double>(n_);
        pf_sn_ = DeviceBuffer<double>(n_);
        upload(pf_cs_.p, cs.data(), n_);
        upload(pf_sn_.p, sn.data(), n_);
        pf_on_ = true;
    }

    void pf_assimilate(double dt, size_t steps, size_t nobs, const int *x, const int *y,
                       const double *value, const double *sd, bool is_last, double *record,
                       double *mean_field) override {
        if (!pf_on_)
            throw std::invalid_argument("particle filter not started");
        if (nobs == 0)
            throw std::invalid_argument("observation group is empty");
        if (mean_field && pf_filters_ != 1)
            throw std::invalid_argument("mean fields are recorded for a single filter only");
        for (size_t o = 0; o < nobs; ++o)
            if (x[o] < 0 || y[o] < 0 || x[o] >= n_ || y[o] >= n_ || !std::isfinite(value[o]) ||
                !std::isfinite(sd[o]) || sd[o] <= 0)
                throw std::invalid_argument("observations must be finite, in bounds, and have "
                                            "positive uncertainty");
        if (!is_last && pf_group_ >= (uint64_t)pf_max_uniforms_)
            throw std::invalid_argument(
                "more observation groups than the filter was started for (max_groups)");
        if (steps)
            validate_dt(dt);
        stage_observations(nobs, x, y, value, sd);
        if (use_graph_ && steps >= 1 && steps <= kGraphSteps) {
            // the whole cycle (forecast, weights, normalisation, resampling
            // copies) replays one captured launch sequence; the step index,
            // group and last-group flag come from ctr_
            flush_parameters();
            ensure_mixed(dt);
            set_counters(pf_group_, is_last ? 1 : 0);
            const GraphKey key{2, dt, steps, nobs, pf_filters_, obs_dev_.p, mean_field != nullptr};
            run_captured(key, [&]() {
                for (size_t s = 0; s < steps; ++s)
                    launch_step(dt, nullptr, 0, ctr_.p, s, nullptr);
                launch_filter(nobs, true, is_last, mean_field != nullptr);
            });
            step_ += steps;
            last_dt_ = dt;
            max_dt_ = std::max(max_dt_, dt);
        } else {
            if (steps)
                step(dt, steps);
            flush_parameters();
            ensure_mixed(dt);
            launch_filter(nobs, false, is_last, mean_field != nullptr);
        }
        check(hipGetLastError(), "particle filter kernels");
        mix_valid_ = false;
        // statistics and mean land in pinned staging (pf_begin): direct copies
        const size_t nst = 8 * pf_filters_;
        double *st = pf_host_, *mean = pf_host_ + nst;
        check(hipMemcpyAsync(st, pf_stats_.p, nst * sizeof(double), hipMemcpyDeviceToHost,
                             stream_),
              "filter statistics");
        if (mean_field)
            check(hipMemcpyAsync(mean, pf_mean_.p, 2 * spec_ * sizeof(double),
                                 hipMemcpyDeviceToHost, stream_),
                  "filter mean");
        check(hipStreamSynchronize(stream_), "particle filter");
        for (size_t f = 0; f < pf_filters_; ++f)
            if (record)
                for (int k = 0; k < 4; ++k)
                    record[4 * f + k] = st[8 * f + k];
        if (mean_field)
            half_to_full<double>(n_, reinterpret_cast<const cx<double> *>(mean), mean_field);
        ++pf_group_;
        // the children's keys were reseeded on the device; d_keys_ is the
        // authoritative copy (keys_ only stages whole or single-member uploads)
    }

    void pf_result(double *log_likelihood, size_t *resampling_events) override {
        if (!pf_on_)
            throw std::invalid_argument("particle filter not started");
        std::vector<double> st(8 * pf_filters_);
        check(hipStreamSynchronize(stream_), "filter result");
        check(hipMemcpy(st.data(), pf_stats_.p, st.size() * sizeof(double),
                        hipMemcpyDeviceToHost),
              "filter result");
        for (size_t f = 0; f < pf_filters_; ++f) {
            if (log_likelihood)
                log_likelihood[f] = st[8 * f + 4];
            if (resampling_events)
                resampling_events[f] = (size_t)st[8 * f + 5];
        }
    }

    // Log-weights of every member for one observation group, with running
    // weights `wprev` and output `logw` in caller-owned device memory (a
    // filter spanning several ensembles), on this ensemble's stream.
    void pf_logweights(size_t nobs, const int *x, const int *y, const double *value,
                       const double *sd, const double *wprev, double *logw) override {
        pf_stage_logweights(nobs, x, y, value, sd);
        if (!mix_valid_) {
            run_col(0, stage_args(1.0, nullptr, 0));
            mix_valid_ = true;
        }
        pf_enqueue_logweights(nobs, wprev, logw);
    }

    void pf_stage_logweights(size_t nobs, const int *x, const int *y, const double *value,
                             const double *sd) override {
        if (nobs == 0)
            throw std::invalid_argument("observation group is empty");
        for (size_t o = 0; o < nobs; ++o)
            if (x[o] < 0 || y[o] < 0 || x[o] >= n_ || y[o] >= n_ || !std::isfinite(value[o]) ||
                !std::isfinite(sd[o]) || sd[o] <= 0)
                throw std::invalid_argument("observations must be finite, in bounds, and have "
                                            "positive uncertainty");
        flush_parameters();
        if (pf_cs_.count != (size_t)n_) {
            check(hipStreamSynchronize(stream_), "table growth");
            clear_graphs();
            std::vector<double> cs(n_), sn(n_);
            for (int j = 0; j < n_; ++j) {
                cs[j] = std::cos(2 * PI * j / n_);
                sn[j] = std::sin(2 * PI * j / n_);
            }
            pf_cs_ = DeviceBuffer<double>(n_);
            pf_sn_ = DeviceBuffer<double>(n_);
            upload(pf_cs_.p, cs.data(), n_);
            upload(pf_sn_.p, sn.data(), n_);
        }
        stage_observations(nobs, x, y, value, sd);
    }

    void pf_enqueue_logweights(size_t nobs, const double *wprev, double *logw) override {
        PfArgs<R> a{(int)B_,  n_,       K_,       mix_.p,    (int)nobs,
                    obs_x_,   obs_y_,   obs_v_,   obs_sd_,   pf_cs_.p,
                    pf_sn_.p, const_cast<double *>(wprev), nullptr, logw, nullptr};
        hipLaunchKernelGGL(k_pf_loglik<R>, dim3((unsigned)B_), dim3(256), 0, stream_, a);
    }

    size_t pf_filters() const override {
        return pf_on_ ? pf_filters_ : 0;
    }

    // Upload one field and replicate it into every member on the device.
    void fill_state(const double *field) override {
        set_state(0, field);
        for (size_t m = 1; m < B_; ++m)
            check(hipMemcpyAsync(w0_.p + m * spec_, w0_.p, spec_ * sizeof(cx<R>),
                                 hipMemcpyDeviceToDevice, stream_),
                  "state fill");
        mix_valid_ = false;
    }

    void set_step_index(uint64_t step) override {
        step_ = step;
    }

    uint64_t step_index() const override {
        return step_;
    }

    void device_state(void **ptr, size_t *stride) override {
        *ptr = w0_.p;
        *stride = spec_ * sizeof(cx<R>);
    }

    void *stream() override {
        return stream_;
    }

    EngineViews views() override {
        flush_parameters();
        EngineViews v{n_,        K_,        (int)B_,        sizeof(R) == 4 ? 0 : 1,
                      w0_.p,     d_nu_.p,   d_alpha_.p,     d_sigma_.p,
                      d_amp_.p,  d_keys_.p, pairs_.p,       p_.forcing_min,
                      p_.forcing_max, p_.kolmogorov_wavenumber, stream_};
        if (ou_) {
            v.ou_F = ou_F_.p;
            v.npairs = npairs_;
            v.ou_off = ou_off_.p;
        }
        return v;
    }

    void prepare_steps(double dt) override {
        validate_dt(dt);
        flush_parameters();
        set_counters(0, 0);
    }

    void enqueue_steps(double dt, size_t steps) override {
        run_col(0, stage_args(dt, nullptr, 0)); // the current state's mixed fields
        for (size_t s = 0; s < steps; ++s)
            launch_step(dt, nullptr, 0, ctr_.p, s, nullptr);
    }

    void commit_steps(double dt, size_t steps) override {
        step_ += steps;
        last_dt_ = dt;
        max_dt_ = std::max(max_dt_, dt);
        mix_valid_ = true; // enqueue_steps left the mixed fields of the new state
    }

    uint64_t launch_epoch() const override { return launch_epoch_; }
    bool graphs_enabled() const override { return use_graph_; }

    void invalidate_mixed() override {
        mix_valid_ = false;
    }

    void advance_counter(uint64_t steps) override {
        step_ += steps;
    }

    void rewind_counter(uint64_t steps) override {
        if (steps > step_)
            throw std::logic_error("step counter rewound past zero");
        step_ -= steps;
    }

  private:
    void member_check(size_t m) const {
        if (m >= B_)
            throw std::invalid_argument("member index out of range");
    }

    template <typename T> void upload(T *dst, const T *src, size_t count) {
        check(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, stream_),
              "upload");
        check(hipStreamSynchronize(stream_), "upload");
    }

    void push_parameters() {
        upload(d_nu_.p, nu_.data(), B_);
        upload(d_alpha_.p, alpha_.data(), B_);
        upload(d_sigma_.p, sigma_.data(), B_);
        upload(d_tau_.p, tau_.data(), B_);
        upload(d_amp_.p, amp_.data(), B_);
        params_dirty_ = false;
    }

    void flush_parameters() {
        if (params_dirty_)
            push_parameters();
    }

    // Explicit-step guard of CudaSolver::advance (src/cuda_solver.cu:203-208):
    // dt * max_retained(nu k^2 + alpha) <= 2.5 for every member.
    void validate_dt(double dt) {
        if (!std::isfinite(dt) || dt <= 0)
            throw std::invalid_argument("invalid GPU time step or diffusion stability limit exceeded");
        const double k2 = 2.0 * K_ * K_;
        for (size_t m = 0; m < B_; ++m)
            if (dt * (nu_[m] * k2 + alpha_[m]) > 2.5)
                throw std::invalid_argument(
                    "invalid GPU time step or diffusion stability limit exceeded");
    }

    ForcingModel<R> forcing_model() const {
        return ForcingModel<R>{noise_.p,       npairs_,        d_amp_.p,
                               pairs_.p,       p_.forcing_min, p_.forcing_max,
                               p_.kolmogorov_wavenumber};
    }

    StageArgs<R> stage_args(double dt, const cx<R> *ext, long long ext_stride) const {
        StageArgs<R> a;
        a.B = (int)B_;
        a.w0 = w0_.p;
        a.w1 = w1_.p;
        a.acc = acc_.p;
        a.nl = nl_.p;
        a.mix = mix_.p;
        a.tw = tw_col_;
        a.nu = d_nu_.p;
        a.alpha = d_alpha_.p;
        a.forcing = forcing_model();
        a.ext = ext;
        a.ext_stride = ext_stride;
        a.h_half = (R)(dt / 2);
        a.h_full = (R)dt;
        a.h_sixth = (R)(dt / 6);
        return a;
    }

    // Column / row launches through the selected launch-shape variant.
    void run_col(int mode, const StageArgs<R> &a) {
        const hipStream_t s = stream_;
        if (mode > 4 && col_var_ > 0) {
            // SSPRK3 stages and drift passes run in the default shape (the
            // mixed and nonlinear fields are shape independent)
            StageArgs<R> d = a;
            d.tw = tw_.p;
            dispatch_col<R>(n_, mode, d, s);
            return;
        }
        if constexpr (std::is_same_v<R, float>) {
            if (col_var_ > 0) {
                if (!launch_col_variant(n_, col_var_, mode, a, s))
                    throw std::logic_error("unknown column variant");
                return;
            }
        }
        dispatch_col<R>(n_, mode, a, s);
    }

    void run_row(const RowArgs<R> &a) {
        const hipStream_t s = stream_;
        if constexpr (std::is_same_v<R, float>) {
            if (row_var_ > 0) {
                if (!launch_row_variant(n_, row_var_, a, s))
                    throw std::logic_error("unknown row variant");
                return;
            }
        }
        dispatch_row<R>(n_, a, s);
    }

    // Twiddle table of the T x E split (the default one is tw_).
    const cx<R> *twiddle_table(int E) {
        if (E == default_split<R>(n_))
            return tw_.p;
        for (auto &t : tw_extra_)
            if (t.first == E)
                return t.second.p;
        const auto tw = twiddles<R>(n_, E);
        tw_extra_.emplace_back(E, DeviceBuffer<cx<R>>(tw.size()));
        upload(tw_extra_.back().second.p, tw.data(), tw.size());
        return tw_extra_.back().second.p;
    }

    // Buffers of the calibration drifts, allocated on first use.
    void calib_buffers() {
        if (caux_.count)
            return;
        caux_ = DeviceBuffer<cx<R>>(B_ * spec_);
        cout_ = DeviceBuffer<cx<R>>(B_ * spec_);
        calA_ = DeviceBuffer<cx<R>>(B_ * 4 * mixf_);
        calB_ = DeviceBuffer<cx<R>>(B_ * 4 * mixf_);
        nlA_ = DeviceBuffer<cx<R>>(B_ * mixf_);
        nlB_ = DeviceBuffer<cx<R>>(B_ * mixf_);
        nlC_ = DeviceBuffer<cx<R>>(B_ * mixf_);
    }

    // Tangent state: two sensitivity fields and the step's work fields, plus
    // the calibration buffers the tangent drift runs through.
    void tangent_buffers() {
        calib_buffers();
        if (tg_[0].count)
            return;
        for (auto &b : tg_)
            b = DeviceBuffer<cx<R>>(B_ * spec_);
        for (auto &b : tw_fields_)
            b = DeviceBuffer<cx<R>>(B_ * spec_);
        slope1_ = DeviceBuffer<cx<R>>(B_ * spec_);
    }

    // cos / sin (2 pi j / n) for point evaluation.
    void point_tables() {
        if (pf_cs_.count == (size_t)n_)
            return;
        check(hipStreamSynchronize(stream_), "table growth");
        clear_graphs();
        std::vector<double> cs(n_), sn(n_);
        for (int j = 0; j < n_; ++j) {
            cs[j] = std::cos(2 * PI * j / n_);
            sn[j] = std::sin(2 * PI * j / n_);
        }
        pf_cs_ = DeviceBuffer<double>(n_);
        pf_sn_ = DeviceBuffer<double>(n_);
        upload(pf_cs_.p, cs.data(), n_);
        upload(pf_sn_.p, sn.data(), n_);
    }

    // Tangent of the drift at X along D into out.
    void linearized_into(const cx<R> *X, const cx<R> *D, cx<R> *out) {
        if (!calB_.p || !nlB_.p || !out)
            throw std::logic_error("tangent pass on unallocated buffers");
        jacobian(D, X, calA_.p, nlA_.p); // J(D, X)
        jacobian(X, D, calB_.p, nlB_.p); // J(X, D)
        StageArgs<R> a = stage_args(1.0, nullptr, 0);
        a.nl = nlA_.p;
        a.nl2 = nlB_.p;
        a.w1 = const_cast<cx<R> *>(D);
        a.out = out;
        run_col(9, a);
    }

    // One step of Solver::step_tangent (src/solver.cpp:324-380) for all
    // members (no forcing): the state w0_ by RK4 through the drift, and the
    // sensitivities tg_[0..1] by RK4 of t' = L_X t - (nu k^2 | alpha) X at the
    // same stages, accumulated in the reference's order.
    void tangent_step(double dt) {
        const size_t SB = B_ * spec_;
        const dim3 grid((unsigned)((SB + 255) / 256));
        auto axpby = [&](cx<R> *y, double a, const cx<R> *x, double b, const cx<R> *z) {
            hipLaunchKernelGGL(k_axpby<R>, grid, dim3(256), 0, stream_, (long)SB, y, (R)a, x,
                               (R)b, z);
        };
        cx<R> *X = tw_fields_[0].p, *kb = tw_fields_[1].p, *kacc = tw_fields_[2].p;
        cx<R> *U[2] = {tw_fields_[3].p, tw_fields_[4].p};
        cx<R> *acc[2] = {tw_fields_[5].p, tw_fields_[6].p};
        cx<R> *slope[2] = {tw_fields_[7].p, slope1_.p}; // this (then the last) stage's slopes
        const double stage_c[4] = {0.0, dt / 2, dt / 2, dt}, weight[4] = {1, 2, 2, 1};
        for (int i = 0; i < 4; ++i) {
            // stage state X = w + c k_{i-1} and tangents U_j = T_j + c t_{i-1,j}
            const cx<R> *Xs = w0_.p;
            const cx<R> *Us[2] = {tg_[0].p, tg_[1].p};
            if (i > 0) {
                axpby(X, 1.0, w0_.p, stage_c[i], kb);
                Xs = X;
                for (int j = 0; j < 2; ++j) {
                    axpby(U[j], 1.0, tg_[j].p, stage_c[i], slope[j]);
                    Us[j] = U[j];
                }
            }
            // drift of the stage into kb
            jacobian(Xs, Xs, calA_.p, nlA_.p);
            StageArgs<R> c = stage_args(dt, nullptr, 0);
            c.nl = nlA_.p;
            c.w1 = const_cast<cx<R> *>(Xs);
            c.out = kb;
            run_col(9, c);
            if (i == 0)
                axpby(kacc, 1.0, kb, 0.0, nullptr);
            else
                axpby(kacc, 1.0, kacc, weight[i], kb);
            for (int j = 0; j < 2; ++j) {
                linearized_into(Xs, Us[j], slope[j]);
                hipLaunchKernelGGL(k_param_source<R>, grid, dim3(256), 0, stream_, (int)B_, K_,
                                   slope[j], Xs, (const double *)d_nu_.p,
                                   (const double *)d_alpha_.p, j);
                if (i == 0)
                    axpby(acc[j], 1.0, slope[j], 0.0, nullptr);
                else
                    axpby(acc[j], 1.0, acc[j], weight[i], slope[j]);
            }
        }
        axpby(w0_.p, 1.0, w0_.p, dt / 6, kacc);
        for (int j = 0; j < 2; ++j)
            axpby(tg_[j].p, 1.0, tg_[j].p, dt / 6, acc[j]);
        mix_valid_ = false;
        ++step_;
    }

    // Drift of the states S into out (retained half planes, B members).
    void drift_into(const cx<R> *S, cx<R> *out) {
        jacobian(S, S, calA_.p, nlA_.p); // J(S, S)
        StageArgs<R> a = stage_args(1.0, nullptr, 0);
        a.nl = nlA_.p;
        a.w1 = const_cast<cx<R> *>(S);
        a.out = out;
        run_col(9, a);
    }

    // Adjoint drift at states S applied to cotangents W, into out.
    void adjoint_into(const cx<R> *S, const cx<R> *W, cx<R> *out) {
        if (!calB_.p || !nlB_.p || !nlC_.p || !out)
            throw std::logic_error("adjoint pass on unallocated buffers");
        jacobian(S, W, calA_.p, nlA_.p); // J(S, W); calA holds W~
        StageArgs<R> m0 = stage_args(1.0, nullptr, 0);
        m0.w0 = const_cast<cx<R> *>(S);
        m0.mix = calB_.p; // mixed fields of the state: w~, (ky w)~
        run_col(0, m0);
        ProductArgs<R> pa{(int)B_, calA_.p, calB_.p, nlB_.p, nlC_.p, tw_.p};
        dispatch_products<R>(n_, pa, stream_);
        StageArgs<R> a = stage_args(1.0, nullptr, 0);
        a.nl = nlA_.p;
        a.nl2 = nlB_.p;
        a.nl3 = nlC_.p;
        a.w1 = const_cast<cx<R> *>(W);
        a.out = out;
        run_col(10, a);
    }

    // J(a, b) = u(a) . grad(b) of every member into the nonlinear mixed field
    // nl: two-state column pass (psi of a, w of b) into mix, then a row pass.
    void jacobian(const cx<R> *a_state, const cx<R> *b_state, cx<R> *mix, cx<R> *nl) {
        if (!a_state || !b_state || !mix || !nl)
            throw std::logic_error("Jacobian pass on unallocated buffers");
        StageArgs<R> c = stage_args(1.0, nullptr, 0);
        c.w0 = const_cast<cx<R> *>(a_state);
        c.w1 = const_cast<cx<R> *>(b_state);
        c.mix = mix;
        run_col(8, c);
        RowArgs<R> r{(int)B_, mix, nl, tw_row_, nullptr};
        run_row(r);
    }

    // members reference Fields (n*n interleaved doubles each) -> retained
    // half planes on the device, dropping masked modes as set_state does.
    void upload_fields(const double *fields, cx<R> *dst) {
        const size_t nn = (size_t)n_ * n_;
        for (size_t i = 0; i < 2 * nn * B_; ++i)
            if (!std::isfinite(fields[i]))
                throw std::invalid_argument("nonfinite GPU field");
        std::vector<cx<R>> half(spec_ * B_);
        for (size_t k = 0; k < B_; ++k)
            full_to_half<R>(n_, fields + 2 * nn * k, half.data() + spec_ * k);
        upload(dst, half.data(), half.size());
    }

    void download_fields(const cx<R> *src, double *fields) {
        check(hipGetLastError(), "drift kernels");
        std::vector<cx<R>> half(spec_ * B_);
        check(hipMemcpyAsync(half.data(), src, half.size() * sizeof(cx<R>),
                             hipMemcpyDeviceToHost, stream_),
              "drift download");
        check(hipStreamSynchronize(stream_), "drift download");
        const size_t nn = (size_t)n_ * n_;
        for (size_t k = 0; k < B_; ++k)
            half_to_full<R>(n_, half.data() + spec_ * k, fields + 2 * nn * k);
    }

    // Mixed fields of the current states (pass A), once after any change.
    void ensure_mixed(double dt) {
        if (!mix_valid_) {
            run_col(0, stage_args(dt, nullptr, 0));
            mix_valid_ = true;
        }
    }

    // The launches of one RK4 step: shell noise, then 4 x (row, col).  The
    // noise uses step index step_off, or *step_base + step_off when
    // step_base is given (captured sequences).  With `events` (9 entries) the
    // launches are bracketed by events so their durations can be read.
    void launch_step(double dt, const cx<R> *ext, long long ext_stride,
                     const uint64_t *step_base, uint64_t step_off, hipEvent_t *events) {
        if (!ext) {
            if (ou_) {
                launch_ou(1, dt / 2, noise_.p);
            } else {
                NoiseArgs<R> na{(int)B_, npairs_, d_sigma_.p, d_keys_.p, dt, step_off, noise_.p,
                                step_base};
                const long count = (long)B_ * npairs_;
                hipLaunchKernelGGL(k_noise<R>, dim3((unsigned)((count + 255) / 256)), dim3(256),
                                   0, stream_, na);
            }
        }
        RowArgs<R> r{(int)B_, mix_.p, nl_.p, tw_row_, nullptr};
        if (events)
            check(hipEventRecord(events[0], stream_), "event");
        for (int stage = 1; stage <= 4; ++stage) {
            r.cfl = stage == 1 ? cfl_.p : nullptr;
            run_row(r);
            if (events)
                check(hipEventRecord(events[2 * stage - 1], stream_), "event");
            run_col(stage, stage_args(dt, ext, ext_stride));
            if (events)
                check(hipEventRecord(events[2 * stage], stream_), "event");
        }
        if (!ext && ou_)
            launch_ou(1, dt / 2, nullptr);
        check(hipGetLastError(), "RK4 kernels");
    }

    // One RK4 step launched eagerly.
    void advance(double dt, const cx<R> *ext, long long ext_stride,
                 hipEvent_t *events = nullptr) {
        flush_parameters();
        if (!mix_valid_) {
            run_col(0, stage_args(dt, ext, ext_stride));
            mix_valid_ = true;
        }
        launch_step(dt, ext, ext_stride, nullptr, step_, events);
        ++step_;
        last_dt_ = dt;
        max_dt_ = std::max(max_dt_, dt);
    }

    // The filter launches of one observation group (observations staged):
    // log-weights, normalisation and resampling decision, optional weighted
    // mean, resampling copies (no-ops on the device without resampling).
    // `device_cycle`: group and last-group flag from ctr_ (captured cycles).
    void launch_filter(size_t nobs, bool device_cycle, bool is_last, bool with_mean) {
        PfArgs<R> a{(int)B_,   n_,        K_,           mix_.p,       (int)nobs,
                    obs_x_,    obs_y_,    obs_v_,       obs_sd_,      pf_cs_.p,
                    pf_sn_.p,  pf_wstate_.p, pf_wnorm_.p, pf_logw_.p, nullptr};
        hipLaunchKernelGGL(k_pf_loglik<R>, dim3((unsigned)B_), dim3(256), 0, stream_, a);
        const int P = (int)(B_ / pf_filters_);
        PfControl c{P,           pf_logw_.p, pf_wstate_.p,      pf_wnorm_.p,
                    pf_stats_.p, pf_anc_.p,  pf_uniforms_.p,    pf_max_uniforms_,
                    d_keys_.p,   pf_seeds_.p, pf_group_,        is_last ? 1 : 0,
                    device_cycle ? ctr_.p + 1 : nullptr};
        hipLaunchKernelGGL(k_pf_normalize, dim3((unsigned)pf_filters_), dim3(256), 0, stream_, c);
        if (with_mean) {
            const long blocks = ((long)spec_ + 255) / 256;
            const int per = pf_mean_chunk((long)B_);
            const int chunks = (int)((B_ + per - 1) / per);
            hipLaunchKernelGGL(k_pf_mean_partial<R>, dim3((unsigned)blocks, (unsigned)chunks),
                               dim3(256), 0, stream_, (int)B_, (long)spec_, per,
                               (const cx<R> *)w0_.p, (const double *)pf_wnorm_.p,
                               pf_mean_part_.p);
            hipLaunchKernelGGL(k_pf_mean_finish, dim3((unsigned)blocks), dim3(256), 0, stream_,
                               (long)spec_, chunks, (const double *)pf_mean_part_.p, pf_mean_.p);
        }
        const dim3 mgrid((unsigned)B_, (unsigned)((spec_ + 255) / 256));
        hipLaunchKernelGGL(k_pf_gather<R>, mgrid, dim3(256), 0, stream_, (int)B_, (long)spec_,
                           (const cx<R> *)w0_.p, w1_.p, (const int *)pf_anc_.p,
                           (const double *)pf_stats_.p, P);
        hipLaunchKernelGGL(k_pf_commit<R>, mgrid, dim3(256), 0, stream_, (int)B_, (long)spec_,
                           (const cx<R> *)w1_.p, w0_.p, (const double *)pf_stats_.p, P);
        if (ou_) {
            // children keep their ancestor's OU coefficients and restart their
            // stream from the key k_pf_normalize gave them (Solver copy +
            // reseed_future, src/filter.cpp:112-115)
            const dim3 fgrid((unsigned)B_, (unsigned)((npairs_ + 255) / 256));
            hipLaunchKernelGGL(k_pf_gather<double>, fgrid, dim3(256), 0, stream_, (int)B_,
                               (long)npairs_, (const cx<double> *)ou_F_.p, ou_tmp_.p,
                               (const int *)pf_anc_.p, (const double *)pf_stats_.p, P);
            hipLaunchKernelGGL(k_pf_commit<double>, fgrid, dim3(256), 0, stream_, (int)B_,
                               (long)npairs_, (const cx<double> *)ou_tmp_.p, ou_F_.p,
                               (const double *)pf_stats_.p, P);
            launch_ou(2, 0.0, nullptr, pf_stats_.p, P);
        }
    }

    // One call of the OU kernel over every member (modes of OuArgs).
    void launch_ou(int mode, double h, cx<R> *noise, const double *stats = nullptr, int P = 1) {
        OuArgs<R> a{(int)B_,   npairs_,  ou_mt_.p, ou_F_.p, ou_map_.p, d_sigma_.p,
                    d_keys_.p, d_tau_.p, h,        mode,    noise,     stats,
                    P};
        hipLaunchKernelGGL(k_ou<R>, dim3((unsigned)B_), dim3(OU_LANES), 0, stream_, a);
    }

    void set_forcing_process(int process) override {
        if (process != HZL_FORCING_WHITE && process != HZL_FORCING_OU)
            throw std::invalid_argument("forcing process must be HZL_FORCING_WHITE or HZL_FORCING_OU");
        if (process == HZL_FORCING_OU && (!(p_.forcing_tau > 0) || !std::isfinite(p_.forcing_tau)))
            throw std::invalid_argument("OU forcing needs a finite forcing_tau > 0");
        check(hipStreamSynchronize(stream_), "forcing process switch");
        clear_graphs(); // captured steps hold the previous process's kernels
        ou_ = process == HZL_FORCING_OU;
        if (!ou_)
            return;
        if (!ou_mt_.count) {
            // forcing_pairs (src/solver.cpp:67-83): first mode i = y n + x of
            // each conjugate pair, i < partner, inside the mask and the shell
            const int n = n_, w = 2 * p_.forcing_max + 1;
            const auto table = pair_table(p_.forcing_min, p_.forcing_max);
            std::vector<int> map, off;
            ou_first_.clear();
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x) {
                    const int a = x < n / 2 ? x : x - n, b = y < n / 2 ? y : y - n;
                    const long i = (long)y * n + x,
                               partner = (long)((n - y) % n) * n + (n - x) % n;
                    const int k2 = a * a + b * b;
                    if (3 * std::abs(a) < n && 3 * std::abs(b) < n && i < partner &&
                        k2 >= p_.forcing_min * p_.forcing_min &&
                        k2 <= p_.forcing_max * p_.forcing_max) {
                        const bool flip = a < 0;
                        const int c = flip ? -a : a, kb = flip ? -b : b;
                        const int p = table[(size_t)(c * w + kb + p_.forcing_max)];
                        if (p < 0)
                            throw std::logic_error("OU pair outside the engine's shell table");
                        map.push_back(flip ? ~p : p);
                        const int o = c * J_ + kb + K_;
                        off.push_back(flip ? ~o : o);
                        ou_first_.push_back(i);
                    }
                }
            if ((int)map.size() != npairs_)
                throw std::logic_error("OU pair count differs from the engine's shell");
            ou_map_ = DeviceBuffer<int>(map.size());
            upload(ou_map_.p, map.data(), map.size());
            ou_off_ = DeviceBuffer<int>(off.size());
            upload(ou_off_.p, off.data(), off.size());
            ou_mt_ = DeviceBuffer<uint64_t>((size_t)(MT_N + 1) * B_);
            ou_F_ = DeviceBuffer<cx<double>>(B_ * (size_t)npairs_);
            ou_tmp_ = DeviceBuffer<cx<double>>(B_ * (size_t)npairs_);
        }
        flush_parameters();
        launch_ou(0, 0.0, nullptr);
        check(hipGetLastError(), "OU start");
    }

    void reseed_future(size_t m, uint64_t seed) override {
        member_check(m);
        keys_[m] = seed;
        upload(d_keys_.p + m, &keys_[m], 1);
        if (!ou_)
            return;
        // mt19937_64 seeding ([rand.eng.mers]), position at the end of the
        // block so the first draw regenerates it
        std::vector<uint64_t> s(MT_N + 1);
        s[0] = seed;
        for (int i = 1; i < MT_N; ++i)
            s[i] = 6364136223846793005ULL * (s[i - 1] ^ (s[i - 1] >> 62)) + (uint64_t)i;
        s[MT_N] = MT_N;
        upload(ou_mt_.p + m * (MT_N + 1), s.data(), s.size());
    }

    void set_forcing_state(size_t m, const uint64_t *engine, const double *field) override {
        if (!ou_)
            throw std::invalid_argument("the OU forcing process is not selected");
        member_check(m);
        if (engine[MT_N] > (uint64_t)MT_N || engine[MT_N] % 2 != 0)
            throw std::invalid_argument("stream position must be even and at most 312");
        std::vector<cx<double>> F((size_t)npairs_);
        for (int j = 0; j < npairs_; ++j) {
            const long i = ou_first_[j];
            F[j] = {field[2 * i], field[2 * i + 1]};
            if (!std::isfinite(F[j].x) || !std::isfinite(F[j].y))
                throw std::invalid_argument("nonfinite forcing");
        }
        upload(ou_F_.p + m * npairs_, F.data(), F.size());
        upload(ou_mt_.p + m * (MT_N + 1), engine, MT_N + 1);
    }

    void get_forcing_state(size_t m, uint64_t *engine, double *field) override {
        get_forcing(m, field); // checks the process and the member
        check(hipMemcpyAsync(engine, ou_mt_.p + m * (MT_N + 1), (MT_N + 1) * sizeof(uint64_t),
                             hipMemcpyDeviceToHost, stream_),
              "stream download");
        check(hipStreamSynchronize(stream_), "stream download");
    }

    int forcing_process() const override {
        return ou_ ? HZL_FORCING_OU : HZL_FORCING_WHITE;
    }

    void get_forcing(size_t m, double *field) override {
        if (!ou_)
            throw std::invalid_argument("the OU forcing process is not selected");
        const size_t first = m == (size_t)-1 ? 0 : m, count = m == (size_t)-1 ? B_ : 1;
        if (m != (size_t)-1)
            member_check(m);
        std::vector<cx<double>> F((size_t)npairs_ * count);
        check(hipMemcpyAsync(F.data(), ou_F_.p + first * npairs_, F.size() * sizeof(cx<double>),
                             hipMemcpyDeviceToHost, stream_),
              "forcing download");
        check(hipStreamSynchronize(stream_), "forcing download");
        const size_t nn = (size_t)n_ * n_;
        for (size_t k = 0; k < count; ++k) {
            double *f = field + 2 * nn * k;
            std::fill(f, f + 2 * nn, 0.0);
            for (int j = 0; j < npairs_; ++j) {
                const long i = ou_first_[j], y = i / n_, x = i % n_;
                const long partner = ((n_ - y) % n_) * n_ + (n_ - x) % n_;
                const cx<double> v = F[k * npairs_ + j];
                f[2 * i] = v.x;
                f[2 * i + 1] = v.y;
                f[2 * partner] = v.x;
                f[2 * partner + 1] = -v.y;
            }
        }
    }

    // ------------------------------------------------ HIP graphs ----
    // A launch sequence is captured on the second consecutive call with the
    // same key (the first runs eagerly, so every lazy initialisation happens
    // outside capture) and replayed afterwards.  Every device address a
    // sequence holds is fixed for the engine's lifetime, except the filter
    // and observation buffers, whose reallocation clears the cache.
    struct GraphKey {
        int kind; // 1: RK4 steps, 2: filter cycle
        double dt;
        size_t steps, nobs, filters;
        const void *obs;
        bool mean;
        bool operator==(const GraphKey &o) const {
            return kind == o.kind && dt == o.dt && steps == o.steps && nobs == o.nobs &&
                   filters == o.filters && obs == o.obs && mean == o.mean;
        }
    };
    struct CachedGraph {
        GraphKey key;
        hipGraphExec_t exec;
        uint64_t used;
    };

    void set_counters(uint64_t group, uint64_t is_last) {
        hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, stream_, ctr_.p, step_, group,
                           is_last);
    }

    template <class F> void run_captured(const GraphKey &key, F &&launch) {
        ++graph_clock_;
        for (auto &g : graphs_)
            if (g.key == key) {
                g.used = graph_clock_;
                check(hipGraphLaunch(g.exec, stream_), "graph launch");
                return;
            }
        if (!(warm_ && warm_key_ == key)) {
            launch();
            warm_ = true;
            warm_key_ = key;
            return;
        }
        if (graphs_.size() >= kGraphCache) { // evict the least recently used
            size_t lru = 0;
            for (size_t i = 1; i < graphs_.size(); ++i)
                if (graphs_[i].used < graphs_[lru].used)
                    lru = i;
            check(hipStreamSynchronize(stream_), "graph eviction");
            (void)hipGraphExecDestroy(graphs_[lru].exec);
            graphs_.erase(graphs_.begin() + (long)lru);
        }
        hipGraph_t g = nullptr;
        check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "capture");
        try {
            launch();
        } catch (...) {
            (void)hipStreamEndCapture(stream_, &g);
            if (g)
                (void)hipGraphDestroy(g);
            throw;
        }
        check(hipStreamEndCapture(stream_, &g), "capture end");
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        check(e, "graph instantiate");
        graphs_.push_back(CachedGraph{key, exec, graph_clock_});
        warm_ = false;
        check(hipGraphLaunch(exec, stream_), "graph launch");
    }

    void clear_graphs() {
        if (!graphs_.empty())
            (void)hipStreamSynchronize(stream_);
        for (auto &g : graphs_)
            (void)hipGraphExecDestroy(g.exec);
        graphs_.clear();
        warm_ = false;
        ++launch_epoch_; // sequences other modules captured are stale too
    }

    int n_;
    size_t B_;
    int K_, KP_, J_;
    hzl_parameters p_;
    size_t spec_ = 0, mixf_ = 0;
    hipStream_t stream_ = nullptr;
    DeviceBuffer<cx<R>> w0_, w1_, acc_, mix_, nl_, ext_, tw_, noise_;
    // particle filter state
    bool pf_on_ = false;
    uint64_t pf_group_ = 0;
    size_t pf_filters_ = 1;
    int pf_max_uniforms_ = 0;
    DeviceBuffer<uint64_t> pf_seeds_;
    size_t obs_cap_ = 0;
    double *obs_host_ = nullptr; // pinned staging of one observation group
    double *pf_host_ = nullptr;  // pinned staging of the filter statistics and mean
    DeviceBuffer<double> obs_dev_;
    hipEvent_t obs_event_ = nullptr;
    const double *obs_v_ = nullptr, *obs_sd_ = nullptr;
    const int *obs_x_ = nullptr, *obs_y_ = nullptr;
    DeviceBuffer<double> pf_wstate_, pf_wnorm_, pf_logw_, pf_stats_, pf_mean_, pf_mean_part_,
        pf_uniforms_, pf_cs_, pf_sn_;
    DeviceBuffer<int> pf_anc_;
    int npairs_ = 0;
    DeviceBuffer<double> cfl_, cfl_diag_;
    DeviceBuffer<int> pairs_;
    DeviceBuffer<double> d_nu_, d_alpha_, d_sigma_, d_tau_, d_amp_;
    DeviceBuffer<uint64_t> d_keys_;
    std::vector<double> nu_, alpha_, sigma_, tau_, amp_;
    // staging for key uploads (creation, pf_begin with new keys, reseed_future);
    // resampling reseeds children on the device only, so d_keys_ is authoritative
    std::vector<uint64_t> keys_;
    // OU forcing process (set_forcing_process): stream words [B][MT_N + 1],
    // coefficients [B][P] in the reference's pair order, their resampling
    // copy, the engine pair index of each (~index when conjugated) and the
    // grid index of each pair's first mode.
    bool ou_ = false;
    DeviceBuffer<uint64_t> ou_mt_;
    DeviceBuffer<cx<double>> ou_F_, ou_tmp_;
    DeviceBuffer<int> ou_map_, ou_off_;
    std::vector<long> ou_first_;
    bool params_dirty_ = false, mix_valid_ = false;
    uint64_t step_ = 0;
    double last_dt_ = 0, max_dt_ = 0, last_cfl_rate_ = 0;
    // tangent state: sensitivities, stage work fields, the second slope field
    DeviceBuffer<cx<R>> tg_[2], tw_fields_[8], slope1_;
    // calibration drifts: input fields, result, mixed and nonlinear fields
    DeviceBuffer<cx<R>> caux_, cout_, calA_, calB_, nlA_, nlB_, nlC_;
    // launch shapes (variants.hpp) and their twiddle tables
    int col_var_ = 0, row_var_ = 0;
    const cx<R> *tw_col_ = nullptr, *tw_row_ = nullptr;
    std::vector<std::pair<int, DeviceBuffer<cx<R>>>> tw_extra_;
    // captured launch sequences
    static constexpr size_t kGraphSteps = 200, kGraphCache = 4;
    DeviceBuffer<uint64_t> ctr_;
    std::vector<CachedGraph> graphs_;
    GraphKey warm_key_{};
    bool warm_ = false, use_graph_ = graphs_by_default();
    int device_ = 0; // the device current at creation (owns every buffer)
    uint64_t launch_epoch_ = 0;
    uint64_t graph_clock_ = 0;
};

// Model-parameter rules of Parameters::validate (src/solver.cpp:31-40): each
// row is a condition that must hold and the reference's message when it does
// not, checked in order.
void validate_model(size_t n, const hzl_parameters &p) {
    const auto nonnegative = [](double v) { return std::isfinite(v) && v >= 0; };
    const long long shell_top = 3LL * p.forcing_max;
    const struct {
        bool holds;
        const char *message;
    } rules[] = {
        {nonnegative(p.viscosity) && nonnegative(p.friction) && nonnegative(p.forcing_sigma) &&
             std::isfinite(p.forcing_tau) && p.forcing_tau > 0,
         "nu, alpha, sigma must be finite and nonnegative; tau must be positive"},
        {p.forcing_min >= 1 && p.forcing_max >= p.forcing_min && shell_top < (long long)n,
         "forcing shell must lie strictly inside the dealiased grid"},
    };
    for (const auto &r : rules)
        if (!r.holds)
            throw std::invalid_argument(r.message);
}

// Everything an ensemble needs: a supported grid and precision, at least one
// member, the model rules, and a Kolmogorov term inside the dealiased grid.
void validate(size_t n, size_t members, int precision, const hzl_parameters &p) {
    (void)elems_per_thread((int)n);
    if (members == 0)
        throw std::invalid_argument("ensemble needs at least one member");
    if (precision != HZL_FP32 && precision != HZL_FP64)
        throw std::invalid_argument("precision must be HZL_FP32 or HZL_FP64");
    validate_model(n, p);
    if (!std::isfinite(p.kolmogorov_amplitude) || p.kolmogorov_wavenumber < 0 ||
        3LL * p.kolmogorov_wavenumber >= (long long)n)
        throw std::invalid_argument("Kolmogorov forcing must lie inside the dealiased grid");
}

} // namespace hzl_amd

// ---------------------------------------------------------------- C-ABI --
struct hzl_ensemble {
    std::unique_ptr<hzl_amd::EngineBase> engine;
};

namespace hzl_amd {
std::string &error_slot() {
    thread_local std::string message;
    return message;
}
} // namespace hzl_amd
using hzl_amd::guarded;

namespace hzl_amd {
EngineBase &engine_of(hzl_ensemble_t h) {
    if (!h || !h->engine)
        throw std::invalid_argument("null ensemble handle");
    return *h->engine;
}
} // namespace hzl_amd
using hzl_amd::engine_of;

extern "C" {

const char *hzl_last_error(void) {
    return hzl_amd::error_slot().c_str();
}

int hzl_ensemble_create(hzl_ensemble_t *out, size_t n, size_t members, int precision,
                        const hzl_parameters *params, uint64_t seed) {
    return guarded([&] {
        if (!out || !params)
            throw std::invalid_argument("null argument");
        *out = nullptr;
        hzl_amd::validate(n, members, precision, *params);
        auto h = std::make_unique<hzl_ensemble>();
        if (precision == HZL_FP32)
            h->engine = std::make_unique<hzl_amd::Engine<float>>((int)n, members, *params, seed);
        else
            h->engine =
                std::make_unique<hzl_amd::Engine<double>>((int)n, members, *params, seed);
        // large FP32 ensembles pick their launch shapes (remembered choice,
        // else timed now, on the zero initial state); HZL_AUTOTUNE=0 keeps
        // the defaults
        const char *env = std::getenv("HZL_AUTOTUNE");
        const bool autotune = !(env && std::strcmp(env, "0") == 0);
        if (autotune && (double)members * (double)n * (double)n >= (double)(1 << 22))
            h->engine->tune(true, nullptr, nullptr);
        *out = h.release();
    });
}

int hzl_ensemble_variant_count(hzl_ensemble_t h, int *col_variants, int *row_variants) {
    return guarded([&] {
        if (!col_variants || !row_variants)
            throw std::invalid_argument("null argument");
        *col_variants = engine_of(h).col_variants();
        *row_variants = engine_of(h).row_variants();
    });
}

int hzl_ensemble_set_tuning(hzl_ensemble_t h, int col_variant, int row_variant) {
    return guarded([&] { engine_of(h).set_tuning(col_variant, row_variant); });
}

int hzl_ensemble_get_tuning(hzl_ensemble_t h, int *col_variant, int *row_variant) {
    return guarded([&] {
        if (!col_variant || !row_variant)
            throw std::invalid_argument("null argument");
        engine_of(h).get_tuning(col_variant, row_variant);
    });
}

int hzl_ensemble_autotune(hzl_ensemble_t h, int use_cache, double *col_ms, double *row_ms) {
    return guarded([&] { engine_of(h).tune(use_cache != 0, col_ms, row_ms); });
}

int hzl_ensemble_destroy(hzl_ensemble_t h) {
    return guarded([&] { delete h; });
}

int hzl_parameters_validate(size_t n, const hzl_parameters *params) {
    return guarded([&] {
        if (!params)
            throw std::invalid_argument("null argument");
        hzl_amd::validate_model(n, *params);
    });
}

uint64_t hzl_stream_seed(uint64_t base, uint64_t index) {
    return hzl_amd::stream_seed(base, index);
}

int hzl_ensemble_set_member_tau(hzl_ensemble_t h, size_t member, double tau) {
    return guarded([&] { engine_of(h).set_member_tau(member, tau); });
}

int hzl_ensemble_set_member_parameters(hzl_ensemble_t h, size_t member, double viscosity,
                                       double friction, double forcing_sigma,
                                       double kolmogorov_amplitude) {
    return guarded([&] {
        engine_of(h).set_member_parameters(member, viscosity, friction, forcing_sigma,
                                           kolmogorov_amplitude);
    });
}

int hzl_ensemble_set_state(hzl_ensemble_t h, size_t member, const double *field) {
    return guarded([&] {
        if (!field)
            throw std::invalid_argument("null field");
        engine_of(h).set_state(member, field);
    });
}

int hzl_ensemble_get_state(hzl_ensemble_t h, size_t member, double *field) {
    return guarded([&] {
        if (!field)
            throw std::invalid_argument("null field");
        engine_of(h).get_state(member, field);
    });
}

int hzl_ensemble_init_grf(hzl_ensemble_t h) {
    return guarded([&] { engine_of(h).init_grf(); });
}

int hzl_ensemble_step(hzl_ensemble_t h, double dt, size_t steps) {
    return guarded([&] { engine_of(h).step(dt, steps); });
}

int hzl_ensemble_step_integrator(hzl_ensemble_t h, double dt, size_t steps, int integrator) {
    return guarded([&] { engine_of(h).step_integrator(dt, steps, integrator); });
}

int hzl_ensemble_set_forcing_process(hzl_ensemble_t h, int process) {
    return guarded([&] { engine_of(h).set_forcing_process(process); });
}

int hzl_ensemble_set_forcing_state(hzl_ensemble_t h, size_t m, const uint64_t *engine,
                                   const double *field) {
    return guarded([&] {
        if (!engine || !field)
            throw std::invalid_argument("null stream or field");
        engine_of(h).set_forcing_state(m, engine, field);
    });
}

int hzl_ensemble_get_forcing_state(hzl_ensemble_t h, size_t m, uint64_t *engine,
                                   double *field) {
    return guarded([&] {
        if (!engine || !field)
            throw std::invalid_argument("null stream or field");
        engine_of(h).get_forcing_state(m, engine, field);
    });
}

int hzl_ensemble_reseed_future(hzl_ensemble_t h, size_t m, uint64_t seed) {
    return guarded([&] { engine_of(h).reseed_future(m, seed); });
}

int hzl_ensemble_get_forcing(hzl_ensemble_t h, size_t m, double *field) {
    return guarded([&] {
        if (!field)
            throw std::invalid_argument("null field");
        engine_of(h).get_forcing(m, field);
    });
}

int hzl_ensemble_drift(hzl_ensemble_t h, double *out) {
    return guarded([&] {
        if (!out)
            throw std::invalid_argument("null argument");
        engine_of(h).drift(out);
    });
}

int hzl_ensemble_linearized_drift(hzl_ensemble_t h, const double *directions, double *out) {
    return guarded([&] {
        if (!directions || !out)
            throw std::invalid_argument("null argument");
        engine_of(h).linearized_drift(directions, out);
    });
}

int hzl_ensemble_adjoint_drift(hzl_ensemble_t h, const double *cotangents, double *out) {
    return guarded([&] {
        if (!cotangents || !out)
            throw std::invalid_argument("null argument");
        engine_of(h).adjoint_drift(cotangents, out);
    });
}

int hzl_ensemble_step_tangent(hzl_ensemble_t h, double dt, size_t steps,
                              const double *tangents_in, double *tangents_out) {
    return guarded([&] {
        if (!tangents_in || !tangents_out)
            throw std::invalid_argument("null argument");
        engine_of(h).step_tangent(dt, steps, tangents_in, tangents_out);
    });
}

int hzl_likelihood_gradient(hzl_ensemble_t h, const double *initial_field, double dt,
                            size_t nobs, const size_t *steps, const int *x, const int *y,
                            const double *value, const double *sd, const double *log_params,
                            double *objective, double *gradient) {
    return guarded([&] {
        if (!initial_field || !steps || !x || !y || !value || !sd || !log_params || !objective ||
            !gradient)
            throw std::invalid_argument("null argument");
        engine_of(h).likelihood_gradient(initial_field, dt, nobs, steps, x, y, value, sd,
                                         log_params, objective, gradient);
    });
}

int hzl_likelihood_adjoint(hzl_ensemble_t h, const double *initial_field, double dt,
                           size_t nobs, const size_t *steps, const int *x, const int *y,
                           const double *value, const double *sd, const double *log_params,
                           size_t stride, double *objective, double *gradient, size_t *counts) {
    return guarded([&] {
        if (!initial_field || !steps || !x || !y || !value || !sd || !log_params || !objective ||
            !gradient)
            throw std::invalid_argument("null argument");
        engine_of(h).likelihood_adjoint(initial_field, dt, nobs, steps, x, y, value, sd,
                                        log_params, stride, objective, gradient, counts);
    });
}

int hzl_ensemble_set_graph(hzl_ensemble_t h, int enabled) {
    return guarded([&] { engine_of(h).set_graph(enabled != 0); });
}

int hzl_ensemble_step_forced(hzl_ensemble_t h, double dt, const double *forcing, int shared) {
    return guarded([&] {
        if (!forcing)
            throw std::invalid_argument("null forcing");
        engine_of(h).step_forced(dt, forcing, shared != 0);
    });
}

int hzl_ensemble_synchronize(hzl_ensemble_t h) {
    return guarded([&] { engine_of(h).synchronize(); });
}

int hzl_ensemble_diagnostics(hzl_ensemble_t h, double *out) {
    return guarded([&] {
        if (!out)
            throw std::invalid_argument("null output");
        engine_of(h).diagnostics(out);
    });
}

}  // extern "C"

// ------------------------------------------- allocation-free diagnostics --
// Staging of hzl_ensemble_diagnostics_noalloc: 6 doubles per member for at
// most kNoAllocMembers members, one copy of the array on every GPU the library
// runs on, held until the process ends.  Host code never names the array in a
// launch or copy: each device's own address comes from hipGetSymbolAddress.
constexpr size_t kNoAllocMembers = 65536;
static __device__ __align__(256) double scratch[6 * kNoAllocMembers];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double *scratch_address[kMaxDevices] = {};

// The device address of `scratch` on `device`, looked up once per device with
// that device current (the caller's current device is restored).
double *scratch_on(int device) {
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (device < 0 || device >= kMaxDevices)
        throw std::runtime_error("device index out of range for the diagnostics scratch");
    if (!scratch_address[device]) {
        int current = 0;
        hzl_amd::check(hipGetDevice(&current), "hipGetDevice");
        if (current != device)
            hzl_amd::check(hipSetDevice(device), "hipSetDevice");
        void *ptr = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (current != device)
            hzl_amd::check(hipSetDevice(current), "hipSetDevice");
        hzl_amd::check(rc, "hipGetSymbolAddress(scratch)");
        scratch_address[device] = static_cast<double *>(ptr);
    }
    return scratch_address[device];
}
} // namespace

extern "C" {

int hzl_ensemble_diagnostics_noalloc(hzl_ensemble_t h, double *out) {
    return guarded([&] {
        if (!out)
            throw std::invalid_argument("null argument");
        hzl_amd::EngineBase &e = engine_of(h);
        if (e.members() > kNoAllocMembers)
            throw std::invalid_argument(
                "diagnostics without allocation take at most 65536 members");
        e.diagnostics_with(out, scratch_on(e.device()));
    });
}

int hzl_ensemble_energy_spectrum(hzl_ensemble_t h, double *out) {
    return guarded([&] {
        if (!out)
            throw std::invalid_argument("null argument");
        engine_of(h).energy_spectrum(out);
    });
}

int hzl_ensemble_profile_step(hzl_ensemble_t h, double dt, float *kernel_ms) {
    return guarded([&] {
        if (!kernel_ms)
            throw std::invalid_argument("null output");
        engine_of(h).profile_step(dt, kernel_ms);
    });
}

int hzl_pf_begin(hzl_ensemble_t h, uint64_t seed, size_t max_groups) {
    return guarded([&] { engine_of(h).pf_begin(seed, max_groups); });
}

int hzl_pf_assimilate(hzl_ensemble_t h, double dt, size_t steps, size_t nobs, const int *x,
                      const int *y, const double *value, const double *sd, int is_last,
                      double *record, double *mean_field) {
    return guarded([&] {
        if (!x || !y || !value || !sd)
            throw std::invalid_argument("null observation array");
        engine_of(h).pf_assimilate(dt, steps, nobs, x, y, value, sd, is_last != 0, record,
                                   mean_field);
    });
}

int hzl_pf_result(hzl_ensemble_t h, double *log_likelihood, size_t *resampling_events) {
    return guarded([&] { engine_of(h).pf_result(log_likelihood, resampling_events); });
}

int hzl_pf_begin_filters(hzl_ensemble_t h, const uint64_t *seeds, size_t filters,
                         size_t max_groups) {
    return guarded([&] {
        if (!seeds)
            throw std::invalid_argument("null seeds");
        engine_of(h).pf_begin_filters(seeds, filters, max_groups, true);
    });
}

int hzl_ensemble_fill_state(hzl_ensemble_t h, const double *field) {
    return guarded([&] {
        if (!field)
            throw std::invalid_argument("null field");
        engine_of(h).fill_state(field);
    });
}

int hzl_ensemble_set_step_index(hzl_ensemble_t h, uint64_t step) {
    return guarded([&] { engine_of(h).set_step_index(step); });
}

int hzl_ensemble_step_index(hzl_ensemble_t h, uint64_t *out) {
    return guarded([&] {
        if (!out)
            throw std::invalid_argument("null output");
        *out = engine_of(h).step_index();
    });
}

int hzl_ensemble_device_state(hzl_ensemble_t h, void **ptr, size_t *member_stride_bytes) {
    return guarded([&] {
        if (!ptr || !member_stride_bytes)
            throw std::invalid_argument("null output");
        engine_of(h).device_state(ptr, member_stride_bytes);
    });
}

int hzl_ensemble_stream(hzl_ensemble_t h, void **stream) {
    return guarded([&] {
        if (!stream)
            throw std::invalid_argument("null output");
        *stream = engine_of(h).stream();
    });
}
}
