// pf_multi.hip -- one bootstrap particle filter whose particles live in
// several ensembles of different grid sizes (e.g. 256 particles at 256^2 and
// 32 at 512^2), restating reference bootstrap_filter (src/filter.cpp:38-126)
// over the union of their members.
//
// Each ensemble advances on its own HIP stream (the forecasts of different
// grids overlap on the GPU) and writes its members' FP64 log-weights into one
// joint array; a single workgroup normalises the union (log-sum-exp, ESS,
// systematic resampling below N/2 except on the last group, children
// reseeded stream_seed(seed, (g+1) N + i) with N the total particle count);
// the gather copies every child's retained modes from its ancestor, whatever
// the ancestor's grid: modes present in both retained sets are copied, modes
// only the child's grid retains are zero (spectral truncation / zero padding
// of the forward-normalised coefficients, which keeps the physical field's
// resolved part).  Observation grid indices are given on the coarsest grid;
// an ensemble r times finer reads the point (r x, r y).
//
// A cycle is one launch sequence over all the streams: from the filter's own
// stream it forks to every ensemble's stream (mixed fields, RK4 steps,
// log-weights) and joins back for the normalisation, resampling gather and key
// scatter.  With graph replay on it is captured as one HIP graph (fork / join
// through events) on its second identical call and replayed afterwards; the
// group index and last-group flag are read from device memory, and the host
// waits once per cycle, for the four record values.
#include "engine.hpp"
#include "gemm_launch.hpp"

#include <algorithm>
#include <limits>
#include <random>
#include <stdexcept>
#include <vector>

namespace hzl_amd {

constexpr int MULTI_MAX = 4;

template <typename R> struct MultiSources {
    int count;
    const cx<R> *w[MULTI_MAX];
    int K[MULTI_MAX];
    long first[MULTI_MAX]; // global index of the ensemble's member 0
    long members[MULTI_MAX];
};

__global__ void __launch_bounds__(256) k_pf_multi_normalize(PfControl a) {
    __shared__ double red[2 * 256];
    __shared__ double cw[kPfLocalWeights];
    pf_normalize_body(DevCtx{}, a, red, 256, 0, cw, kPfLocalWeights);
}

// dst member i of ensemble `self` (grid K) <- ancestor anc[first_self + i],
// when the group resampled (stats[3] != 0).
template <typename R>
__global__ void k_pf_multi_gather(MultiSources<R> src, int self, long B, int K,
                                  const int *anc, const double *stats, cx<R> *dst) {
    if (stats[3] == 0)
        return;
    const int J = 2 * K + 1;
    const long spec = (long)(K + 1) * J;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= B * spec)
        return;
    const long m = i / spec, e = i % spec;
    const int c = (int)(e / J), ky = (int)(e % J) - K;
    const long a = anc[src.first[self] + m];
    int s = 0;
    for (int t = 1; t < src.count; ++t)
        if (a >= src.first[t])
            s = t;
    const int Ks = src.K[s], Js = 2 * Ks + 1;
    cx<R> v = {0, 0};
    if (c <= Ks && ky >= -Ks && ky <= Ks)
        v = src.w[s][(a - src.first[s]) * (long)(Ks + 1) * Js + (long)c * Js + ky + Ks];
    dst[i] = v;
}

template <typename R>
__global__ void k_pf_multi_copy(const double *stats, const cx<R> *src, cx<R> *dst, long n) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (stats[3] != 0 && i < n)
        dst[i] = src[i];
}

// The group index and last-group flag of the coming normalisation.
__global__ void k_set_cycle(uint64_t *cycle, uint64_t group, uint64_t is_last) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cycle[0] = group;
        cycle[1] = is_last;
    }
}

__global__ void k_copy_keys(const uint64_t *src, uint64_t *dst, long n) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        dst[i] = src[i];
}

class MultiFilter {
  public:
    MultiFilter(EngineBase **engines, size_t count) {
        if (count < 1 || count > (size_t)MULTI_MAX)
            throw std::invalid_argument("a joint filter spans 1 to 4 ensembles");
        for (size_t e = 0; e < count; ++e) {
            if (!engines[e])
                throw std::invalid_argument("null ensemble handle");
            engs_.push_back(engines[e]);
        }
        const EngineViews v0 = engs_[0]->views();
        coarse_ = v0.n;
        for (EngineBase *e : engs_) {
            const EngineViews v = e->views();
            if (v.precision != v0.precision)
                throw std::invalid_argument("the ensembles of a joint filter share one precision");
            coarse_ = std::min(coarse_, v.n);
            first_.push_back(total_);
            total_ += (size_t)v.B;
        }
        for (EngineBase *e : engs_) {
            const int n = e->views().n;
            if (n % coarse_ != 0)
                throw std::invalid_argument("grids must be multiples of the coarsest grid");
        }
        if (total_ < 2)
            throw std::invalid_argument("particle filter needs at least two particles");
        // device addresses, sizes and streams are fixed for an ensemble's
        // lifetime: read once, so launch sequences touch no engine state
        for (EngineBase *e : engs_)
            views_.push_back(e->views());
        hcheck(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "hipStreamCreate");
        hcheck(hipEventCreateWithFlags(&done_, hipEventDisableTiming), "hipEventCreate");
        hcheck(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "hipEventCreate");
        for (size_t e = 0; e < engs_.size(); ++e) {
            hipEvent_t ev, ej;
            hcheck(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
            hcheck(hipEventCreateWithFlags(&ej, hipEventDisableTiming), "hipEventCreate");
            ready_.push_back(ev);
            join_.push_back(ej);
        }
        hcheck(hipHostMalloc(reinterpret_cast<void **>(&host_stats_), 8 * sizeof(double)),
               "hipHostMalloc");
    }
    ~MultiFilter() {
        if (st_)
            (void)hipStreamSynchronize(st_);
        if (exec_)
            (void)hipGraphExecDestroy(exec_);
        for (hipEvent_t e : ready_)
            (void)hipEventDestroy(e);
        for (hipEvent_t e : join_)
            (void)hipEventDestroy(e);
        if (done_)
            (void)hipEventDestroy(done_);
        if (fork_)
            (void)hipEventDestroy(fork_);
        if (host_stats_)
            (void)hipHostFree(host_stats_);
        if (st_)
            (void)hipStreamDestroy(st_);
    }
    MultiFilter(const MultiFilter &) = delete;
    MultiFilter &operator=(const MultiFilter &) = delete;

    // Particle i of the union starts on noise key stream_seed(seed, i).
    void begin(uint64_t seed, size_t max_groups) {
        // children copy states across grids with their own gather; the OU
        // forcing state (coefficients, streams) does not follow them
        for (EngineBase *e : engs_)
            if (e->forcing_process() == HZL_FORCING_OU)
                throw std::invalid_argument("a joint filter runs with the white forcing process");
        seed_ = seed;
        group_ = 0;
        const size_t P = total_;
        std::vector<double> w(P, 1.0 / (double)P), st(8, 0.0);
        up(wstate_, w.data(), P);
        up(stats_, st.data(), 8);
        const size_t U = std::max<size_t>(max_groups, 1);
        std::mt19937_64 rng(stream_seed(seed, std::numeric_limits<uint64_t>::max()));
        std::vector<double> u(U);
        for (auto &x : u)
            x = (double)(rng() >> 11) * 0x1.0p-53;
        up(uniforms_, u.data(), U);
        max_uniforms_ = (int)U;
        std::vector<uint64_t> keys(P);
        for (size_t i = 0; i < P; ++i)
            keys[i] = stream_seed(seed, i);
        up(keys_, keys.data(), P);
        up(seeds_, &seed_, 1);
        wnorm_.alloc(P);
        logw_.alloc(P);
        anc_.alloc(P);
        cycle_.alloc(2);
        size_t need = 0; // resampling scratch: every child's retained modes
        for (const EngineViews &v : views_)
            need += (size_t)v.B * (v.K + 1) * (2 * v.K + 1);
        scratch_.alloc(need * (views_[0].precision == 0 ? sizeof(cx<float>) : sizeof(cx<double>)) /
                           sizeof(cx<double>) + 1);
        scatter_keys();
        hcheck(hipStreamSynchronize(st_), "joint filter start");
        on_ = true;
    }

    void assimilate(double dt, size_t steps, size_t nobs, const int *x, const int *y,
                    const double *value, const double *sd, bool is_last, double *record) {
        if (!on_)
            throw std::invalid_argument("particle filter not started");
        if (nobs == 0)
            throw std::invalid_argument("observation group is empty");
        for (size_t o = 0; o < nobs; ++o)
            if (x[o] < 0 || y[o] < 0 || x[o] >= coarse_ || y[o] >= coarse_)
                throw std::invalid_argument("observations must lie on the coarsest grid");
        if (!is_last && group_ >= (uint64_t)max_uniforms_)
            throw std::invalid_argument(
                "more observation groups than the filter was started for (max_groups)");
        // eager preparation on every ensemble's stream: observations staged at
        // its resolution, time-step checks and the device step counter
        std::vector<int> xs(nobs), ys(nobs);
        for (size_t k = 0; k < engs_.size(); ++k) {
            const int r = views_[k].n / coarse_;
            for (size_t o = 0; o < nobs; ++o) {
                xs[o] = x[o] * r;
                ys[o] = y[o] * r;
            }
            engs_[k]->pf_stage_logweights(nobs, xs.data(), ys.data(), value, sd);
            if (steps)
                engs_[k]->prepare_steps(dt);
        }
        hipLaunchKernelGGL(k_set_cycle, dim3(1), dim3(64), 0, st_, cycle_.p, group_,
                           (uint64_t)(is_last ? 1 : 0));
        for (size_t k = 0; k < engs_.size(); ++k) {
            hcheck(hipEventRecord(ready_[k], views_[k].stream), "event");
            hcheck(hipStreamWaitEvent(st_, ready_[k], 0), "event wait");
        }
        Key key{dt, steps, nobs, alloc_generation(), {}};
        for (size_t k = 0; k < engs_.size(); ++k)
            key.epochs[k] = engs_[k]->launch_epoch();
        bool graphs = true;
        for (EngineBase *e : engs_)
            graphs = graphs && e->graphs_enabled();
        if (graphs && exec_ && exec_key_ == key) {
            hcheck(hipGraphLaunch(exec_, st_), "joint graph launch");
        } else if (graphs && warm_ && warm_key_ == key) {
            if (exec_) {
                (void)hipGraphExecDestroy(exec_);
                exec_ = nullptr;
            }
            hipGraph_t g = nullptr;
            hcheck(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal), "capture");
            try {
                launches(dt, steps, nobs);
            } catch (...) {
                (void)hipStreamEndCapture(st_, &g);
                if (g)
                    (void)hipGraphDestroy(g);
                throw;
            }
            hcheck(hipStreamEndCapture(st_, &g), "capture end");
            const hipError_t e = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hcheck(e, "joint graph instantiate");
            exec_key_ = key;
            warm_ = false;
            hcheck(hipGraphLaunch(exec_, st_), "joint graph launch");
        } else {
            launches(dt, steps, nobs);
            warm_ = graphs;
            warm_key_ = key;
            warm_key_.generation = alloc_generation(); // after this run's allocations
        }
        hcheck(hipGetLastError(), "joint filter kernels");
        // every ensemble's next work waits for the resampling copies
        hcheck(hipEventRecord(done_, st_), "event");
        for (size_t k = 0; k < engs_.size(); ++k) {
            hcheck(hipStreamWaitEvent(views_[k].stream, done_, 0), "event wait");
            if (steps)
                engs_[k]->commit_steps(dt, steps);
            engs_[k]->invalidate_mixed();
        }
        hcheck(hipMemcpyAsync(host_stats_, stats_.p, 8 * sizeof(double), hipMemcpyDeviceToHost,
                              st_),
               "stats");
        hcheck(hipStreamSynchronize(st_), "joint filter"); // the cycle's one wait
        if (record)
            for (int k = 0; k < 4; ++k)
                record[k] = host_stats_[k];
        ++group_;
    }

    void result(double *ll, size_t *events) {
        if (!on_)
            throw std::invalid_argument("particle filter not started");
        double st[8];
        hcheck(hipMemcpy(st, stats_.p, sizeof(st), hipMemcpyDeviceToHost), "result");
        if (ll)
            *ll = st[4];
        if (events)
            *events = (size_t)st[5];
    }

    size_t particles() const {
        return total_;
    }

  private:
    template <typename T> void up(Buf<T> &b, const T *h, size_t n) {
        b.alloc(n);
        hcheck(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, st_), "upload");
    }

    // One cycle's launches (capturable): fork to the ensembles' streams for
    // their mixed fields, steps and log-weights, join back for the
    // normalisation, the resampling gather and the key scatter.
    void launches(double dt, size_t steps, size_t nobs) {
        hcheck(hipEventRecord(fork_, st_), "event");
        for (size_t k = 0; k < engs_.size(); ++k) {
            EngineBase *e = engs_[k];
            const hipStream_t s = views_[k].stream;
            hcheck(hipStreamWaitEvent(s, fork_, 0), "event wait");
            e->enqueue_steps(dt, steps);
            e->pf_enqueue_logweights(nobs, wstate_.p + first_[k], logw_.p + first_[k]);
            hcheck(hipEventRecord(join_[k], s), "event");
            hcheck(hipStreamWaitEvent(st_, join_[k], 0), "event wait");
        }
        PfControl c{(int)total_, logw_.p, wstate_.p,      wnorm_.p,       stats_.p,
                    anc_.p,      uniforms_.p, max_uniforms_, keys_.p,       seeds_.p,
                    0,           0,           cycle_.p};
        hipLaunchKernelGGL(k_pf_multi_normalize, dim3(1), dim3(256), 0, st_, c);
        if (views_[0].precision == 0)
            gather<float>();
        else
            gather<double>();
        scatter_keys();
    }

    void scatter_keys() {
        for (size_t k = 0; k < engs_.size(); ++k) {
            const EngineViews &v = views_[k];
            hipLaunchKernelGGL(k_copy_keys, grid1d(v.B), dim3(256), 0, st_, keys_.p + first_[k],
                               v.keys, (long)v.B);
        }
    }

    template <typename R> void gather() {
        MultiSources<R> src{};
        src.count = (int)engs_.size();
        for (size_t k = 0; k < engs_.size(); ++k) {
            const EngineViews &v = views_[k];
            src.w[k] = reinterpret_cast<const cx<R> *>(v.w0);
            src.K[k] = v.K;
            src.first[k] = (long)first_[k];
            src.members[k] = v.B;
        }
        // children first into scratch (every source is read before any write)
        std::vector<long> counts(engs_.size());
        size_t need = 0;
        for (size_t k = 0; k < engs_.size(); ++k) {
            const EngineViews &v = views_[k];
            counts[k] = (long)v.B * (v.K + 1) * (2 * v.K + 1);
            need += (size_t)counts[k];
        }
        scratch_.alloc(need * sizeof(cx<R>) / sizeof(cx<double>) + 1);
        cx<R> *base = reinterpret_cast<cx<R> *>(scratch_.p);
        long off = 0;
        for (size_t k = 0; k < engs_.size(); ++k) {
            const EngineViews &v = views_[k];
            hipLaunchKernelGGL(k_pf_multi_gather<R>, grid1d(counts[k]), dim3(256), 0, st_, src,
                               (int)k, (long)v.B, v.K, anc_.p, stats_.p, base + off);
            off += counts[k];
        }
        off = 0;
        for (size_t k = 0; k < engs_.size(); ++k) {
            const EngineViews &v = views_[k];
            hipLaunchKernelGGL(k_pf_multi_copy<R>, grid1d(counts[k]), dim3(256), 0, st_,
                               stats_.p, base + off, reinterpret_cast<cx<R> *>(v.w0), counts[k]);
            off += counts[k];
        }
    }

    std::vector<EngineBase *> engs_;
    std::vector<EngineViews> views_;
    std::vector<size_t> first_;
    size_t total_ = 0;
    int coarse_ = 0;
    hipStream_t st_ = nullptr;
    hipEvent_t done_ = nullptr, fork_ = nullptr;
    std::vector<hipEvent_t> ready_, join_; // before the cycle (eager), inside it (captured)
    double *host_stats_ = nullptr;
    Buf<uint64_t> cycle_; // group index, last-group flag
    // the captured cycle: key of its launch sequence
    struct Key {
        double dt;
        size_t steps, nobs;
        uint64_t generation;
        uint64_t epochs[MULTI_MAX];
        bool operator==(const Key &o) const {
            for (int k = 0; k < MULTI_MAX; ++k)
                if (epochs[k] != o.epochs[k])
                    return false;
            return dt == o.dt && steps == o.steps && nobs == o.nobs &&
                   generation == o.generation;
        }
    };
    hipGraphExec_t exec_ = nullptr;
    Key exec_key_{}, warm_key_{};
    bool warm_ = false;
    Buf<double> wstate_, wnorm_, logw_, stats_, uniforms_;
    Buf<int> anc_;
    Buf<uint64_t> keys_, seeds_;
    Buf<cx<double>> scratch_;
    int max_uniforms_ = 0;
    uint64_t seed_ = 0, group_ = 0;
    bool on_ = false;
};

} // namespace hzl_amd

// ---------------------------------------------------------------- C-ABI --
struct hzl_pf_joint {
    std::unique_ptr<hzl_amd::MultiFilter> f;
};

using hzl_amd::engine_of;
using hzl_amd::guarded;

namespace {
hzl_amd::MultiFilter &joint_of(hzl_pf_joint_t h) {
    if (!h || !h->f)
        throw std::invalid_argument("null joint filter handle");
    return *h->f;
}
} // namespace

extern "C" {

int hzl_pf_joint_create(hzl_pf_joint_t *out, const hzl_ensemble_t *ensembles, size_t count) {
    return guarded([&] {
        if (!out || !ensembles)
            throw std::invalid_argument("null argument");
        *out = nullptr;
        std::vector<hzl_amd::EngineBase *> e;
        for (size_t k = 0; k < count; ++k)
            e.push_back(&engine_of(ensembles[k]));
        auto h = std::make_unique<hzl_pf_joint>();
        h->f = std::make_unique<hzl_amd::MultiFilter>(e.data(), e.size());
        *out = h.release();
    });
}

int hzl_pf_joint_destroy(hzl_pf_joint_t h) {
    return guarded([&] { delete h; });
}

int hzl_pf_joint_begin(hzl_pf_joint_t h, uint64_t seed, size_t max_groups) {
    return guarded([&] { joint_of(h).begin(seed, max_groups); });
}

int hzl_pf_joint_assimilate(hzl_pf_joint_t h, double dt, size_t steps, size_t nobs, const int *x,
                            const int *y, const double *value, const double *sd, int is_last,
                            double *record) {
    return guarded([&] {
        if (!x || !y || !value || !sd)
            throw std::invalid_argument("null observations");
        joint_of(h).assimilate(dt, steps, nobs, x, y, value, sd, is_last != 0, record);
    });
}

int hzl_pf_joint_result(hzl_pf_joint_t h, double *log_likelihood, size_t *resampling_events) {
    return guarded([&] { joint_of(h).result(log_likelihood, resampling_events); });
}

} // extern "C"
