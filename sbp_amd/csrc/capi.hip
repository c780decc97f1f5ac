// C-ABI of the SBP hot path: handles, finetune loops with early stopping, block-merge carry-out, the
// agglomerative driver with its golden-section triplet, sampling.  Kernels live in build.hip, mdl.hip,
// finetune.hip, merge.hip and mh.hip.
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <numeric>
#include <queue>
#include <vector>

#include "device_common.hpp"
#include "internal.hpp"
#include "primitives.hpp"
#include "carry.hpp"
#include "host_blockmodel.hpp"
#include "snowball.hpp"
#include "../../include/sbp/sbp.hpp"

using namespace sbpimpl;

namespace sbpimpl {

double elapsed_ms(sbp_ctx ctx, hipEvent_t a, hipEvent_t b) {
    (void)ctx;
    float ms = 0.f;
    SBP_HIP(hipEventSynchronize(b));
    SBP_HIP(hipEventElapsedTime(&ms, a, b));
    return (double)ms;
}

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static sbp_params default_params() {
    sbp_params p;
    p.algorithm = 2;
    p.batches = 1;
    p.greedy = 1;
    p.approximate = 0;
    p.ordered = 0;
    p.asynciterations = 2147483647;
    p.mh_percent = 0.075;
    p.seed = 0;
    p.max_iterations = 100;
    p.threads = 1;
    p.nonparametric = 0;
    p.degreecorrected = 0;
    return p;
}

static sbp_bm new_bm(sbp_ctx ctx) {
    sbp_bm bm = new sbp_bm_s();
    bm->ctx = ctx;
    bm->pool = ctx->pool;
    return bm;
}

// entropy::mdl (src/entropy.cpp:597-604): nonparametric::mdl for nonparametric blockmodels
static double mdl_of(sbp_ctx ctx, sbp_bm bm, sbp_graph g, sbp_stats *st) {
    double t0 = now_ms();
    double m;
    if (bm->np) {
        m = np_mdl(ctx, g, bm, nullptr);
    } else {
        double lp = log_posterior(ctx, bm);
        m = model_term(g->E, g->N, bm->B) - lp;
    }
    if (st) st->mdl_ms += now_ms() - t0;
    return m;
}

// ------------------------------------------------------------------ finetune

static int32_t *iota_of(sbp_ctx ctx, sbp_graph g);

// One async-Gibbs batch: proposals for items[0, count) against the frozen blockmodel, then commit.
static void gibbs_batch(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int greedy, uint32_t epoch, const int32_t *d_items,
                        int64_t count, sbp_stats *st) {
    hipStream_t s = ctx->stream;
    int32_t *move_s = ctx->scratch[S_MOVES].as<int32_t>(g->N);
    SBP_HIP(sbpprim::fill<int32_t>(move_s, g->N, -1, s));
    double t0 = now_ms();
    FtLaunchInfo info = ft_propose(ctx, bm, g, epoch, greedy, d_items, count, move_s, nullptr, nullptr);
    int64_t refused = 0;
    double t1 = now_ms();
    const int64_t updates = ctx->cell_updates;
    int64_t moved = ft_commit(ctx, bm, g, move_s, &refused);
    double t2 = now_ms();
    if (st) {
        st->cell_updates += ctx->cell_updates - updates;
        st->proposals += count;
        st->moves += moved;
        st->refused += refused;
        st->ft_bytes += info.bytes;
        st->ft_launches += 1;
        st->ft_kernel_ms += info.kernel_ms;
        if (info.high_launched) {
            st->ft_high_bytes += info.high_bytes;
            st->ft_high_launches += 1;
            st->ft_high_kernel_ms += info.high_kernel_ms;
        }
        for (int b = 0; b < FT_NB; ++b)
            if (info.class_launched[b]) {
                st->ft_class_bytes[b] += info.class_bytes[b];
                st->ft_class_launches[b] += 1;
                st->ft_class_ms[b] += info.class_ms[b];
            }
        st->finetune_ms += (t1 - t0) + (t2 - t1);
    }
}

static void async_gibbs_sweep(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params &p, uint32_t epoch,
                              sbp_stats *st) {
    const int64_t N = g->N;
    int32_t *ident = iota_of(ctx, g);
    const int32_t *order = ident;
    if (!p.ordered && p.batches > 1) {
        // the shuffle only changes which vertices share a batch; with one batch it has no effect
        int32_t *sh = ctx->scratch[S_ORDER].as<int32_t>(N);
        shuffle_keys(ctx, ident, N, epoch, sh);
        order = sh;
    }
    int64_t batch_size = (int64_t)ceil((double)N / (double)p.batches);
    for (int64_t batch = 0; batch_size > 0 && batch < N / batch_size; ++batch) {
        int64_t start = batch * batch_size;
        int64_t end = std::min<int64_t>(N, (batch + 1) * batch_size);
        gibbs_batch(ctx, bm, g, p.greedy, epoch, order + start, end - start, st);
    }
    if (st) st->sweeps++;
}

// ------------------------------------------------------------------ recorded sweeps

// One asynchronous-Gibbs iteration (finetune::asynchronous_gibbs' loop body, src/finetune.cpp:50-88: the batches'
// proposals against the frozen blockmodel, each batch's commit in vertex-id order with the rebuild, then the MDL)
// recorded once into a HIP graph and replayed every sweep: a replay is one graph launch plus one read of its
// results, where the sweep run call by call waits on the device five times per batch.  The recorded launches read
// the epoch from device memory (written before each replay) and size nothing from host-side values of the
// changing state: the commit's event sort and segment scan read their counts on the device, and the rebuild and
// the MDL reduction are launched for the largest block matrix of B blocks (min(E, B^2) cells), with the cells,
// the cell hash and every scratch buffer reserved before recording.  Parametric entropy only (the nonparametric
// description length takes host-side constants of the state, the degree-corrected histograms host bookkeeping).
struct SweepResult {
    unsigned long long class_bytes[FT_NB];  // algorithmic bytes of the proposal launches per degree class
    unsigned long long err;                 // a proposal found degrees inconsistent with the block matrix
    unsigned long long moved, proposed;     // commits summed over the batches
    sbpdev::Fx lp;                          // log posterior, fixed point
    int32_t lp_err;                         // a non-zero cell with a zero degree
    uint32_t epoch;                         // written by the host before each replay
};

__global__ void k_sweep_begin(SweepResult *r) {
    if (threadIdx.x < FT_NB) r->class_bytes[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        r->err = 0;
        r->moved = 0;
        r->proposed = 0;
    }
}

__global__ void k_sweep_accumulate(SweepResult *r, const unsigned long long *__restrict__ prop,
                                   const unsigned long long *__restrict__ commit) {
    if (threadIdx.x < FT_NB) r->class_bytes[threadIdx.x] += prop[threadIdx.x];
    if (threadIdx.x == 0) {
        r->err |= prop[FT_NB];
        r->moved += commit[0];
        r->proposed += commit[2];
    }
}

struct RecordedSweep {
    hipGraphExec_t exec = nullptr;
    SweepResult *d = nullptr;      // device
    SweepResult *h = nullptr;      // pinned host copy
    uint32_t *h_epoch = nullptr;   // pinned host epoch for the copy before a replay
    // what the recording depends on
    sbp_bm bm = nullptr;
    sbp_graph g = nullptr;
    int32_t B = 0, greedy = 0, batches = 0, ordered = 0;
    int64_t proposals = 0;           // proposals per sweep (the batches' sizes)
    std::vector<const void *> deps;  // every device buffer the recorded launches address
    ~RecordedSweep() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (d) (void)dev_free(d);
        if (h) (void)dev_host_free(h);
        if (h_epoch) (void)dev_host_free(h_epoch);
    }
    static std::vector<const void *> buffers_of(sbp_ctx ctx, sbp_bm b) {
        std::vector<const void *> v = {b->assign,  b->sizes,   b->d_out,    b->d_in,      b->d_tot,    b->row_ptr,
                                       b->col_ptr, b->row_cells, b->col_cells, b->row_cum, b->col_cum, b->cell_hash};
        for (const Scratch &s : ctx->scratch) v.push_back(s.ptr);
        return v;
    }
    bool matches(sbp_ctx ctx, sbp_bm b, sbp_graph gr, const sbp_params &p) const {
        return exec && b == bm && gr == g && b->B == B && p.greedy == greedy && p.batches == batches &&
               p.ordered == ordered && b->hash_log2 == hash_log2 && buffers_of(ctx, b) == deps;
    }
    int hash_log2 = 0;
};

// Whether the partition driver runs this blockmodel's sweeps as graph replays (SBP_AMD_RECORDED_SWEEPS=0 turns
// them off, =1 on for every size).  A recording serves one finetune phase (each merge phase makes a new
// blockmodel) and costs milliseconds plus a pause of the other threads' allocations and frees; it pays for
// itself when sweeps are long, not for the few short sweeps of a small graph (measured: 200,000 vertices 6%
// faster per partition; the batched workload's graphs of up to 50,000 vertices 40% slower on 12 worker slots).
// By size, graphs whose eager commits update the cells in place (2^22 edges and more) run eagerly: the update
// is faster than the recorded rebuild (measured: the 5,000,000-vertex workload's 2,500,000-vertex sample, finetune
// 2.47 s eager against 2.78 s recorded).
constexpr int32_t RECORD_MIN_VERTICES = 100000;
static bool recorded_sweeps_enabled(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params &p) {
    static const int env_mode = [] {
        const char *e = getenv("SBP_AMD_RECORDED_SWEEPS");
        return e ? (e[0] == '0' ? 0 : 1) : -1;
    }();
    const int mode = ctx->record_mode >= 0 ? ctx->record_mode : env_mode;
    if (mode == 0 || bm->np || bm->dc || p.algorithm != 0 || bm->N <= 0 || bm->E <= 0) return false;
    return mode == 1 || (bm->N >= RECORD_MIN_VERTICES && !bm_update_wanted(g));
}

static void record_sweep(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params &p) {
    hipStream_t st = ctx->stream;
    auto rs = std::make_shared<RecordedSweep>();
    SBP_HIP(dev_malloc(&rs->d, sizeof(SweepResult)));
    SBP_HIP(hipHostMalloc(&rs->h, sizeof(SweepResult), hipHostMallocDefault));
    SBP_HIP(hipHostMalloc(&rs->h_epoch, 64, hipHostMallocDefault));
    const int64_t N = g->N;
    const int64_t batch_size = (int64_t)ceil((double)N / (double)p.batches);
    // reserve every buffer at its largest, then record
    bm_reserve_cells(ctx, g, bm);
    ft_reserve(ctx, bm, g, batch_size);
    log_posterior_reserve(ctx, std::min<int64_t>(g->E, (int64_t)bm->B * bm->B));
    SBP_HIP(hipStreamSynchronize(st));
    int32_t *ident = iota_of(ctx, g);
    uint32_t *d_epoch = &rs->d->epoch;
    const int64_t max_nnz = std::min<int64_t>(g->E, (int64_t)bm->B * bm->B);
    hipGraph_t graph = nullptr;
    // no hipFree or hipDeviceSynchronize of another thread may run while the stream is being recorded
    std::unique_lock<std::shared_mutex> quiet(device_wide_mu());
    SBP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    g_recording = true;
    try {
        k_sweep_begin<<<1, 64, 0, st>>>(rs->d);
        const int32_t *order = ident;
        if (!p.ordered && p.batches > 1) {
            int32_t *sh = ctx->scratch[S_ORDER].as<int32_t>(N);
            shuffle_keys(ctx, ident, N, 0, sh, d_epoch);
            order = sh;
        }
        int32_t *move_s = ctx->scratch[S_MOVES].as<int32_t>(N);
        for (int64_t batch = 0; batch_size > 0 && batch < N / batch_size; ++batch) {
            const int64_t start = batch * batch_size;
            const int64_t end = std::min<int64_t>(N, (batch + 1) * batch_size);
            SBP_HIP(sbpprim::fill<int32_t>(move_s, N, -1, st));
            ft_propose(ctx, bm, g, 0, p.greedy, order + start, end - start, move_s, nullptr, nullptr, d_epoch);
            ft_commit(ctx, bm, g, move_s, nullptr, true);
            k_sweep_accumulate<<<1, 64, 0, st>>>(rs->d, ft_propose_counters(ctx), ft_commit_counters(ctx));
        }
        log_posterior_device(ctx, bm, max_nnz, &rs->d->lp, &rs->d->lp_err);
        SBP_HIP(hipGetLastError());
    } catch (...) {
        g_recording = false;
        hipGraph_t aborted = nullptr;
        (void)hipStreamEndCapture(st, &aborted);
        if (aborted) (void)hipGraphDestroy(aborted);
        (void)hipGetLastError();
        throw;
    }
    g_recording = false;
    SBP_HIP(hipStreamEndCapture(st, &graph));
    const hipError_t e = hipGraphInstantiate(&rs->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SBP_HIP(e);
    rs->bm = bm;
    rs->g = g;
    rs->B = bm->B;
    rs->greedy = p.greedy;
    rs->batches = p.batches;
    rs->ordered = p.ordered;
    rs->hash_log2 = bm->hash_log2;
    for (int64_t b = 0; batch_size > 0 && b < N / batch_size; ++b)
        rs->proposals += std::min<int64_t>(N, (b + 1) * batch_size) - b * batch_size;
    rs->deps = RecordedSweep::buffers_of(ctx, bm);
    ctx->recorded = std::move(rs);
}

// One asynchronous-Gibbs iteration as a replay of the recorded graph (recorded first when the blockmodel, its
// buffers or the parameters changed); returns the log posterior after the sweep.
static double replay_sweep(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params &p, uint32_t epoch, sbp_stats *st) {
    if (!ctx->recorded || !ctx->recorded->matches(ctx, bm, g, p)) {
        ctx->recorded.reset();
        record_sweep(ctx, bm, g, p);
    }
    RecordedSweep &rs = *ctx->recorded;
    hipStream_t s = ctx->stream;
    *rs.h_epoch = epoch;
    SBP_HIP(hipMemcpyAsync(&rs.d->epoch, rs.h_epoch, sizeof(uint32_t), hipMemcpyHostToDevice, s));
    SBP_HIP(hipEventRecord(ctx->ev[0], s));
    SBP_HIP(hipGraphLaunch(rs.exec, s));
    SBP_HIP(hipEventRecord(ctx->ev[1], s));
    int32_t nnz = 0;
    SBP_HIP(hipMemcpyAsync(rs.h, rs.d, sizeof(SweepResult), hipMemcpyDeviceToHost, s));
    read_small(ctx, {{&nnz, bm->row_ptr + bm->B, sizeof(int32_t)}});
    const SweepResult &r = *rs.h;
    bm->nnz = nnz;
    bm->hash_valid = bm->dense_valid = false;
    SBP_CHECK(r.err == 0, SBP_E_STATE, "finetune proposal: block degrees inconsistent with the block matrix");
    SBP_CHECK(r.lp_err == 0, SBP_E_STATE, "log_posterior: non-zero cell with zero block degree");
    if (st) {
        st->proposals += rs.proposals;
        st->moves += (int64_t)r.moved;
        st->refused += (int64_t)r.proposed - (int64_t)r.moved;
        for (int b = 0; b < FT_NB; ++b) {
            st->ft_class_bytes[b] += r.class_bytes[b];
            if (b < FT_NB - 2) st->ft_bytes += r.class_bytes[b];
            else st->ft_high_bytes += r.class_bytes[b];
        }
        st->sweeps++;
        st->recorded_sweeps++;
        st->finetune_ms += elapsed_ms(ctx, ctx->ev[0], ctx->ev[1]);
    }
    return (double)r.lp.hi + (double)r.lp.lo * 0x1p-64;
}

// finetune::early_stop, 4-argument form (src/finetune.cpp:265-283).
static bool early_stop4(int64_t iteration, bool golden_ratio_not_reached, double initial, const std::vector<double> &de) {
    size_t last = de.size() - 1;
    if (de[last] == 0.0) return true;
    if (iteration < 3) return false;
    double average = de[last] + de[last - 1] + de[last - 2];
    average /= -3.0;
    double threshold = golden_ratio_not_reached ? 5e-4 * initial : 1e-4 * initial;
    return average < threshold;
}

// finetune::early_stop, 3-argument form (src/finetune.cpp:285-294).
static bool early_stop3(int64_t iteration, double initial, const std::vector<double> &de) {
    if (iteration < 3) return false;
    size_t last = de.size() - 1;
    double average = de[last] + de[last - 1] + de[last - 2];
    average /= -3.0;
    return average < 1e-4 * initial;
}

// Graph::sort_vertices (src/graph.cpp:235-268): stable descending degree; the first ceil(mh_percent * N)
// vertices are the high-degree ones.
static int64_t high_degree_count(sbp_graph g, double mh_percent) {
    int64_t nh = 0;
    while (nh < g->N && (double)nh < mh_percent * (double)g->N) ++nh;
    return nh;
}

void mh_run(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params &p, uint32_t epoch, const int32_t *d_order,
            int64_t count, double *sum_dS, sbp_stats *st);

static void mcmc(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params &p, bool golden_ratio_not_reached,
                 uint32_t *epoch, sbp_stats *st) {
    if (bm->B == 1) return;
    const int max_it = p.max_iterations > 0 ? p.max_iterations : 100;
    std::vector<double> de;
    if (p.algorithm == 0 || p.algorithm == 2) {
        double initial = mdl_of(ctx, bm, g, st), last = initial;
        bm->overall_entropy = initial;
        const int32_t *d_h = nullptr, *d_l = nullptr;
        int64_t nh = 0, nl = 0;
        if (p.algorithm == 2) {
            d_h = degree_order(ctx, g);
            nh = high_degree_count(g, p.mh_percent);
            nl = g->N - nh;
            d_l = d_h + nh;
        }
        const bool replay = p.algorithm == 0 && recorded_sweeps_enabled(ctx, bm, g, p);
        for (int64_t it = 0; it < max_it; ++it) {
            uint32_t ep = (*epoch)++;
            if (replay) {
                // the sweep and the MDL that closes the iteration, as one graph launch
                const double lp = replay_sweep(ctx, bm, g, p, ep, st);
                const double e = model_term(g->E, g->N, bm->B) - lp;
                de.push_back(e - last);
                last = e;
                if (early_stop4(it, golden_ratio_not_reached, initial, de)) break;
                continue;
            }
            if (p.algorithm == 0) {
                async_gibbs_sweep(ctx, bm, g, p, ep, st);
            } else {
                // hybrid_mcmc (src/finetune.cpp:563-646)
                int32_t *h = ctx->scratch[S_ORDER].as<int32_t>(nh + nl + 2);
                int32_t *l = h + nh + 1;
                if (!p.ordered) {
                    shuffle_keys(ctx, d_h, nh, ep, h);
                    shuffle_keys(ctx, d_l, nl, ep, l);
                } else {
                    SBP_HIP(hipMemcpyAsync(h, d_h, sizeof(int32_t) * nh, hipMemcpyDeviceToDevice, ctx->stream));
                    SBP_HIP(hipMemcpyAsync(l, d_l, sizeof(int32_t) * nl, hipMemcpyDeviceToDevice, ctx->stream));
                }
                double sum = 0.0;
                mh_run(ctx, bm, g, p, ep, h, nh, &sum, st);
                int64_t batch_size = (int64_t)ceil((double)nl / (double)p.batches);
                for (int64_t batch = 0; batch_size > 0 && batch < nl / batch_size; ++batch) {
                    int64_t start = batch * batch_size;
                    int64_t end = std::min<int64_t>(nl, (batch + 1) * batch_size);
                    gibbs_batch(ctx, bm, g, p.greedy, ep, l + start, end - start, st);
                }
                if (st) st->sweeps++;
            }
            double e = mdl_of(ctx, bm, g, st);
            de.push_back(e - last);
            last = e;
            if (early_stop4(it, golden_ratio_not_reached, initial, de)) break;
        }
        bm->overall_entropy = last;
    } else {
        // metropolis_hastings (src/finetune.cpp:773-820)
        bm->overall_entropy = mdl_of(ctx, bm, g, st);
        int32_t *ident = iota_of(ctx, g);
        int32_t *order = ctx->scratch[S_ORDER].as<int32_t>(g->N);
        for (int64_t it = 0; it < max_it; ++it) {
            uint32_t ep = (*epoch)++;
            const int32_t *ord = ident;
            if (!p.ordered) {
                shuffle_keys(ctx, ident, g->N, ep, order);
                ord = order;
            }
            double sum = 0.0;
            mh_run(ctx, bm, g, p, ep, ord, g->N, &sum, st);
            if (st) st->sweeps++;
            de.push_back(sum);
            if (early_stop4(it, golden_ratio_not_reached, bm->overall_entropy, de)) break;
        }
        bm->overall_entropy = mdl_of(ctx, bm, g, st);
    }
}

static void finetune_assignment(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params &p, uint32_t *epoch,
                                sbp_stats *st) {
    const int max_it = p.max_iterations > 0 ? p.max_iterations : 100;
    std::vector<double> de;
    bm->overall_entropy = mdl_of(ctx, bm, g, st);
    // One block: no vertex has another block to move to.  The reference walks the sweeps anyway (every proposal
    // is the current block, and an isolated vertex draws from the empty range U[0, B-2]); the result is the same
    // blockmodel, so the walk is skipped, as metropolis_hastings does (src/finetune.cpp:775-777).
    if (bm->B < 2) return;
    int32_t *ident = iota_of(ctx, g);
    for (int64_t it = 0; it < max_it; ++it) {
        uint32_t ep = (*epoch)++;
        double sum = 0.0;
        mh_run(ctx, bm, g, p, ep, ident, g->N, &sum, st);
        if (st) st->sweeps++;
        de.push_back(sum);
        if (early_stop3(it, bm->overall_entropy, de)) break;
    }
    bm->overall_entropy = mdl_of(ctx, bm, g, st);
}

// ------------------------------------------------------------------ block merge

static sbp_bm merge_apply(sbp_ctx ctx, sbp_bm bm, sbp_graph g, std::vector<int32_t> &best_s,
                          std::vector<double> &best_dS, int64_t num_to_merge, bool approximate, sbp_stats *st);

static sbp_bm merge_blocks(sbp_ctx ctx, sbp_bm bm, sbp_graph g, uint32_t epoch, int64_t num_to_merge,
                           bool approximate, sbp_stats *st) {
    hipStream_t s = ctx->stream;
    const int32_t B = bm->B;
    int32_t *d_best_s = ctx->scratch[S_BEST].as<int32_t>(3 * (size_t)B + 4);
    double *d_best_dS = reinterpret_cast<double *>(ctx->scratch[S_DS].as<double>((size_t)B + 2));
    SBP_HIP(hipEventRecord(ctx->ev[2], s));
    uint64_t bytes = merge_propose(ctx, bm, epoch, d_best_s, d_best_dS);
    SBP_HIP(hipEventRecord(ctx->ev[3], s));
    std::vector<int32_t> best_s(B);
    std::vector<double> best_dS(B);
    SBP_HIP(hipMemcpyAsync(best_s.data(), d_best_s, sizeof(int32_t) * B, hipMemcpyDeviceToHost, s));
    SBP_HIP(hipMemcpyAsync(best_dS.data(), d_best_dS, sizeof(double) * B, hipMemcpyDeviceToHost, s));
    SBP_HIP(hipStreamSynchronize(s));
    if (st) {
        st->merge_ms += elapsed_ms(ctx, ctx->ev[2], ctx->ev[3]);
        st->merge_bytes += bytes;
        st->merge_proposals += 10 * (int64_t)B;
        st->merge_phases++;
    }
    return merge_apply(ctx, bm, g, best_s, best_dS, num_to_merge, approximate, st);
}

// The carry-out of given proposals (best_s / best_dS per block), label compaction and rebuild: a new blockmodel
// with B - num_to_merge blocks.
static sbp_bm merge_apply(sbp_ctx ctx, sbp_bm bm, sbp_graph g, std::vector<int32_t> &best_s,
                          std::vector<double> &best_dS, int64_t num_to_merge, bool approximate, sbp_stats *st) {
    hipStream_t s = ctx->stream;
    const int32_t B = bm->B;
    int32_t *d_best_s = ctx->scratch[S_BEST].as<int32_t>(3 * (size_t)B + 4);
    double t0 = now_ms();
    std::vector<int32_t> block_map, sizes(B);
    SBP_HIP(dev_copy(sizes.data(), bm->sizes, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
    sbpcarry::Model model;
    sbpcarry::Pool *pool = nullptr;
    double t_copied = t0;
    if (!approximate) {  // the advanced carry-out re-evaluates merges against the changing blockmodel
        // the CSR, CSC and degrees in one pinned staging area that the context keeps (no pageable copies, no
        // freshly faulted pages per merge phase); int2 cells start at 16 B + 8 bytes, a multiple of 8
        const size_t words = 2 * (size_t)(B + 1) + 2 * (size_t)B;
        const size_t need = sizeof(int32_t) * words + 2 * sizeof(int2) * (size_t)bm->nnz;
        if (ctx->stage_bytes < need) {
            if (ctx->stage) SBP_HIP(dev_host_free(ctx->stage));
            ctx->stage = nullptr;
            ctx->stage_bytes = 0;
            const size_t bytes = need + need / 4 + 4096;
            SBP_HIP(hipHostMalloc(&ctx->stage, bytes, hipHostMallocDefault));
            ctx->stage_bytes = bytes;
        }
        int32_t *rp = static_cast<int32_t *>(ctx->stage), *cp = rp + (B + 1), *dout = cp + (B + 1), *din = dout + B;
        int2 *rc = reinterpret_cast<int2 *>(din + B), *cc = rc + bm->nnz;
        SBP_HIP(hipMemcpyAsync(rp, bm->row_ptr, sizeof(int32_t) * (B + 1), hipMemcpyDeviceToHost, s));
        SBP_HIP(hipMemcpyAsync(cp, bm->col_ptr, sizeof(int32_t) * (B + 1), hipMemcpyDeviceToHost, s));
        SBP_HIP(hipMemcpyAsync(dout, bm->d_out, sizeof(int32_t) * B, hipMemcpyDeviceToHost, s));
        SBP_HIP(hipMemcpyAsync(din, bm->d_in, sizeof(int32_t) * B, hipMemcpyDeviceToHost, s));
        if (bm->nnz) {
            SBP_HIP(hipMemcpyAsync(rc, bm->row_cells, sizeof(int2) * bm->nnz, hipMemcpyDeviceToHost, s));
            SBP_HIP(hipMemcpyAsync(cc, bm->col_cells, sizeof(int2) * bm->nnz, hipMemcpyDeviceToHost, s));
        }
        SBP_HIP(hipStreamSynchronize(s));
        t_copied = now_ms();
        int threads = ctx->carry_threads > 0 ? ctx->carry_threads : sbpcarry::carry_threads();
        if (ctx->carry_share) {  // the host's carry threads shared among the requests running now
            const int active = std::max(1, (int)ctx->carry_share->load(std::memory_order_relaxed));
            threads = std::max(1, sbpcarry::carry_threads() / active);
        }
        // the parametric carry-out applies merges on a thread of its own beside the pool: leave it a core
        if (!bm->np && B >= sbpcarry::PIPE_MIN_BLOCKS && bm->nnz < sbpcarry::PIPE_MAX_ROW_CELLS * (int64_t)B &&
            threads > 2 && sbpcarry::pipe_enabled())
            --threads;
        if (!ctx->carry_pool || ctx->carry_pool->threads() != threads)
            ctx->carry_pool = std::make_shared<sbpcarry::Pool>(threads);
        pool = ctx->carry_pool.get();
        const int32_t *rcp = reinterpret_cast<const int32_t *>(rc);
        const int32_t *ccp = reinterpret_cast<const int32_t *>(cc);
        model.load(B, rp, rcp, rcp + 1, cp, ccp, ccp + 1, 2, dout, din, pool);
        if (bm->np) model.set_nonparametric(sizes.data(), bm->N);
        if (bm->dc) {
            SBP_CHECK(bm->hist_ready, SBP_E_STATE, "degree histograms not built");
            const int64_t n = int64_t(1) << bm->hist_log2;
            std::vector<sbpdev::HistSlot> h((size_t)n);
            SBP_HIP(dev_copy(h.data(), bm->hist, sizeof(sbpdev::HistSlot) * n, hipMemcpyDeviceToHost));
            std::vector<uint64_t> keys((size_t)n);
            std::vector<int32_t> counts((size_t)n);
            for (int64_t i = 0; i < n; ++i) {
                keys[i] = h[i].key;
                counts[i] = h[i].val;
            }
            model.set_degree_corrected(keys.data(), counts.data(), n);
        }
    }
    const double t_loaded = now_ms();
    sbpcarry::CarryStats cs;
    int32_t new_B = sbpcarry::carry_out(approximate ? nullptr : &model, B, best_s.data(), best_dS.data(),
                                        sizes.data(), num_to_merge, approximate, block_map, &cs, pool);
    if (st) {
        st->carry_ms += now_ms() - t0;
        st->carry_load_ms += t_loaded - t0;
        st->carry_evaluations += cs.evaluations;
        st->carry_speculative += cs.speculative;
        st->carry_reused += cs.reused;
        st->carry_requeues += cs.requeues;
    }
    if (getenv("SBP_AMD_TRACE"))
        fprintf(stderr, "[sbp] merge B=%d -> %d: propose %.1f ms, carry %.1f ms (load %.1f, copy %.1f), pops %lld, "
                "evals %lld, speculative %lld, reused %lld, requeues %lld\n", B, new_B,
                elapsed_ms(ctx, ctx->ev[2], ctx->ev[3]), now_ms() - t0, t_loaded - t0, t_copied - t0,
                (long long)cs.pops, (long long)cs.evaluations,
                (long long)cs.speculative, (long long)cs.reused, (long long)cs.requeues);
    double t1 = now_ms();
    int32_t *d_map = d_best_s;  // reuse
    SBP_HIP(hipMemcpyAsync(d_map, block_map.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice, s));
    sbp_bm out = new_bm(ctx);
    bm_alloc(out, new_B, bm->N);
    out->np = bm->np;
    out->dc = bm->dc;  // the rebuild below builds its degree histograms afresh (initialize_edge_counts)
    relabel_assignment(ctx, bm, d_map, new_B, out->assign);
    bm_rebuild(ctx, g, out);
    SBP_HIP(hipStreamSynchronize(s));
    out->overall_entropy = bm->overall_entropy;
    if (st) st->build_ms += now_ms() - t1;
    return out;
}

static int32_t *iota_of(sbp_ctx ctx, sbp_graph g) {
    (void)ctx;
    return g->edge_src + g->E;  // identity vector stored after the edge sources at upload
}

}  // namespace sbpimpl

// ====================================================================== C-ABI

#define SBP_API_BEGIN try {
#define SBP_API_END(ctx)                                  \
    }                                                     \
    catch (const sbpimpl::Error &e) {                     \
        if (ctx) (ctx)->err = e.what();                   \
        return e.code;                                    \
    }                                                     \
    catch (const std::bad_alloc &) {                      \
        if (ctx) (ctx)->err = "host out of memory";       \
        return SBP_E_OOM;                                 \
    }                                                     \
    catch (const std::exception &e) {                     \
        if (ctx) (ctx)->err = e.what();                   \
        return SBP_E_STATE;                               \
    }                                                     \
    return SBP_OK;

extern "C" {

// finetune launch shape that new contexts start with (sbp_set_finetune_shape(NULL, ...))
static std::mutex g_shape_mu;
static int32_t g_default_lanes[5] = {2, 4, 16, 64, 256};
static int32_t g_default_occ[4] = {4, 4, 4, 4};

int sbp_create(sbp_ctx *out, int device, uint64_t seed) {
    if (!out) return SBP_E_ARG;
    *out = nullptr;
    sbp_ctx ctx = nullptr;
    try {
        int n = 0;
        SBP_HIP(hipGetDeviceCount(&n));
        SBP_CHECK(device >= 0 && device < n, SBP_E_ARG, "no such HIP device");
        SBP_HIP(hipSetDevice(device));
        noalloc_setup(device);
        ctx = new sbp_ctx_s();
        ctx->device = device;
        ctx->seed = seed;
        {
            std::lock_guard<std::mutex> lk(g_shape_mu);
            std::copy(g_default_lanes, g_default_lanes + 5, ctx->ft_lanes);
            std::copy(g_default_occ, g_default_occ + 4, ctx->ft_occ);
        }
        SBP_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        for (auto &e : ctx->ev) SBP_HIP(hipEventCreate(&e));
        for (auto &e : ctx->ft_ev) SBP_HIP(hipEventCreate(&e));
        SBP_HIP(hipHostMalloc(&ctx->pinned, PINNED_BYTES, hipHostMallocDefault));
        *out = ctx;
        return SBP_OK;
    } catch (const sbpimpl::Error &e) {
        delete ctx;
        return e.code;
    }
}

int sbp_destroy(sbp_ctx ctx) {
    if (ctx) ctx->recorded.reset();
    if (!ctx) return SBP_OK;
    for (sbp_ctx w : ctx->slots) sbp_destroy(w);
    ctx->slots.clear();
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : ctx->ft_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &s : ctx->scratch) {
        if (s.ptr) (void)dev_free(s.ptr);
        s.ptr = nullptr;
        s.bytes = 0;
    }
    if (ctx->lgamma) (void)dev_free(ctx->lgamma);
    if (ctx->pinned) (void)dev_host_free(ctx->pinned);
    if (ctx->stage) (void)dev_host_free(ctx->stage);
    ctx->pool.reset();  // freed here unless blockmodels of this context still hold buffers
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SBP_OK;
}

const char *sbp_last_error(sbp_ctx ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sbp_set_seed(sbp_ctx ctx, uint64_t seed) {
    if (!ctx) return SBP_E_ARG;
    ctx->seed = seed;
    return SBP_OK;
}

int sbp_synchronize(sbp_ctx ctx) {
    if (!ctx) return SBP_E_ARG;
    SBP_API_BEGIN
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    SBP_API_END(ctx)
}

int sbp_graph_upload(sbp_ctx ctx, int32_t N, int64_t E, const int64_t *out_off, const int32_t *out_adj,
                     const int64_t *in_off, const int32_t *in_adj, sbp_graph *out) {
    if (!ctx || !out || N <= 0 || E < 0 || !out_off || !in_off || (E > 0 && (!out_adj || !in_adj))) return SBP_E_ARG;
    *out = nullptr;
    SBP_API_BEGIN
    SBP_CHECK(out_off[0] == 0 && out_off[N] == E && in_off[0] == 0 && in_off[N] == E, SBP_E_ARG,
              "graph offsets do not match E");
    for (int32_t v = 0; v < N; ++v) {
        SBP_CHECK(out_off[v + 1] >= out_off[v] && in_off[v + 1] >= in_off[v], SBP_E_ARG, "offsets not monotone");
        for (int64_t e = out_off[v]; e < out_off[v + 1]; ++e)
            SBP_CHECK(out_adj[e] >= 0 && out_adj[e] < N, SBP_E_ARG, "out_adj id out of range");
        for (int64_t e = in_off[v]; e < in_off[v + 1]; ++e)
            SBP_CHECK(in_adj[e] >= 0 && in_adj[e] < N, SBP_E_ARG, "in_adj id out of range");
    }
    auto g = std::unique_ptr<sbp_graph_s, int (*)(sbp_graph)>(new sbp_graph_s(), sbp_graph_free);
    g->ctx = ctx;
    g->N = N;
    g->E = E;
    SBP_HIP(dev_malloc(&g->out_off, sizeof(int64_t) * (N + 1)));
    SBP_HIP(dev_malloc(&g->in_off, sizeof(int64_t) * (N + 1)));
    SBP_HIP(dev_malloc(&g->out_adj, sizeof(int32_t) * (E + 1)));
    SBP_HIP(dev_malloc(&g->in_adj, sizeof(int32_t) * (E + 1)));
    SBP_HIP(dev_copy(g->out_off, out_off, sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice));
    SBP_HIP(dev_copy(g->in_off, in_off, sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice));
    if (E > 0) {
        SBP_HIP(dev_copy(g->out_adj, out_adj, sizeof(int32_t) * E, hipMemcpyHostToDevice));
        SBP_HIP(dev_copy(g->in_adj, in_adj, sizeof(int32_t) * E, hipMemcpyHostToDevice));
    }
    graph_finish(ctx, g.get());
    *out = g.release();
    SBP_API_END(ctx)
}

int sbp_graph_free(sbp_graph g) {
    if (!g) return SBP_OK;
    void *p[] = {g->out_off, g->in_off,    g->out_adj,    g->in_adj,          g->edge_src, g->degree,
                 g->deg_order, g->parent_map, g->sample_vertices, g->ft_blist, g->ft_bmeta};
    for (void *x : p)
        if (x) (void)dev_free(x);
    delete g;
    return SBP_OK;
}

int sbp_graph_info(sbp_graph g, int32_t *num_vertices, int64_t *num_edges) {
    if (!g) return SBP_E_ARG;
    if (num_vertices) *num_vertices = g->N;
    if (num_edges) *num_edges = g->E;
    return SBP_OK;
}

int sbp_graph_download(sbp_ctx ctx, sbp_graph g, int64_t *out_off, int32_t *out_adj, int64_t *in_off,
                       int32_t *in_adj) {
    if (!ctx || !g) return SBP_E_ARG;
    SBP_API_BEGIN
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    const size_t no = sizeof(int64_t) * ((size_t)g->N + 1), na = sizeof(int32_t) * (size_t)g->E;
    if (out_off) SBP_HIP(dev_copy(out_off, g->out_off, no, hipMemcpyDeviceToHost));
    if (in_off) SBP_HIP(dev_copy(in_off, g->in_off, no, hipMemcpyDeviceToHost));
    if (out_adj && na) SBP_HIP(dev_copy(out_adj, g->out_adj, na, hipMemcpyDeviceToHost));
    if (in_adj && na) SBP_HIP(dev_copy(in_adj, g->in_adj, na, hipMemcpyDeviceToHost));
    SBP_API_END(ctx)
}

int sbp_sample(sbp_ctx ctx, sbp_graph g, int32_t alg, double fraction, uint64_t seed, sbp_graph *sample_out,
               int32_t *num_sampled, int32_t *mapping_out) {
    if (!ctx || !g || !sample_out || !(fraction >= 0.0 && fraction <= 1.0) || alg < 0 || alg > 5) return SBP_E_ARG;
    *sample_out = nullptr;
    SBP_API_BEGIN
    sbp_graph s = sample_graph(ctx, g, alg, fraction, seed);
    if (num_sampled) *num_sampled = s->N;
    if (mapping_out) SBP_HIP(dev_copy(mapping_out, s->parent_map, sizeof(int32_t) * g->N, hipMemcpyDeviceToHost));
    *sample_out = s;
    SBP_API_END(ctx)
}

int sbp_detach(sbp_ctx ctx, sbp_graph g, sbp_graph *detached_out, int32_t *num_kept, int32_t *mapping_out) {
    if (!ctx || !g || !detached_out) return SBP_E_ARG;
    *detached_out = nullptr;
    SBP_API_BEGIN
    sbp_graph s = sample_graph(ctx, g, SAMPLE_DETACH, 1.0, 0);
    if (num_kept) *num_kept = s->N;
    if (mapping_out) SBP_HIP(dev_copy(mapping_out, s->parent_map, sizeof(int32_t) * g->N, hipMemcpyDeviceToHost));
    *detached_out = s;
    SBP_API_END(ctx)
}

int sbp_reattach(int32_t N, const int64_t *out_off, const int32_t *out_adj, const int64_t *in_off,
                 const int32_t *in_adj, const int32_t *mapping, const int32_t *detached_assignment,
                 int32_t detached_blocks, uint64_t seed, int32_t *assignment_out) {
    if (N <= 0 || !out_off || !in_off || !mapping || !detached_assignment || detached_blocks < 1 || !assignment_out)
        return SBP_E_ARG;
    if ((out_off[N] > 0 && !out_adj) || (in_off[N] > 0 && !in_adj)) return SBP_E_ARG;
    try {
        const sbpsnow::Csr c{N, out_off, in_off, out_adj, in_adj};
        sbpsnow::reattach(c, mapping, detached_assignment, detached_blocks, seed, assignment_out);
    } catch (const std::bad_alloc &) {
        return SBP_E_OOM;
    }
    return SBP_OK;
}

int sbp_extend(sbp_ctx ctx, sbp_graph g, sbp_graph sample, sbp_bm sample_bm, uint64_t seed, sbp_bm *out,
               int32_t *assignment_out) {
    if (!ctx || !g || !sample || !sample_bm || !out) return SBP_E_ARG;
    *out = nullptr;
    SBP_API_BEGIN
    sbp_bm bm = new_bm(ctx);
    try {
        bm_alloc(bm, sample_bm->B, g->N);
        extend_assignment(ctx, g, sample, sample_bm, seed, bm->assign);
        bm_rebuild(ctx, g, bm);
        SBP_HIP(hipStreamSynchronize(ctx->stream));
        bm->num_to_merge = (int64_t)((double)bm->B * 0.5);
        if (assignment_out)
            SBP_HIP(dev_copy(assignment_out, bm->assign, sizeof(int32_t) * g->N, hipMemcpyDeviceToHost));
    } catch (...) {
        bm_release(bm);
        delete bm;
        throw;
    }
    *out = bm;
    SBP_API_END(ctx)
}

int sbp_bm_build(sbp_ctx ctx, sbp_graph g, int32_t B, const int32_t *assignment, sbp_bm *out) {
    if (!ctx || !g || !assignment || !out || B <= 0) return SBP_E_ARG;
    *out = nullptr;
    SBP_API_BEGIN
    for (int32_t v = 0; v < g->N; ++v)
        SBP_CHECK(assignment[v] >= 0 && assignment[v] < B, SBP_E_ARG, "assignment value out of [0, B)");
    sbp_bm bm = new_bm(ctx);
    try {
        bm_alloc(bm, B, g->N);
        // on the context's stream: the buffer may be recycled from a blockmodel that queued work still reads
        SBP_HIP(hipMemcpyAsync(bm->assign, assignment, sizeof(int32_t) * g->N, hipMemcpyHostToDevice, ctx->stream));
        bm_rebuild(ctx, g, bm);
        SBP_HIP(hipStreamSynchronize(ctx->stream));
        bm->num_to_merge = (int64_t)((double)B * 0.5);
    } catch (...) {
        bm_release(bm);
        delete bm;
        throw;
    }
    *out = bm;
    SBP_API_END(ctx)
}

int sbp_bm_copy(sbp_ctx ctx, sbp_bm src, sbp_bm *out) {
    if (!ctx || !src || !out) return SBP_E_ARG;
    *out = nullptr;
    SBP_API_BEGIN
    sbp_bm bm = new_bm(ctx);
    bm_copy_into(ctx, src, bm);
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    *out = bm;
    SBP_API_END(ctx)
}

int sbp_bm_free(sbp_bm bm) {
    if (!bm) return SBP_OK;
    bm_release(bm);
    delete bm;
    return SBP_OK;
}

int sbp_bm_info(sbp_bm bm, int32_t *num_blocks, int64_t *nnz) {
    if (!bm) return SBP_E_ARG;
    if (num_blocks) *num_blocks = bm->B;
    if (nnz) *nnz = bm->nnz;
    return SBP_OK;
}

int sbp_bm_download(sbp_ctx ctx, sbp_bm bm, int32_t *assignment, int32_t *sizes, int32_t *d_out, int32_t *d_in,
                    int32_t *d_tot, int32_t *row_ptr, int32_t *cols, int32_t *cnts) {
    if (!ctx || !bm) return SBP_E_ARG;
    SBP_API_BEGIN
    hipStream_t s = ctx->stream;
    SBP_HIP(hipStreamSynchronize(s));
    const int32_t B = bm->B;
    if (assignment) SBP_HIP(dev_copy(assignment, bm->assign, sizeof(int32_t) * bm->N, hipMemcpyDeviceToHost));
    if (sizes) SBP_HIP(dev_copy(sizes, bm->sizes, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
    if (d_out) SBP_HIP(dev_copy(d_out, bm->d_out, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
    if (d_in) SBP_HIP(dev_copy(d_in, bm->d_in, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
    if (d_tot) SBP_HIP(dev_copy(d_tot, bm->d_tot, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
    if (row_ptr) SBP_HIP(dev_copy(row_ptr, bm->row_ptr, sizeof(int32_t) * (B + 1), hipMemcpyDeviceToHost));
    if ((cols || cnts) && bm->nnz > 0) {
        std::vector<int2> cells(bm->nnz);
        SBP_HIP(dev_copy(cells.data(), bm->row_cells, sizeof(int2) * bm->nnz, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < bm->nnz; ++i) {
            if (cols) cols[i] = cells[i].x;
            if (cnts) cnts[i] = cells[i].y;
        }
    }
    SBP_API_END(ctx)
}

int sbp_bm_validate(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int32_t *ok) {
    if (!ctx || !bm || !g || !ok) return SBP_E_ARG;
    SBP_API_BEGIN
    sbp_bm fresh = new_bm(ctx);
    bm_alloc(fresh, bm->B, bm->N);
    SBP_HIP(hipMemcpyAsync(fresh->assign, bm->assign, sizeof(int32_t) * bm->N, hipMemcpyDeviceToDevice, ctx->stream));
    bm_rebuild(ctx, g, fresh);
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    HostBlockmodel a, b;
    a.download(ctx, bm);
    b.download(ctx, fresh);
    *ok = a.equals(b) ? 1 : 0;
    bm_release(fresh);
    delete fresh;
    SBP_API_END(ctx)
}

int sbp_mdl(sbp_ctx ctx, sbp_bm bm, sbp_graph g, double *mdl, double *lp) {
    if (!ctx || !bm || !g) return SBP_E_ARG;
    SBP_API_BEGIN
    if (bm->np) {
        const double m = np_mdl(ctx, g, bm, nullptr);
        if (lp) *lp = NAN;
        if (mdl) *mdl = m;
    } else {
        double l = log_posterior(ctx, bm);
        if (lp) *lp = l;
        if (mdl) *mdl = model_term(g->E, g->N, bm->B) - l;
    }
    SBP_API_END(ctx)
}

int sbp_mdl_noalloc(sbp_ctx ctx, sbp_bm bm, sbp_graph g, double *mdl, double *lp) {
    if (!ctx || !bm || !g) return SBP_E_ARG;
    SBP_API_BEGIN
    SBP_CHECK(!bm->np, SBP_E_ARG, "sbp_mdl_noalloc: the parametric description length only (entropy::mdl)");
    const double l = log_posterior_noalloc(ctx, bm);
    if (lp) *lp = l;
    if (mdl) *mdl = model_term(g->E, g->N, bm->B) - l;
    SBP_API_END(ctx)
}

int sbp_bm_set_model(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int32_t model) {
    if (!ctx || !bm || !g || model < 0 || model > 2 || g->N != bm->N) return SBP_E_ARG;
    SBP_API_BEGIN
    bm_set_model(ctx, g, bm, model >= 1, model == 2);
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    SBP_API_END(ctx)
}

int sbp_np_description_length(sbp_ctx ctx, sbp_bm bm, sbp_graph g, double *mdl, double *sparse_entropy,
                              double *partition_dl, double *edges_dl, double *degree_dl) {
    if (!ctx || !bm || !g) return SBP_E_ARG;
    SBP_API_BEGIN
    double parts[4];
    const double m = np_mdl(ctx, g, bm, parts);
    if (mdl) *mdl = m;
    if (sparse_entropy) *sparse_entropy = parts[0];
    if (partition_dl) *partition_dl = parts[1];
    if (edges_dl) *edges_dl = parts[2];
    if (degree_dl) *degree_dl = parts[3];
    SBP_API_END(ctx)
}

int sbp_ft_propose_all(sbp_ctx ctx, sbp_bm bm, sbp_graph g, uint32_t epoch, int32_t greedy, int32_t *s_out,
                       double *dS_out, int32_t *accept_out) {
    if (!ctx || !bm || !g) return SBP_E_ARG;
    SBP_API_BEGIN
    hipStream_t s = ctx->stream;
    const int32_t N = g->N;
    int32_t *move_s = ctx->scratch[S_MOVES].as<int32_t>(N);
    double *dS = ctx->scratch[S_DS].as<double>(N);
    int32_t *sv = ctx->scratch[S_ACCEPT].as<int32_t>(N);
    SBP_HIP(hipMemsetAsync(move_s, 0xff, sizeof(int32_t) * N, s));
    SBP_HIP(hipMemsetAsync(dS, 0, sizeof(double) * N, s));
    ft_propose(ctx, bm, g, epoch, greedy, iota_of(ctx, g), N, move_s, dS, sv);
    std::vector<int32_t> hm(N);
    SBP_HIP(dev_copy(hm.data(), move_s, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
    if (s_out) SBP_HIP(dev_copy(s_out, sv, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
    if (dS_out) SBP_HIP(dev_copy(dS_out, dS, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (accept_out)
        for (int32_t v = 0; v < N; ++v) accept_out[v] = hm[v] >= 0 ? 1 : 0;
    SBP_API_END(ctx)
}

int sbp_async_gibbs_sweep(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params *p, uint32_t epoch,
                          sbp_stats *stats) {
    if (!ctx || !bm || !g || !p || p->batches < 1) return SBP_E_ARG;
    SBP_API_BEGIN
    bm_set_model(ctx, g, bm, p->nonparametric, p->nonparametric && p->degreecorrected);
    async_gibbs_sweep(ctx, bm, g, *p, epoch, stats);
    SBP_API_END(ctx)
}

int sbp_async_gibbs_sweep_recorded(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params *p, uint32_t epoch,
                                   double *mdl_out, sbp_stats *stats) {
    if (!ctx || !bm || !g || !p || p->batches < 1 || p->nonparametric || bm->np || bm->dc || bm->E == 0 || g->E == 0)
        return SBP_E_ARG;
    SBP_API_BEGIN
    sbp_params q = *p;
    q.algorithm = 0;
    const double lp = replay_sweep(ctx, bm, g, q, epoch, stats);
    const double m = model_term(g->E, g->N, bm->B) - lp;
    bm->overall_entropy = m;
    if (mdl_out) *mdl_out = m;
    SBP_API_END(ctx)
}

int sbp_mh_sweep(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params *p, uint32_t epoch, int32_t ordered,
                 double *sum_dS, sbp_stats *stats) {
    if (!ctx || !bm || !g || !p) return SBP_E_ARG;
    SBP_API_BEGIN
    bm_set_model(ctx, g, bm, p->nonparametric, p->nonparametric && p->degreecorrected);
    if (bm->B < 2) {  // one block: nothing can move (src/finetune.cpp:775-777)
        if (sum_dS) *sum_dS = 0.0;
        if (stats) stats->sweeps++;
        return SBP_OK;
    }
    int32_t *ident = iota_of(ctx, g);
    const int32_t *ord = ident;
    if (!ordered) {
        int32_t *order = ctx->scratch[S_ORDER].as<int32_t>(g->N);
        shuffle_keys(ctx, ident, g->N, epoch, order);
        ord = order;
    }
    double sum = 0.0;
    mh_run(ctx, bm, g, *p, epoch, ord, g->N, &sum, stats);
    if (stats) stats->sweeps++;
    if (sum_dS) *sum_dS = sum;
    SBP_API_END(ctx)
}

int sbp_mcmc(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params *p, int32_t golden_ratio_not_reached,
             uint32_t *epoch, double *mdl_out, sbp_stats *stats) {
    if (!ctx || !bm || !g || !p || !epoch || p->batches < 1) return SBP_E_ARG;
    SBP_API_BEGIN
    bm_set_model(ctx, g, bm, p->nonparametric, p->nonparametric && p->degreecorrected);
    mcmc(ctx, bm, g, *p, golden_ratio_not_reached != 0, epoch, stats);
    if (mdl_out) *mdl_out = bm->overall_entropy;
    SBP_API_END(ctx)
}

int sbp_finetune_assignment(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params *p, uint32_t *epoch,
                            double *mdl_out, sbp_stats *stats) {
    if (!ctx || !bm || !g || !p || !epoch) return SBP_E_ARG;
    SBP_API_BEGIN
    bm_set_model(ctx, g, bm, p->nonparametric, p->nonparametric && p->degreecorrected);
    finetune_assignment(ctx, bm, g, *p, epoch, stats);
    if (mdl_out) *mdl_out = bm->overall_entropy;
    SBP_API_END(ctx)
}

int sbp_merge_propose(sbp_ctx ctx, sbp_bm bm, uint32_t epoch, int32_t *best_s, double *best_dS) {
    if (!ctx || !bm) return SBP_E_ARG;
    SBP_API_BEGIN
    SBP_CHECK(bm->B >= 2, SBP_E_ARG, "merge proposals need at least two blocks");
    const int32_t B = bm->B;
    int32_t *d_s = ctx->scratch[S_BEST].as<int32_t>(3 * (size_t)B + 4);
    double *d_dS = ctx->scratch[S_DS].as<double>((size_t)B + 2);
    merge_propose(ctx, bm, epoch, d_s, d_dS);
    if (best_s) SBP_HIP(dev_copy(best_s, d_s, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
    if (best_dS) SBP_HIP(dev_copy(best_dS, d_dS, sizeof(double) * B, hipMemcpyDeviceToHost));
    SBP_API_END(ctx)
}

int sbp_merge_blocks(sbp_ctx ctx, sbp_bm bm, sbp_graph g, uint32_t epoch, int64_t num_to_merge, int32_t approximate,
                     sbp_bm *out, sbp_stats *stats) {
    if (!ctx || !bm || !g || !out || num_to_merge < 0 || num_to_merge >= bm->B) return SBP_E_ARG;
    *out = nullptr;
    SBP_API_BEGIN
    *out = merge_blocks(ctx, bm, g, epoch, num_to_merge, approximate != 0, stats);
    SBP_API_END(ctx)
}

int sbp_merge_apply(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const int32_t *best_s, const double *best_dS,
                    int64_t num_to_merge, int32_t approximate, sbp_bm *out) {
    if (!ctx || !bm || !g || !best_s || !best_dS || !out || num_to_merge < 0 || num_to_merge >= bm->B) return SBP_E_ARG;
    *out = nullptr;
    SBP_API_BEGIN
    for (int32_t b = 0; b < bm->B; ++b)
        SBP_CHECK(best_s[b] >= 0 && best_s[b] < bm->B, SBP_E_ARG, "sbp_merge_apply: proposed block out of range");
    std::vector<int32_t> s(best_s, best_s + bm->B);
    std::vector<double> d(best_dS, best_dS + bm->B);
    *out = merge_apply(ctx, bm, g, s, d, num_to_merge, approximate != 0, nullptr);
    SBP_API_END(ctx)
}

int sbp_split_communities(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int32_t target, uint32_t epoch,
                          int32_t *num_blocks_out, double *delta_out) {
    if (!ctx || !bm || !g || !num_blocks_out) return SBP_E_ARG;
    SBP_API_BEGIN
    std::vector<double> delta;
    const int32_t B = bm->B;
    *num_blocks_out = split_communities(ctx, g, bm, target, epoch, delta_out ? &delta : nullptr);
    if (delta_out) std::copy(delta.begin(), delta.begin() + B, delta_out);
    SBP_API_END(ctx)
}

int sbp_ft_evaluate(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int64_t n, const int32_t *vertices, const int32_t *targets,
                    double *dS_out, double *hastings_out) {
    if (!ctx || !bm || !g || n < 0 || (n > 0 && (!vertices || !targets))) return SBP_E_ARG;
    SBP_API_BEGIN
    if (n == 0) return SBP_OK;
    for (int64_t i = 0; i < n; ++i)
        SBP_CHECK(vertices[i] >= 0 && vertices[i] < g->N && targets[i] >= 0 && targets[i] < bm->B, SBP_E_ARG,
                  "sbp_ft_evaluate: vertex or block out of range");
    int32_t *d_io = ctx->scratch[S_MISC].as<int32_t>(2 * (size_t)n);
    double *d_out = ctx->scratch[S_PARTIAL].as<double>(2 * (size_t)n);
    SBP_HIP(hipMemcpyAsync(d_io, vertices, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    SBP_HIP(hipMemcpyAsync(d_io + n, targets, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    ft_evaluate(ctx, bm, g, n, d_io, d_io + n, d_out, d_out + n);
    if (dS_out) SBP_HIP(hipMemcpyAsync(dS_out, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (hastings_out)
        SBP_HIP(hipMemcpyAsync(hastings_out, d_out + n, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    SBP_API_END(ctx)
}

int sbp_merge_evaluate(sbp_ctx ctx, sbp_bm bm, int64_t n, const int32_t *blocks, const int32_t *targets,
                       double *dS_out) {
    if (!ctx || !bm || n < 0 || (n > 0 && (!blocks || !targets || !dS_out))) return SBP_E_ARG;
    SBP_API_BEGIN
    if (n == 0) return SBP_OK;
    for (int64_t i = 0; i < n; ++i)
        SBP_CHECK(blocks[i] >= 0 && blocks[i] < bm->B && targets[i] >= 0 && targets[i] < bm->B && blocks[i] != targets[i],
                  SBP_E_ARG, "sbp_merge_evaluate: blocks out of range or equal");
    int32_t *d_io = ctx->scratch[S_MISC].as<int32_t>(2 * (size_t)n);
    double *d_out = ctx->scratch[S_PARTIAL].as<double>((size_t)n);
    SBP_HIP(hipMemcpyAsync(d_io, blocks, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    SBP_HIP(hipMemcpyAsync(d_io + n, targets, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    merge_evaluate(ctx, bm, n, d_io, d_io + n, d_out);
    SBP_HIP(hipMemcpyAsync(dS_out, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    SBP_API_END(ctx)
}

int sbp_bm_rebuild(sbp_ctx ctx, sbp_bm bm, sbp_graph g) {
    if (!ctx || !bm || !g || g->N != bm->N) return SBP_E_ARG;
    SBP_API_BEGIN
    bm_rebuild(ctx, g, bm);
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    SBP_API_END(ctx)
}

int sbp_bm_move_vertex(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int32_t vertex, int32_t block, int32_t *moved) {
    if (!ctx || !bm || !g || g->N != bm->N || vertex < 0 || vertex >= bm->N || block < 0 || block >= bm->B)
        return SBP_E_ARG;
    SBP_API_BEGIN
    int32_t r = 0, size_r = 0;
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    SBP_HIP(dev_copy(&r, bm->assign + vertex, sizeof(int32_t), hipMemcpyDeviceToHost));
    SBP_HIP(dev_copy(&size_r, bm->sizes + r, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (size_r == 1) {  // refused: the move would empty its block (src/blockmodel/blockmodel.cpp:439-443)
        if (moved) *moved = 0;
        return SBP_OK;
    }
    SBP_HIP(dev_copy(bm->assign + vertex, &block, sizeof(int32_t), hipMemcpyHostToDevice));
    bm_rebuild(ctx, g, bm);  // the counts and degree histograms of the new assignment, as the moves leave them
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    if (moved) *moved = 1;
    SBP_API_END(ctx)
}

int sbp_bm_merge_block(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int32_t from, int32_t to) {
    if (!ctx || !bm || !g || g->N != bm->N || from < 0 || from >= bm->B || to < 0 || to >= bm->B) return SBP_E_ARG;
    SBP_API_BEGIN
    std::vector<int32_t> map(bm->B);
    std::iota(map.begin(), map.end(), 0);
    map[from] = to;
    int32_t *d_map = ctx->scratch[S_BEST].as<int32_t>((size_t)bm->B + 4);
    SBP_HIP(hipMemcpyAsync(d_map, map.data(), sizeof(int32_t) * bm->B, hipMemcpyHostToDevice, ctx->stream));
    relabel_assignment(ctx, bm, d_map, bm->B, bm->assign);
    bm_rebuild(ctx, g, bm);
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    SBP_API_END(ctx)
}

int sbp_partition(sbp_ctx ctx, sbp_graph g, const sbp_params *p, int32_t *assignment_out, int32_t *num_blocks_out,
                  sbp_stats *stats) {
    if (!ctx || !g || !p || p->batches < 1 || p->algorithm < 0 || p->algorithm > 2) return SBP_E_ARG;
    SBP_API_BEGIN
    double t0 = now_ms();
    const uint64_t saved = ctx->seed;
    ctx->seed = p->seed;
    try {
        Blockmodel bm = sbp_amd::drive(ctx, g, g->N, *p, stats);
        SBP_HIP(hipStreamSynchronize(ctx->stream));
        if (assignment_out)
            SBP_HIP(dev_copy(assignment_out, bm.handle()->assign, sizeof(int32_t) * g->N, hipMemcpyDeviceToHost));
        if (num_blocks_out) *num_blocks_out = bm.handle()->B;
    } catch (...) {
        ctx->seed = saved;
        throw;
    }
    ctx->seed = saved;
    ctx->peak_bytes = std::max<uint64_t>(ctx->peak_bytes, ctx_device_bytes(ctx));
    if (getenv("SBP_AMD_TRACE")) {
        fprintf(stderr, "[sbp] device bytes: pool %zu (cached %zu), lgamma %lld, scratch", ctx->pool->total_bytes,
                ctx->pool->cached_bytes, (long long)ctx->lgamma_n * 8);
        for (int i = 0; i < S_NUM_SLOTS; ++i) fprintf(stderr, " %d:%zu", i, ctx->scratch[i].bytes);
        fprintf(stderr, "\n");
    }
    if (stats) stats->total_ms += now_ms() - t0;
    SBP_API_END(ctx)
}

int sbp_ctx_memory(sbp_ctx ctx, uint64_t *current_bytes, uint64_t *peak_bytes) {
    if (!ctx) return SBP_E_ARG;
    const uint64_t now = ctx_device_bytes(ctx);
    ctx->peak_bytes = std::max(ctx->peak_bytes, now);
    if (current_bytes) *current_bytes = now;
    if (peak_bytes) *peak_bytes = ctx->peak_bytes;
    return SBP_OK;
}

int sbp_ctx_trim(sbp_ctx ctx) {
    if (!ctx) return SBP_E_ARG;
    SBP_API_BEGIN
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->pool) ctx->pool->trim();
    for (auto &sc : ctx->scratch) {  // the next use allocates afresh under a new generation
        if (sc.ptr) SBP_HIP(dev_free(sc.ptr));
        sc.ptr = nullptr;
        sc.bytes = 0;
    }
    SBP_API_END(ctx)
}

int sbp_set_carry_threads(sbp_ctx ctx, int32_t threads) {
    if (!ctx || threads < 0 || threads > 64) return SBP_E_ARG;
    ctx->carry_threads = threads;
    return SBP_OK;
}

int sbp_set_finetune_shape(sbp_ctx ctx, const int32_t lanes[5], const int32_t waves_per_simd[4]) {
    if (!lanes || !waves_per_simd) return SBP_E_ARG;
    if (ctx) return set_finetune_shape(ctx, lanes, waves_per_simd) ? SBP_OK : SBP_E_ARG;
    sbp_ctx_s probe;
    if (!set_finetune_shape(&probe, lanes, waves_per_simd)) return SBP_E_ARG;
    std::lock_guard<std::mutex> lk(g_shape_mu);
    std::copy(lanes, lanes + 5, g_default_lanes);
    std::copy(waves_per_simd, waves_per_simd + 4, g_default_occ);
    return SBP_OK;
}

int sbp_set_recorded_sweeps(sbp_ctx ctx, int32_t mode) {
    if (!ctx || mode < -1 || mode > 1) return SBP_E_ARG;
    ctx->record_mode = mode;
    return SBP_OK;
}

int sbp_get_finetune_shape(sbp_ctx ctx, int32_t lanes[5], int32_t waves_per_simd[4]) {
    if (!lanes || !waves_per_simd) return SBP_E_ARG;
    if (ctx) {
        std::copy(ctx->ft_lanes, ctx->ft_lanes + 5, lanes);
        std::copy(ctx->ft_occ, ctx->ft_occ + 4, waves_per_simd);
    } else {
        std::lock_guard<std::mutex> lk(g_shape_mu);
        std::copy(g_default_lanes, g_default_lanes + 5, lanes);
        std::copy(g_default_occ, g_default_occ + 4, waves_per_simd);
    }
    return SBP_OK;
}

// Device buffers of one debug call, freed on return.
struct DebugBufs {
    std::vector<void *> p;
    void *get(size_t bytes) {
        void *x = nullptr;
        SBP_HIP(dev_malloc(&x, bytes > 0 ? bytes : 1));
        p.push_back(x);
        return x;
    }
    ~DebugBufs() {
        for (void *x : p) (void)dev_free(x);
    }
};

int sbp_debug_sort(sbp_ctx ctx, int64_t n, int32_t key_bits, const uint64_t *keys, const int32_t *vals,
                   int32_t begin_bit, int32_t end_bit, uint64_t *keys_out, int32_t *vals_out) {
    if (!ctx || n < 0 || (n > 0 && (!keys || !keys_out)) || (vals && !vals_out) || (key_bits != 32 && key_bits != 64) ||
        begin_bit < 0 || end_bit > key_bits || begin_bit > end_bit)
        return SBP_E_ARG;
    SBP_API_BEGIN
    hipStream_t st = ctx->stream;
    DebugBufs b;
    const size_t kb = (size_t)n * (key_bits / 8);
    void *dk = b.get(kb), *dko = b.get(kb);
    int32_t *dv = vals ? static_cast<int32_t *>(b.get(sizeof(int32_t) * n)) : nullptr;
    int32_t *dvo = vals ? static_cast<int32_t *>(b.get(sizeof(int32_t) * n)) : nullptr;
    std::vector<uint32_t> k32;
    if (key_bits == 32) {
        k32.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) k32[i] = (uint32_t)keys[i];
        SBP_HIP(dev_copy(dk, k32.data(), kb, hipMemcpyHostToDevice));
    } else if (n > 0) {
        SBP_HIP(dev_copy(dk, keys, kb, hipMemcpyHostToDevice));
    }
    if (vals && n > 0) SBP_HIP(dev_copy(dv, vals, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    size_t tmp = 0;
    if (key_bits == 64) {
        auto *a = static_cast<const unsigned long long *>(dk);
        auto *o = static_cast<unsigned long long *>(dko);
        SBP_HIP(sbpprim::radix_sort(nullptr, tmp, a, o, (const int32_t *)dv, dvo, n, begin_bit, end_bit, st));
        SBP_HIP(sbpprim::radix_sort(b.get(tmp), tmp, a, o, (const int32_t *)dv, dvo, n, begin_bit, end_bit, st));
    } else {
        auto *a = static_cast<const uint32_t *>(dk);
        auto *o = static_cast<uint32_t *>(dko);
        SBP_HIP(sbpprim::radix_sort(nullptr, tmp, a, o, (const int32_t *)dv, dvo, n, begin_bit, end_bit, st));
        SBP_HIP(sbpprim::radix_sort(b.get(tmp), tmp, a, o, (const int32_t *)dv, dvo, n, begin_bit, end_bit, st));
    }
    SBP_HIP(hipStreamSynchronize(st));
    if (key_bits == 32) {
        SBP_HIP(dev_copy(k32.data(), dko, kb, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) keys_out[i] = k32[i];
    } else if (n > 0) {
        SBP_HIP(dev_copy(keys_out, dko, kb, hipMemcpyDeviceToHost));
    }
    if (vals && n > 0) SBP_HIP(dev_copy(vals_out, dvo, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    SBP_API_END(ctx)
}

int sbp_debug_scan(sbp_ctx ctx, int64_t n, const int64_t *in, int32_t inclusive, int64_t *out) {
    if (!ctx || n < 0 || (n > 0 && (!in || !out))) return SBP_E_ARG;
    SBP_API_BEGIN
    hipStream_t st = ctx->stream;
    DebugBufs b;
    int64_t *di = static_cast<int64_t *>(b.get(sizeof(int64_t) * n));
    int64_t *dout = static_cast<int64_t *>(b.get(sizeof(int64_t) * n));
    if (n > 0) SBP_HIP(dev_copy(di, in, sizeof(int64_t) * n, hipMemcpyHostToDevice));
    size_t tmp = 0;
    SBP_HIP(sbpprim::scan(nullptr, tmp, (const int64_t *)di, dout, n, inclusive != 0, st));
    SBP_HIP(sbpprim::scan(b.get(tmp), tmp, (const int64_t *)di, dout, n, inclusive != 0, st));
    SBP_HIP(hipStreamSynchronize(st));
    if (n > 0) SBP_HIP(dev_copy(out, dout, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
    SBP_API_END(ctx)
}

int sbp_debug_rle(sbp_ctx ctx, int64_t n, const uint64_t *keys, uint64_t *uniq, int32_t *counts, int64_t *num_runs) {
    if (!ctx || n < 0 || n >= (int64_t(1) << 31) || !num_runs || (n > 0 && (!keys || !uniq || !counts)))
        return SBP_E_ARG;
    SBP_API_BEGIN
    hipStream_t st = ctx->stream;
    DebugBufs b;
    auto *dk = static_cast<unsigned long long *>(b.get(sizeof(uint64_t) * n));
    auto *du = static_cast<unsigned long long *>(b.get(sizeof(uint64_t) * n));
    int32_t *dc = static_cast<int32_t *>(b.get(sizeof(int32_t) * n));
    int64_t *dn = static_cast<int64_t *>(b.get(sizeof(int64_t)));
    if (n > 0) SBP_HIP(dev_copy(dk, keys, sizeof(uint64_t) * n, hipMemcpyHostToDevice));
    size_t tmp = 0;
    SBP_HIP(sbpprim::run_length_encode(nullptr, tmp, (const unsigned long long *)dk, du, dc, dn, n, st));
    SBP_HIP(sbpprim::run_length_encode(b.get(tmp), tmp, (const unsigned long long *)dk, du, dc, dn, n, st));
    SBP_HIP(hipStreamSynchronize(st));
    SBP_HIP(dev_copy(num_runs, dn, sizeof(int64_t), hipMemcpyDeviceToHost));
    if (*num_runs > 0) {
        SBP_HIP(dev_copy(uniq, du, sizeof(uint64_t) * *num_runs, hipMemcpyDeviceToHost));
        SBP_HIP(dev_copy(counts, dc, sizeof(int32_t) * *num_runs, hipMemcpyDeviceToHost));
    }
    SBP_API_END(ctx)
}

int sbp_debug_logf(sbp_ctx ctx, uint32_t first_bits, uint32_t count, float *out) {
    if (!ctx || !out) return SBP_E_ARG;
    SBP_API_BEGIN
    float *d = ctx->scratch[S_MISC].as<float>(count + 1);
    debug_logf(ctx, first_bits, count, d);
    SBP_HIP(hipMemcpyAsync(out, d, sizeof(float) * count, hipMemcpyDeviceToHost, ctx->stream));
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    SBP_API_END(ctx)
}

}  // extern "C"
