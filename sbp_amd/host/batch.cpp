// Many graphs on one device: a batch call (sbp_partition_batch) and a request server (sbp_server_*) with
// continuous admission.  The reference partitions one graph per process (src/main.cpp:60-177); neither entry
// point has a counterpart there.
//
// Both run each graph through the single-graph driver (sbp_partition) on a worker context of its own: a HIP
// stream, a device buffer pool and pinned staging area per worker, driven by one host thread.  While one
// worker waits on a device-to-host read or runs its merge carry-out on the host, the others keep the device
// busy, and the kernels of small graphs, which fill few of the 256 CUs, overlap on the device.  A graph's
// result depends only on the graph and its parameters (its own seed, replay counters keyed by its own vertex
// and block ids, no reduction across graphs), so it is the same alone, in a batch or among other requests.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../csrc/internal.hpp"

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

constexpr int32_t MAX_SLOTS = 64;

// Default number of worker slots: the host threads that drive the device (one per worker) and run the
// carry-out; SBP_AMD_SLOTS (1..64) overrides it.
int32_t default_slots() {
    if (const char *e = getenv("SBP_AMD_SLOTS")) {
        const int v = atoi(e);
        if (v >= 1 && v <= MAX_SLOTS) return v;
    }
    const unsigned hw = std::thread::hardware_concurrency();
    return (int32_t)std::max(1u, std::min(12u, hw ? hw : 1u));
}

int make_worker(int device, sbp_ctx *out) {
    int rc = sbp_create(out, device, 0);
    if (rc != SBP_OK) return rc;
    (*out)->carry_threads = 1;  // the workers themselves are the host parallelism
    return SBP_OK;
}

}  // namespace

extern "C" {

uint64_t sbp_graph_bytes(int32_t N, int64_t E) {
    // out/in offsets, out/in adjacency, per-edge sources (E + N), degrees and the degree order
    return 16 * ((uint64_t)N + 1) + 8 * ((uint64_t)E + 1) + 4 * ((uint64_t)E + N) + 8 * ((uint64_t)N + 1);
}

uint64_t sbp_footprint_bytes(int32_t N, int64_t E, int64_t max_degree, int64_t n_high_degree, int32_t algorithm) {
    // Upper bound of the device bytes of one partition run: the graph, a worker's buffer pool and its scratch
    // slots (each grown to 5/4 of its largest request).
    const uint64_t n = (uint64_t)std::max<int32_t>(N, 0), e = (uint64_t)std::max<int64_t>(E, 0);
    // A blockmodel with nnz <= E cells takes 40 nnz (row and column cells with their cumulative counts and sorted
    // keys) at capacity 9/8 nnz, 28 N (assignment and six per-block arrays, B <= N) and a cell hash of up to
    // 96 nnz (16-byte slots, 3 to 6 per cell).  The pool holds at most six at once with the released buffers it
    // keeps for reuse: the triplet's three, the current one, the merged one and a copy.
    uint64_t pool = 6 * (141 * e + 28 * n + 4096);
    // Graphs of at least 2^22 edges, and every Metropolis-Hastings walk (alone or in hybrid_mcmc), also keep a
    // dense copy of each blockmodel with at most 8,192 blocks: up to min(N, 8192)^2 counts for each of the six.
    if (e >= (uint64_t(1) << 22) || algorithm == 1 || algorithm == 2) {
        const uint64_t b = std::min<uint64_t>(n, 8192);
        pool += 6 * 4 * b * b;
    }
    // Rebuild sort keys and values, move records and per-block arrays; the update after moves (at most E / 2
    // signed deltas: their keys, runs and nets, merged key and count lists of up to 3/2 E cells, ranks).
    uint64_t scratch = 72 * e + 128 * n + (uint64_t(32) << 20) + (68 * e + 8 * n) * 5 / 4;
    // Per-wave neighbour-block tables of the finetune proposals of vertices with |out| + |in| > 256: one per
    // such vertex (rounded up to a block of 4), 16 bytes per slot at 2^lg >= 2 min(max degree, N) + 2 slots,
    // at most 1 GiB; Metropolis-Hastings takes one.
    if (max_degree > 256) {
        int lg = 10;
        while ((int64_t(1) << lg) < 2 * std::min<int64_t>(max_degree, (int64_t)n) + 2) ++lg;
        const uint64_t table = 4 * ((uint64_t(4) << lg) + 1);
        uint64_t tables = (uint64_t)std::max<int64_t>(n_high_degree, 1);
        tables = std::min<uint64_t>(tables, std::max<uint64_t>(1, (uint64_t(1) << 30) / table));
        tables = (tables + 3) / 4 * 4;
        scratch += tables * table * 5 / 4 + 4096;
    }
    // Metropolis-Hastings (alone or in hybrid_mcmc) lays rows and columns out with slack: each capacity is
    // len + max(8, len / 2), and only a row (column) that runs short gets more, at least doubling its extra
    // room and covering the request, so the extra room stays within twice the cells the sweep adds (at most
    // 4 per edge): two int2 arrays of at most 10 E + 32 N entries each.
    if (algorithm == 1 || algorithm == 2) scratch += 2 * 8 * (10 * e + 32 * n) * 5 / 4;
    return sbp_graph_bytes(N, E) + pool + scratch;
}

int sbp_partition_batch(sbp_ctx ctx, const sbp_graph *graphs, int32_t n, const sbp_params *params, int32_t slots,
                        int32_t *const *assignments, int32_t *num_blocks, sbp_stats *stats, double *finish_ms) {
    if (!ctx || n < 0 || (n > 0 && (!graphs || !params)) || slots < 0 || slots > MAX_SLOTS) return SBP_E_ARG;
    for (int32_t i = 0; i < n; ++i)
        if (!graphs[i] || params[i].batches < 1 || params[i].algorithm < 0 || params[i].algorithm > 2)
            return SBP_E_ARG;
    if (n == 0) return SBP_OK;
    if (slots == 0) slots = default_slots();
    slots = std::min(slots, n);
    try {
        while ((int32_t)ctx->slots.size() < slots) {
            sbp_ctx w = nullptr;
            const int rc = make_worker(ctx->device, &w);
            if (rc != SBP_OK) {
                ctx->err = "sbp_partition_batch: cannot create a worker context";
                return rc;
            }
            ctx->slots.push_back(w);
        }
        for (sbp_ctx w : ctx->slots) {
            std::copy(ctx->ft_lanes, ctx->ft_lanes + 5, w->ft_lanes);
            std::copy(ctx->ft_occ, ctx->ft_occ + 4, w->ft_occ);
            w->record_mode = ctx->record_mode;
        }
        SBP_HIP(hipStreamSynchronize(ctx->stream));  // graphs uploaded on ctx are complete
        // longest first (by edges, then vertices), which bounds the last worker's tail
        std::vector<int32_t> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            if (graphs[a]->E != graphs[b]->E) return graphs[a]->E > graphs[b]->E;
            return graphs[a]->N > graphs[b]->N;
        });
        std::atomic<int32_t> next{0}, done{0};
        // SBP_AMD_BATCH_PROGRESS set: one line on stderr each time another 1/32 of the graphs has finished
        const bool progress = getenv("SBP_AMD_BATCH_PROGRESS") != nullptr;
        const int32_t step = std::max<int32_t>(1, n / 32);
        std::mutex mu;
        int first_rc = SBP_OK;
        int32_t first_bad = -1;
        std::string first_err;
        const double t0 = now_ms();
        auto work = [&](sbp_ctx w) {
            if (hipSetDevice(w->device) != hipSuccess) {
                std::lock_guard<std::mutex> lk(mu);
                if (first_rc == SBP_OK) first_rc = SBP_E_HIP, first_err = "hipSetDevice failed in a worker";
                return;
            }
            for (int32_t k; (k = next.fetch_add(1)) < n;) {
                const int32_t i = order[k];
                sbp_stats local{};
                int32_t nb = 0;
                const int rc = sbp_partition(w, graphs[i], &params[i], assignments ? assignments[i] : nullptr, &nb,
                                             stats ? &stats[i] : &local);
                if (num_blocks) num_blocks[i] = nb;
                if (finish_ms) finish_ms[i] = now_ms() - t0;
                const int32_t d = done.fetch_add(1) + 1;
                if (progress && (d % step == 0 || d == n))
                    fprintf(stderr, "[sbp] batch: %d of %d graphs done at %.1f s\n", d, n, (now_ms() - t0) / 1e3);
                if (rc != SBP_OK) {
                    std::lock_guard<std::mutex> lk(mu);
                    if (first_rc == SBP_OK || i < first_bad) first_rc = rc, first_bad = i, first_err = w->err;
                }
            }
        };
        std::vector<std::thread> threads;
        threads.reserve(slots);
        for (int32_t s = 0; s < slots; ++s) threads.emplace_back(work, ctx->slots[s]);
        for (auto &t : threads) t.join();
        if (first_rc != SBP_OK) {
            ctx->err = "sbp_partition_batch: graph " + std::to_string(first_bad) + ": " + first_err;
            return first_rc;
        }
        return SBP_OK;
    } catch (const sbpimpl::Error &e) {
        ctx->err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        ctx->err = "host out of memory";
        return SBP_E_OOM;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return SBP_E_STATE;
    }
}

}  // extern "C"

// ------------------------------------------------------------------ request server

namespace {

enum ReqState { QUEUED, RUNNING, DONE };

struct Request {
    int32_t N = 0;
    int64_t E = 0;
    const int64_t *out_off = nullptr, *in_off = nullptr;
    const int32_t *out_adj = nullptr, *in_adj = nullptr;
    sbp_params p{};
    uint64_t footprint = 0;
    ReqState state = QUEUED;
    bool claimed = false;  // a sbp_server_wait call owns the result
    double t_submit = 0, t_start = 0, t_done = 0;
    std::vector<int32_t> assignment;
    int32_t num_blocks = 0;
    sbp_stats stats{};
    int rc = SBP_OK;
    std::string err;
};

}  // namespace

struct sbp_server_s {
    int device = 0;
    uint64_t budget = 0;       // bytes the requests in flight may reserve
    uint64_t worker_cache = 0;  // bytes each worker may keep cached between requests (reserved up front)
    std::vector<sbp_ctx> workers;
    std::vector<std::thread> threads;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;  // queued tickets, oldest first
    std::unordered_map<int64_t, Request> reqs;
    int64_t next_ticket = 0;
    uint64_t in_use = 0, peak_in_use = 0;
    int32_t running = 0, max_running = 0;
    std::atomic<int32_t> active{0};  // running, read by the workers' carry-outs without the lock
    bool stop = false;
    std::string err;

    void work(sbp_ctx w) {
        const bool dev_ok = hipSetDevice(device) == hipSuccess;
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            // strict arrival order: the oldest request starts as soon as a worker is free and its footprint
            // fits in what the requests in flight leave of the budget
            cv_work.wait(lk, [&] {
                return stop || (!queue.empty() && in_use + reqs[queue.front()].footprint <= budget);
            });
            if (stop) return;
            const int64_t t = queue.front();
            queue.pop_front();
            Request &r = reqs[t];
            r.state = RUNNING;
            r.t_start = now_ms();
            in_use += r.footprint;
            peak_in_use = std::max(peak_in_use, in_use);
            max_running = std::max(max_running, ++running);
            active.store(running, std::memory_order_relaxed);
            Request job = r;  // the fields read below; the map may rehash while unlocked
            lk.unlock();
            int rc = dev_ok ? SBP_OK : SBP_E_HIP;
            std::string msg = dev_ok ? "" : "hipSetDevice failed in a worker";
            std::vector<int32_t> assign(job.N);
            int32_t nb = 0;
            sbp_stats st{};
            sbp_graph g = nullptr;
            if (rc == SBP_OK)
                rc = sbp_graph_upload(w, job.N, job.E, job.out_off, job.out_adj, job.in_off, job.in_adj, &g);
            if (rc == SBP_OK) rc = sbp_partition(w, g, &job.p, assign.data(), &nb, &st);
            if (rc != SBP_OK && msg.empty()) msg = w->err;
            if (g) sbp_graph_free(g);
            if (ctx_cached(w) > worker_cache) sbp_ctx_trim(w);
            lk.lock();
            Request &d = reqs[t];
            d.assignment.swap(assign);
            d.num_blocks = nb;
            d.stats = st;
            d.rc = rc;
            d.err = msg;
            d.t_done = now_ms();
            d.state = DONE;
            in_use -= d.footprint;
            --running;
            active.store(running, std::memory_order_relaxed);
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
    static uint64_t ctx_cached(sbp_ctx w) { return sbpimpl::ctx_device_bytes(w); }
};

extern "C" {

int sbp_server_create(int device, int32_t slots, uint64_t hbm_budget_bytes, sbp_server *out) {
    if (!out || slots < 0 || slots > MAX_SLOTS) return SBP_E_ARG;
    *out = nullptr;
    if (slots == 0) slots = default_slots();
    auto *s = new (std::nothrow) sbp_server_s();
    if (!s) return SBP_E_OOM;
    s->device = device;
    // each worker keeps up to budget / (4 slots), at most 4 GiB, of pooled buffers and scratch between requests;
    // that much is set aside from the budget, the rest is what requests in flight may reserve
    s->worker_cache = std::min<uint64_t>(uint64_t(4) << 30, hbm_budget_bytes / (4 * (uint64_t)slots));
    const uint64_t reserved = s->worker_cache * (uint64_t)slots;
    s->budget = hbm_budget_bytes > reserved ? hbm_budget_bytes - reserved : 0;
    for (int32_t i = 0; i < slots; ++i) {
        sbp_ctx w = nullptr;
        const int rc = make_worker(device, &w);
        if (rc != SBP_OK) {
            for (sbp_ctx x : s->workers) sbp_destroy(x);
            delete s;
            return rc;
        }
        w->carry_share = &s->active;  // a request's carry-outs share the host with the others running
        s->workers.push_back(w);
    }
    for (sbp_ctx w : s->workers) s->threads.emplace_back([s, w] { s->work(w); });
    *out = s;
    return SBP_OK;
}

const char *sbp_server_last_error(sbp_server srv) { return srv ? srv->err.c_str() : "null server"; }

int sbp_server_submit(sbp_server srv, int32_t N, int64_t E, const int64_t *out_off, const int32_t *out_adj,
                      const int64_t *in_off, const int32_t *in_adj, const sbp_params *p, int64_t *ticket) {
    if (!srv || !ticket || N <= 0 || E < 0 || !out_off || !in_off || (E > 0 && (!out_adj || !in_adj)) || !p ||
        p->batches < 1 || p->algorithm < 0 || p->algorithm > 2)
        return SBP_E_ARG;
    int64_t max_degree = 0, n_high = 0;
    for (int32_t v = 0; v < N; ++v) {
        const int64_t k = (out_off[v + 1] - out_off[v]) + (in_off[v + 1] - in_off[v]);
        max_degree = std::max(max_degree, k);
        n_high += k > 256;
    }
    const uint64_t fp = sbp_footprint_bytes(N, E, max_degree, n_high, p->algorithm);
    std::lock_guard<std::mutex> lk(srv->mu);
    if (fp > srv->budget) {
        srv->err = "sbp_server_submit: request footprint " + std::to_string(fp) + " bytes exceeds the budget of " +
                   std::to_string(srv->budget) + " bytes";
        return SBP_E_OOM;
    }
    const int64_t t = srv->next_ticket++;
    Request &r = srv->reqs[t];
    r.N = N;
    r.E = E;
    r.out_off = out_off;
    r.out_adj = out_adj;
    r.in_off = in_off;
    r.in_adj = in_adj;
    r.p = *p;
    r.footprint = fp;
    r.t_submit = now_ms();
    srv->queue.push_back(t);
    srv->cv_work.notify_all();
    *ticket = t;
    return SBP_OK;
}

int sbp_server_wait(sbp_server srv, int64_t ticket, int32_t *assignment_out, int32_t *num_blocks_out,
                    sbp_stats *stats, double *latency_ms, double *queue_ms) {
    if (!srv) return SBP_E_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) {
        srv->err = "sbp_server_wait: unknown ticket";
        return SBP_E_ARG;
    }
    if (it->second.claimed) {  // a ticket is waited for once: a second waiter would find it erased
        srv->err = "sbp_server_wait: ticket " + std::to_string(ticket) + " is already being waited for";
        return SBP_E_ARG;
    }
    it->second.claimed = true;
    // std::unordered_map keeps references to its elements valid across inserts of other keys, and only this
    // waiter erases the entry
    Request &r = it->second;
    srv->cv_done.wait(lk, [&] { return r.state == DONE; });
    const int rc = r.rc;
    if (rc == SBP_OK) {
        if (assignment_out) memcpy(assignment_out, r.assignment.data(), sizeof(int32_t) * r.N);
        if (num_blocks_out) *num_blocks_out = r.num_blocks;
        if (stats) *stats = r.stats;
    } else {
        srv->err = "request " + std::to_string(ticket) + ": " + r.err;
    }
    if (latency_ms) *latency_ms = r.t_done - r.t_submit;
    if (queue_ms) *queue_ms = r.t_start - r.t_submit;
    srv->reqs.erase(ticket);
    return rc;
}

int sbp_server_info(sbp_server srv, int32_t *running, int32_t *queued, int32_t *max_running, uint64_t *budget_bytes,
                    uint64_t *in_use_bytes, uint64_t *peak_in_use_bytes) {
    if (!srv) return SBP_E_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    if (running) *running = srv->running;
    if (queued) *queued = (int32_t)srv->queue.size();
    if (max_running) *max_running = srv->max_running;
    if (budget_bytes) *budget_bytes = srv->budget;
    if (in_use_bytes) *in_use_bytes = srv->in_use;
    if (peak_in_use_bytes) *peak_in_use_bytes = srv->peak_in_use;
    return SBP_OK;
}

int sbp_server_destroy(sbp_server srv) {
    if (!srv) return SBP_OK;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        // requests already running finish; queued ones end with SBP_E_STATE
        for (int64_t t : srv->queue) {
            Request &r = srv->reqs[t];
            r.rc = SBP_E_STATE;
            r.err = "server destroyed before the request started";
            r.t_start = r.t_done = now_ms();
            r.state = DONE;
        }
        srv->queue.clear();
        srv->stop = true;
        srv->cv_work.notify_all();
        srv->cv_done.notify_all();
    }
    for (auto &t : srv->threads) t.join();
    for (sbp_ctx w : srv->workers) sbp_destroy(w);
    delete srv;
    return SBP_OK;
}

}  // extern "C"
