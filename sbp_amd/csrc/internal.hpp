// Host-side handle layouts and helpers shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <initializer_list>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/sbp_amd.h"
#include "device_common.hpp"
#include "np_entropy.hpp"

namespace sbpcarry {
class Pool;
}
namespace sbpimpl {
struct RecordedSweep;
}

namespace sbpimpl {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define SBP_HIP(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t _e = (expr);                                                                           \
        if (_e != hipSuccess)                                                                             \
            throw ::sbpimpl::Error(SBP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e) + " at " + \
                                                  __FILE__ + ":" + std::to_string(__LINE__));             \
    } while (0)

#define SBP_CHECK(cond, code, msg)                                  \
    do {                                                            \
        if (!(cond)) throw ::sbpimpl::Error((code), std::string(msg)); \
    } while (0)

// Set on the thread that records a graph (stream capture): growing a buffer then would synchronize and allocate
// inside the capture, so every buffer the recorded work needs is reserved beforehand and growth is an error.
inline thread_local bool g_recording = false;

// hipFree and hipDeviceSynchronize wait for every stream of the device, a stream being recorded by another thread
// included, which ends that recording with an error.  They take this lock shared; a recording takes it exclusively
// (record_sweep), so none runs while a graph is being recorded.
inline std::shared_mutex &device_wide_mu() {
    static std::shared_mutex m;
    return m;
}
inline hipError_t dev_free(void *p) {
    if (g_recording) return hipErrorStreamCaptureUnsupported;
    std::shared_lock<std::shared_mutex> lk(device_wide_mu());
    return hipFree(p);
}
inline hipError_t dev_host_free(void *p) {
    if (g_recording) return hipErrorStreamCaptureUnsupported;
    std::shared_lock<std::shared_mutex> lk(device_wide_mu());
    return hipHostFree(p);
}
// Synchronous copies and allocations likewise: they are outside any stream's order.
inline hipError_t dev_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    if (g_recording) return hipErrorStreamCaptureUnsupported;
    std::shared_lock<std::shared_mutex> lk(device_wide_mu());
    return hipMemcpy(dst, src, bytes, kind);
}
template <typename T>
inline hipError_t dev_malloc(T **p, size_t bytes) {
    if (g_recording) return hipErrorStreamCaptureUnsupported;
    std::shared_lock<std::shared_mutex> lk(device_wide_mu());
    return hipMalloc(reinterpret_cast<void **>(p), bytes);
}
inline hipError_t dev_sync() {
    if (g_recording) return hipErrorStreamCaptureUnsupported;
    std::shared_lock<std::shared_mutex> lk(device_wide_mu());
    return hipDeviceSynchronize();
}

// Device buffer that only grows; contents are not preserved across growth.
struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
    uint64_t generation = 0;  // incremented whenever the buffer is replaced
    void *get(size_t need) {
        if (need > bytes) {
            if (g_recording) throw Error(SBP_E_STATE, "device scratch growth while recording a graph");
            ++generation;
            // queued kernels may still read the old buffer
            if (ptr) {
                SBP_HIP(dev_sync());
                SBP_HIP(dev_free(ptr));
            }
            ptr = nullptr;
            size_t n = need + need / 4 + 256;
            SBP_HIP(dev_malloc(&ptr, n));
            bytes = n;
        }
        return ptr;
    }
    template <typename T>
    T *as(size_t count) {
        return static_cast<T *>(get(count * sizeof(T) + 16));
    }
    ~Scratch() {
        if (ptr) (void)dev_free(ptr);
    }
};

// Recycles the device buffers of blockmodels.  Blockmodels are created, copied and dropped in every merge
// phase; hipMalloc and hipFree cost tens of microseconds each and hipFree waits for the whole device.  A
// released buffer is kept and handed out again for a request of at most its size and more than half of it.
// Reuse is ordered by the context's stream, on which every write to a blockmodel buffer is issued.  The
// context and each of its blockmodels share the pool, so it outlives the context while blockmodels remain.
struct DevicePool {
    static constexpr size_t MAX_CACHED = size_t(16) << 30;
    std::multimap<size_t, void *> cached;
    std::unordered_map<void *, size_t> size_of;
    size_t cached_bytes = 0;
    size_t total_bytes = 0;  // every buffer the pool owns, handed out or cached
    void *alloc(size_t bytes) {
        if (g_recording) throw Error(SBP_E_STATE, "device allocation while recording a graph");
        bytes = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
        auto it = cached.lower_bound(bytes);
        if (it != cached.end() && it->first < 2 * bytes) {
            void *p = it->second;
            cached_bytes -= it->first;
            cached.erase(it);
            return p;
        }
        void *p = nullptr;
        if (dev_malloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            trim();
            SBP_HIP(dev_malloc(&p, bytes));
        }
        size_of[p] = bytes;
        total_bytes += bytes;
        return p;
    }
    void release(void *p) {
        if (!p) return;
        auto it = size_of.find(p);
        if (it == size_of.end()) {
            (void)dev_free(p);
            return;
        }
        cached.emplace(it->second, p);
        cached_bytes += it->second;
        if (cached_bytes > MAX_CACHED) trim();
    }
    void trim() {
        if (cached.empty()) return;
        (void)dev_sync();
        for (auto &kv : cached) {
            size_of.erase(kv.second);
            (void)dev_free(kv.second);
        }
        total_bytes -= cached_bytes;
        cached.clear();
        cached_bytes = 0;
    }
    ~DevicePool() { trim(); }
};

enum ScratchSlot {
    S_KEYS_A, S_KEYS_B, S_VALS_A, S_VALS_B, S_UNIQ, S_COUNTS, S_SORT_TMP, S_MISC, S_PARTIAL,
    S_MOVES, S_DS, S_ACCEPT, S_COUNTER, S_HASH, S_RISK, S_EVENTS_A, S_EVENTS_B, S_ORDER, S_BEST,
    S_ASSIGN_TMP, S_FLAGS, S_NP, S_BUCKETS, S_PREV, S_UPD_CNT, S_UPD_MOVED, S_UPD_KEYS, S_UPD_SORTED, S_UPD_UNIQ,
    S_UPD_CNTS, S_UPD_NET, S_UPD_RKEYS, S_UPD_CKEYS, S_UPD_RCNT, S_UPD_CCNT, S_UPD_POS, S_UPD_AT, S_UPD_REM,
    S_NUM_SLOTS
};

}  // namespace sbpimpl

struct sbp_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t seed = 0;
    std::string err;
    sbpimpl::Scratch scratch[sbpimpl::S_NUM_SLOTS];
    std::shared_ptr<sbpimpl::DevicePool> pool = std::make_shared<sbpimpl::DevicePool>();
    hipEvent_t ev[8] = {};
    hipEvent_t ft_ev[8] = {};  // boundaries of the finetune proposal launches, one per degree class
    void *pinned = nullptr;  // pinned staging area for small reads (sbpimpl::PINNED_BYTES)
    void *stage = nullptr;   // pinned staging of the block matrix for the host carry-out, grown on demand
    size_t stage_bytes = 0;
    double *lgamma = nullptr;  // fastlgamma table for the nonparametric entropy, grown on demand
    int64_t lgamma_n = 0;
    // the global-memory neighbour-block tables in scratch slot S_HASH that are known to be empty (every kernel
    // that uses them leaves them empty): buffer, table geometry (log2 capacity) and count
    uint64_t hash_clean_gen = 0;  // generation of slot S_HASH the state below refers to (0 = none)
    int hash_clean_lg = -1;
    int64_t hash_clean_tables = 0;
    int carry_threads = 0;    // host threads of the merge carry-out (0: SBP_AMD_CARRY_THREADS or up to 16)
    // a server worker's count of the requests running now: each merge phase's carry-out then takes its share of
    // the host's carry threads (at least one) instead of carry_threads
    const std::atomic<int32_t> *carry_share = nullptr;
    std::shared_ptr<sbpcarry::Pool> carry_pool;  // the carry-out's threads, kept between merge phases
    uint64_t peak_bytes = 0;  // largest device footprint seen at the end of a partition (sbp_ctx_memory)
    std::vector<sbp_ctx> slots;  // worker contexts of sbp_partition_batch, created on first use
    std::shared_ptr<sbpimpl::RecordedSweep> recorded;  // the asynchronous-Gibbs sweep recorded as a graph (capi.hip)
    int64_t cell_updates = 0;  // commits whose cells were updated from the moves (bm_update_moves)
    int32_t record_mode = -1;  // sbp_set_recorded_sweeps: -1 by size (or SBP_AMD_RECORDED_SWEEPS), 0 off, 1 on
    // launch shape of the finetune degree classes 0..4 (k <= 8, 32, 128, 256, 1,024): lanes (threads) per vertex,
    // and wavefronts per SIMD of the greedy parametric lane-group launches (sbp_set_finetune_shape)
    int32_t ft_lanes[5] = {2, 4, 16, 64, 256};
    int32_t ft_occ[4] = {4, 4, 4, 4};
};

namespace sbpimpl {

// Device bytes a context holds: its buffer pool (blockmodels, handed out or cached), scratch slots and the
// lgamma table.  Graphs are allocated apart and not counted.
inline uint64_t ctx_device_bytes(sbp_ctx ctx) {
    uint64_t b = ctx->pool ? ctx->pool->total_bytes : 0;
    for (const Scratch &s : ctx->scratch) b += s.bytes;
    return b + sizeof(double) * (uint64_t)ctx->lgamma_n;
}

constexpr size_t PINNED_BYTES = 4096;
struct ReadPiece {
    void *dst;
    const void *src;
    size_t bytes;
};
// Small device-to-host reads (counters, flags, totals) through the context's pinned staging area: the pieces
// are copied on the context's stream, then one wait.  A pageable destination would make each copy a staged,
// blocking transfer of its own.
inline void read_small(sbp_ctx ctx, std::initializer_list<ReadPiece> pieces) {
    char *stage = static_cast<char *>(ctx->pinned);
    size_t off = 0;
    for (const ReadPiece &p : pieces) {
        SBP_CHECK(off + p.bytes <= PINNED_BYTES, SBP_E_STATE, "read_small: staging area exceeded");
        SBP_HIP(hipMemcpyAsync(stage + off, p.src, p.bytes, hipMemcpyDeviceToHost, ctx->stream));
        off += (p.bytes + 15) & ~size_t(15);
    }
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    off = 0;
    for (const ReadPiece &p : pieces) {
        memcpy(p.dst, stage + off, p.bytes);
        off += (p.bytes + 15) & ~size_t(15);
    }
}

}  // namespace sbpimpl

struct sbp_graph_s {
    sbp_ctx ctx = nullptr;
    int32_t N = 0;
    int64_t E = 0;
    int64_t *out_off = nullptr, *in_off = nullptr;
    int32_t *out_adj = nullptr, *in_adj = nullptr;
    int32_t *edge_src = nullptr;    // [E + N] source vertex of each out-edge (CSR order), then 0..N-1
    int32_t *degree = nullptr;      // [N] |out| + |in| - self-loop flag (Graph::degrees)
    int32_t *deg_order = nullptr;   // [N] stable descending degree order, computed on first use; published
                                    // under deg_mu once complete (workers on other streams may share the graph)
    std::mutex deg_mu;
    // every vertex grouped by finetune degree class (ascending id within a class), built on first use and published
    // under deg_mu once complete: FT_NB counts then FT_NB offsets into ft_blist (device), the counts on the host
    int32_t *ft_blist = nullptr;
    int32_t *ft_bmeta = nullptr;
    int64_t ft_bcount[6] = {};
    int64_t max_degree = 0;         // max over v of |out(v)| + |in(v)|
    int64_t n_high_degree = 0;      // vertices with |out(v)| + |in(v)| > FT_SMALL_DEGREE
    // set when this graph is a sample of `parent` (sbp_sample)
    sbp_graph parent = nullptr;
    int32_t parent_N = 0;
    int32_t *parent_map = nullptr;       // [parent_N] parent vertex -> sample vertex or -1
    int32_t *sample_vertices = nullptr;  // [N] sample vertex -> parent vertex
};

// Vertices up to this total degree keep their neighbour-block histogram in LDS during finetune proposals and run
// on lane groups of a wavefront; the others run one workgroup per vertex.
constexpr int FT_SMALL_DEGREE = 256;
// Finetune degree classes: four lane-group classes (k <= 8, 32, 128, 256) and two workgroup classes (k <= 1,024,
// and the rest).
constexpr int FT_NB = 6;


struct sbp_bm_s {
    sbp_ctx ctx = nullptr;
    std::shared_ptr<sbpimpl::DevicePool> pool;  // the context's, holding every buffer below
    int32_t B = 0, N = 0;
    int64_t nnz = 0, cap = 0;
    int32_t *assign = nullptr;   // [N]
    int32_t *sizes = nullptr;    // [B]
    int32_t *d_out = nullptr;    // [B]
    int32_t *d_in = nullptr;     // [B]
    int32_t *d_tot = nullptr;    // [B]
    int32_t *row_ptr = nullptr;  // [B+1]
    int32_t *col_ptr = nullptr;  // [B+1]
    int2 *row_cells = nullptr;   // [cap] {col, count}, ascending col per row
    int2 *col_cells = nullptr;   // [cap] {row, count}, ascending row per column
    int32_t *row_cum = nullptr;  // [cap] inclusive cumulative counts within each row
    int32_t *col_cum = nullptr;  // [cap] inclusive cumulative counts within each column, diagonal excluded
    int32_t bcap = 0;            // capacity of the per-block arrays
    double overall_entropy = 1.7976931348623157e308;
    int64_t num_to_merge = 0;
    int32_t np = 0;  // entropy model: 0 parametric (entropy::mdl), 1 nonparametric (entropy::nonparametric)
    int32_t dc = 0;  // nonparametric with degree correction: the degree histograms below are kept
    sbpdev::HistSlot *hist = nullptr;  // dc: (block, direction, vertex degree) -> vertex count
    int hist_log2 = 0;
    bool hist_ready = false;  // dc: built from the current assignment (and kept up to date by moves)
    int64_t E = 0;   // edges of the graph the blockmodel was built on
    // cell hash for O(1) lookups (built on demand; any change to the cells invalidates it)
    sbpdev::CellSlot *cell_hash = nullptr;
    int hash_log2 = 0;
    bool hash_valid = false;
    int64_t hash_used = 0;  // claimed slots of the hash
    // dense copy of the block matrix (B x B counts, row-major) for lookups when B is small; kept by the in-place
    // cell update, dropped with the hash by every other change to the cells
    int32_t *dense = nullptr;
    int32_t dense_B = 0;
    bool dense_valid = false;
    // sorted cell keys (row << keys_cbits | col and col << keys_cbits | row) of the current cells, kept by eager
    // rebuilds for the update after moves
    unsigned long long *row_keys = nullptr, *col_keys = nullptr;  // [cap]
    int keys_cbits = 0;
    bool keys_valid = false;
};

namespace sbpimpl {

// build.hip
void bm_alloc(sbp_bm bm, int32_t B, int32_t N);
int bits_for_blocks(int64_t B);
void bm_fill_cells(sbp_ctx ctx, sbp_bm bm, const unsigned long long *row_keys, const int32_t *row_counts,
                   const unsigned long long *col_keys, const int32_t *col_counts);
void bm_count_sizes(sbp_ctx ctx, sbp_bm bm);
// the dense block matrix for lookups when it is wanted (B small, large graph or any_graph; SBP_AMD_DENSE=0/1
// turns it off/on for any B up to the bound), built on demand; nullptr when not wanted
const int32_t *bm_ensure_dense(sbp_ctx ctx, sbp_bm bm, sbp_graph g, bool any_graph = false);
// update.hip: the cells after committed moves (prev: the assignment before them); false when the rebuild must run
bool bm_update_moves(sbp_ctx ctx, sbp_graph g, sbp_bm bm, const int32_t *prev, int64_t moved_hint);
bool bm_update_wanted(sbp_graph g);  // graph large enough for the update (or SBP_AMD_INCREMENTAL forcing it)
void bm_release(sbp_bm bm);
// From bm->assign (device), bm->B set.  The degree histograms of a degree-corrected blockmodel are rebuilt too,
// unless keep_hist (a finetune commit, which updates them move by move as the reference does).
// capture: no host reads and no allocation (the cell arrays must hold the largest block matrix the graph allows,
// bm_reserve_cells); bm->nnz is then stale until bm_sync_nnz.
void bm_rebuild(sbp_ctx ctx, sbp_graph g, sbp_bm bm, bool keep_hist = false, bool capture = false);
// Cell capacity for any block matrix of B blocks over g (nnz <= min(E, B^2)), for captured rebuilds.
void bm_reserve_cells(sbp_ctx ctx, sbp_graph g, sbp_bm bm);
// bm->nnz from the device (after captured rebuilds).
void bm_sync_nnz(sbp_ctx ctx, sbp_bm bm);
void bm_copy_into(sbp_ctx ctx, sbp_bm src, sbp_bm dst);
void bm_compute_cum(sbp_ctx ctx, sbp_bm bm);
// Degree histograms (initialize_edge_counts, src/blockmodel/blockmodel.cpp:234-235) from the assignment and g.
void bm_build_hist(sbp_ctx ctx, sbp_graph g, sbp_bm bm);
// After a batch of moves: fails on a full table, compacts it when half of its slots are claimed (claimed and
// full are the table's bookkeeping slot, read with the batch's other counters).
void bm_hist_after_moves(sbp_ctx ctx, sbp_bm bm, int32_t claimed, int32_t full);
// Sets the entropy model (args.nonparametric / args.degreecorrected), building the histograms when needed.
void bm_set_model(sbp_ctx ctx, sbp_graph g, sbp_bm bm, int np, int dc);
// capture: rebuild the cell hash in place (its table must hold 2 x the largest nnz: bm_reserve_cells), with no
// host reads; it is then valid only for the graph.
void bm_ensure_hash(sbp_ctx ctx, sbp_bm bm, bool capture = false);
void bm_ensure_cells(sbp_bm bm, int64_t nnz);
// mdl.hip
double log_posterior(sbp_ctx ctx, sbp_bm bm);
// The same reduction with nnz read on the device (row_ptr[B]) and the total (fixed point) and error flag left in
// device memory at out_total / out_err: no host reads, for graph capture.  Work area in slot S_PARTIAL, sized for
// max_nnz cells.
void log_posterior_device(sbp_ctx ctx, sbp_bm bm, int64_t max_nnz, sbpdev::Fx *out_total, int32_t *out_err);
void log_posterior_reserve(sbp_ctx ctx, int64_t max_nnz);
// The same reduction with its work area in a fixed per-device array (sbp_mdl_noalloc): allocates nothing.
double log_posterior_noalloc(sbp_ctx ctx, sbp_bm bm);
// Looks up the fixed array's address on the current device (once per device); sbp_create calls it.
void noalloc_setup(int device);
double model_term(int64_t E, int32_t N, int32_t B);
double host_lgamma(int64_t x);
double host_lbinom(int64_t n, int64_t k);
double host_edges_dl(int64_t B, int64_t E);
int64_t count_nonempty(sbp_ctx ctx, sbp_bm bm);
sbpdev::NpConst np_const(sbp_ctx ctx, sbp_bm bm);
double np_mdl(sbp_ctx ctx, sbp_graph g, sbp_bm bm, double *parts);
// split.hip
int32_t split_communities(sbp_ctx ctx, sbp_graph g, sbp_bm bm, int32_t target, uint32_t epoch,
                          std::vector<double> *delta_out);
// finetune.hip
struct FtLaunchInfo {
    uint64_t bytes;       // algorithmic bytes of the lane-group launches (vertices up to FT_SMALL_DEGREE)
    uint64_t high_bytes;  // ... of the workgroup-per-vertex launches
    int64_t proposals;
    float kernel_ms;      // device time of the lane-group launches
    float high_kernel_ms; // device time of the workgroup-per-vertex launches (0 when not launched)
    bool high_launched;
    uint64_t class_bytes[FT_NB];  // per degree class
    float class_ms[FT_NB];
    bool class_launched[FT_NB];
};
// epoch_ptr: the launches read the epoch there and are recorded for a graph (no events, no host reads; the
// counters stay on the device, ft_propose_counters).
FtLaunchInfo ft_propose(sbp_ctx ctx, sbp_bm bm, sbp_graph g, uint32_t epoch, int greedy, const int32_t *d_vertices,
                        int64_t count, int32_t *d_move_s, double *d_dS, int32_t *d_s_out,
                        const uint32_t *epoch_ptr = nullptr);
const unsigned long long *ft_propose_counters(sbp_ctx ctx);
// Reserves every buffer a recorded sweep's proposal and commit launches use (batches of up to batch_max items).
void ft_reserve(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int64_t batch_max);
// Commit and rebuild; capture: no host reads (counters left on the device, ft_commit_counters), for graph capture.
int64_t ft_commit(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int32_t *d_move_s, int64_t *refused_out, bool capture = false);
const unsigned long long *ft_commit_counters(sbp_ctx ctx);
// delta-MDL and Hastings correction of moving d_vertices[i] to d_targets[i] (device arrays of n), against bm.
void ft_evaluate(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int64_t n, const int32_t *d_vertices, const int32_t *d_targets,
                 double *d_dS, double *d_H);
void shuffle_keys(sbp_ctx ctx, const int32_t *d_items, int64_t n, uint32_t epoch, int32_t *d_out,
                  const uint32_t *epoch_ptr = nullptr);
// Sets the launch shape of the degree classes (lanes or threads per vertex of classes 0-4, wavefronts per SIMD of the
// greedy parametric lane-group launches); false (nothing set) when one is not allowed.
bool set_finetune_shape(sbp_ctx ctx, const int32_t *lanes, const int32_t *waves_per_simd);
// graph.hip
void graph_finish(sbp_ctx ctx, sbp_graph g);
const int32_t *degree_order(sbp_ctx ctx, sbp_graph g);
// alg: sbp_sample's 0..5, or SAMPLE_DETACH (sample::detach; fraction and seed unused)
constexpr int32_t SAMPLE_DETACH = 6;
sbp_graph sample_graph(sbp_ctx ctx, sbp_graph g, int32_t alg, double fraction, uint64_t seed);
void extend_assignment(sbp_ctx ctx, sbp_graph g, sbp_graph sample, sbp_bm sample_bm, uint64_t seed, int32_t *d_assign);
// merge.hip
uint64_t merge_propose(sbp_ctx ctx, sbp_bm bm, uint32_t epoch, int32_t *d_best_s, double *d_best_dS);
// delta-MDL of merging d_blocks[i] into d_targets[i] (device arrays of n), against bm.
void merge_evaluate(sbp_ctx ctx, sbp_bm bm, int64_t n, const int32_t *d_blocks, const int32_t *d_targets, double *d_dS);
void relabel_assignment(sbp_ctx ctx, sbp_bm bm, const int32_t *d_block_map, int32_t new_B, int32_t *d_out);
void debug_logf(sbp_ctx ctx, uint32_t first, uint32_t count, float *d_out);

double elapsed_ms(sbp_ctx ctx, hipEvent_t a, hipEvent_t b);

}  // namespace sbpimpl
