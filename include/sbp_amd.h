/*
 * sbp_amd.h -- C-ABI of the MI355X (gfx950) implementation of SBP's hot path.
 *
 * Plain pointers and sizes only.  Every call returns an int status
 * (SBP_OK = 0, negative on error); sbp_last_error() returns the message of the
 * last failing call on a context.  No C++ exception crosses this boundary.
 *
 * Ownership: the caller owns every handle and frees it; device memory belongs
 * to the handle.  A context owns one HIP stream; calls on one context are
 * serialised by the caller; separate contexts may run concurrently.
 *
 * Host arrays passed in are read before the call returns; host arrays passed
 * out are written before it returns.  Vertex ids are int32, graph offsets int64.
 *
 * Each entry point names the reference interface it replaces
 * (vtsynergy/SBP, file:line).  The random draws follow a replay contract:
 * Philox4x32-10 keyed by the context seed, counter {unit, epoch, tag, sub}.
 */
#ifndef SBP_AMD_H
#define SBP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBP_OK 0
#define SBP_E_HIP (-1)
#define SBP_E_OOM (-2)
#define SBP_E_ARG (-3)
#define SBP_E_NAN (-4)
#define SBP_E_STATE (-5)

typedef struct sbp_ctx_s *sbp_ctx;
typedef struct sbp_graph_s *sbp_graph;
typedef struct sbp_bm_s *sbp_bm;

/* Run parameters: the hot-path subset of the reference's Args (include/args.hpp:61-212). */
typedef struct {
    int32_t algorithm;        /* -a: 0 async_gibbs, 1 metropolis_hastings, 2 hybrid_mcmc (args.hpp:65-68) */
    int32_t batches;          /* --batches (args.hpp:75-78), default 1 */
    int32_t greedy;           /* --greedy (args.hpp:111-112), default 1 */
    int32_t approximate;      /* --approximate (args.hpp:69-71), default 0 */
    int32_t ordered;          /* --ordered (args.hpp:130), default 0 */
    int32_t asynciterations;  /* --asynciterations (args.hpp:72-74), default INT32_MAX */
    double mh_percent;        /* --mh_percent (args.hpp:115-116), default 0.075 */
    uint64_t seed;            /* replay seed (the reference seeds from random_device; src/rng.cpp:34-46) */
    int32_t max_iterations;   /* MAX_NUM_ITERATIONS (include/finetune.hpp:34), default 100 */
    int32_t threads;          /* --threads: accepted for flag parity; the device ignores it */
    int32_t nonparametric;    /* --nonparametric (args.hpp:126-127): entropy::nonparametric description length */
    int32_t degreecorrected;  /* --degreecorrected (args.hpp:87-88): with nonparametric, the degree-corrected terms */
} sbp_params;

/* Counters and phase timers (the reference's timers::*, include/globals.hpp:19-67). */
typedef struct {
    int64_t proposals;        /* vertex-move proposals evaluated */
    int64_t moves;            /* accepted and committed vertex moves */
    int64_t refused;          /* async-Gibbs moves refused at commit (source block of size 1) */
    int64_t sweeps;           /* finetune sweeps */
    int64_t merge_proposals;  /* block-merge proposals evaluated */
    int64_t merge_phases;
    double mdl;               /* final description length */
    double finetune_ms;       /* device time in finetune proposal + commit kernels */
    double merge_ms;          /* merge proposal kernels */
    double carry_ms;          /* merge carry-out */
    double build_ms;          /* blockmodel (re)builds */
    double mdl_ms;            /* MDL reductions */
    double total_ms;          /* wall time of the call */
    uint64_t ft_bytes;        /* algorithmic bytes of the finetune proposal launches on lane groups (vertices of
                               * total degree <= 256: degree classes 0-3) */
    uint64_t merge_bytes;     /* algorithmic bytes of the merge proposal kernels */
    int64_t ft_launches;      /* proposal batches with lane-group launches */
    double ft_kernel_ms;      /* device time of those lane-group launches */
    int64_t carry_evaluations;   /* merge carry-out: delta-MDL re-evaluations on the sequential path */
    int64_t carry_speculative;   /* merge carry-out: re-evaluations run ahead in parallel */
    int64_t carry_reused;        /* merge carry-out: pops answered by a still-exact earlier evaluation */
    int64_t carry_requeues;      /* merge carry-out: merges pushed back because they got worse */
    double carry_load_ms;        /* merge carry-out: device-to-host copy and hash-map build */
    uint64_t ft_high_bytes;      /* bytes, batches, device time of the workgroup-per-vertex finetune proposal
                                  * launches (total degree > 256: degree classes 4-5) */
    int64_t ft_high_launches;
    double ft_high_kernel_ms;
    uint64_t mh_bytes;           /* bytes, launches, device time of the Metropolis-Hastings kernel */
    int64_t mh_launches;
    double mh_kernel_ms;
    /* finetune proposal launches per degree class of k = |out(v)| + |in(v)| (k <= 8, 32, 128, 256, 1,024, more):
     * algorithmic bytes, launches, device time */
    uint64_t ft_class_bytes[6];
    int64_t ft_class_launches[6];
    double ft_class_ms[6];
    int64_t recorded_sweeps;     /* asynchronous-Gibbs sweeps run as replays of a recorded HIP graph (their
                                  * per-class launch times are not taken; finetune_ms holds the replays' time) */
    int64_t cell_updates;        /* asynchronous-Gibbs commits whose block matrix was updated from the moved
                                  * vertices' edges instead of rebuilt from every edge */
} sbp_stats;

/* ----------------------------------------------------------------- context */

/* Replaces rng::init_generators (src/rng.cpp:34-46): one context = one device stream + replay seed. */
int sbp_create(sbp_ctx *out, int device, uint64_t seed);
int sbp_destroy(sbp_ctx ctx);
const char *sbp_last_error(sbp_ctx ctx);
int sbp_set_seed(sbp_ctx ctx, uint64_t seed);
/* Blocks until all work queued on the context's stream is complete. */
int sbp_synchronize(sbp_ctx ctx);

/* ----------------------------------------------------------------- graph */

/* Replaces Graph(out_neighbors, in_neighbors, ...) (include/graph.hpp:15-140): CSR with each
 * vertex's neighbours in file order.  out_off/in_off have N+1 entries. */
int sbp_graph_upload(sbp_ctx ctx, int32_t N, int64_t E, const int64_t *out_off, const int32_t *out_adj,
                     const int64_t *in_off, const int32_t *in_adj, sbp_graph *out);
int sbp_graph_free(sbp_graph g);

int sbp_graph_info(sbp_graph g, int32_t *num_vertices, int64_t *num_edges);
/* Copies a graph's CSR to host (e.g. a sample's induced subgraph); any pointer may be NULL. */
int sbp_graph_download(sbp_ctx ctx, sbp_graph g, int64_t *out_off, int32_t *out_adj, int64_t *in_off,
                       int32_t *in_adj);

/* ----------------------------------------------------------------- sampling */

/* Replaces sample::sample (src/sample.cpp:352-365) with sample::from_vertices (:202-217): draws
 * (long)(fraction * N) vertices and builds their induced subgraph on the device.  alg: 0 max_degree
 * (:219-233), 1 random (:235-247), 2 degree_product (:28-49), 3 degree_weighted (new: weighted sampling
 * without replacement, weight = degree), 4 expansion_snowball (:94-155), 5 degree_product_snowball
 * (:51-92); the two snowball samplers grow the sample sequentially on the host, the subgraph is built on
 * the device.  The sample keeps its vertex mapping for sbp_extend.
 * mapping_out: host int32[N] (vertex -> sample vertex or -1), may be NULL. */
int sbp_sample(sbp_ctx ctx, sbp_graph g, int32_t alg, double fraction, uint64_t seed, sbp_graph *sample_out,
               int32_t *num_sampled, int32_t *mapping_out);
/* Replaces sample::detach (src/sample.cpp:12-26): the induced subgraph of the vertices of degree > 1
 * (Graph::degrees, |out| + |in| - self-loop flag) in id order.  mapping_out: host int32[N] (vertex ->
 * detached vertex or -1), may be NULL. */
int sbp_detach(sbp_ctx ctx, sbp_graph g, sbp_graph *detached_out, int32_t *num_kept, int32_t *mapping_out);
/* Replaces sample::reattach (src/sample.cpp:367-405) on the host (sequential by definition): the assignment
 * of every vertex of the full graph (host CSR) from the detached graph's assignment.  Vertices left out draw
 * random_integer(0, detached_blocks - 1) from the replay stream Philox(seed; j, 0, 0x52410001, 0), j counting
 * the draws, and take the blocks of their neighbours as the reference's loop does.  The caller builds the
 * blockmodel with sbp_bm_build(ctx, g, detached_blocks, assignment_out, ...). */
int sbp_reattach(int32_t N, const int64_t *out_off, const int32_t *out_adj, const int64_t *in_off,
                 const int32_t *in_adj, const int32_t *mapping, const int32_t *detached_assignment,
                 int32_t detached_blocks, uint64_t seed, int32_t *assignment_out);
/* Replaces sample::extend (src/sample.cpp:157-200) followed by Blockmodel(B, graph, 0.5, assignment)
 * (src/main.cpp:139): extends sample_bm, a blockmodel of `sample`, to g and builds the full-graph
 * blockmodel with the same B.  assignment_out: host int32[N], may be NULL. */
int sbp_extend(sbp_ctx ctx, sbp_graph g, sbp_graph sample, sbp_bm sample_bm, uint64_t seed, sbp_bm *out,
               int32_t *assignment_out);

/* ----------------------------------------------------------------- blockmodel */

/* Replaces Blockmodel(num_blocks, graph, rate, assignment) + initialize_edge_counts
 * (include/blockmodel/blockmodel.hpp:82-88; src/blockmodel/blockmodel.cpp:193-276).
 * assignment: host int32[N] with values in [0, B). */
int sbp_bm_build(sbp_ctx ctx, sbp_graph g, int32_t B, const int32_t *assignment, sbp_bm *out);
/* Replaces Blockmodel::copy (src/blockmodel/blockmodel.cpp:139-153). */
int sbp_bm_copy(sbp_ctx ctx, sbp_bm src, sbp_bm *out);
int sbp_bm_free(sbp_bm bm);
int sbp_bm_info(sbp_bm bm, int32_t *num_blocks, int64_t *nnz);
/* Replaces Blockmodel::initialize_edge_counts (src/blockmodel/blockmodel.cpp:193-276): recounts the block matrix,
 * degrees, sizes (and degree histograms) from the blockmodel's assignment. */
int sbp_bm_rebuild(sbp_ctx ctx, sbp_bm bm, sbp_graph g);
/* Replaces Blockmodel::move_vertex(const VertexMove_v3&) (src/blockmodel/blockmodel.cpp:437-494): moves `vertex` to
 * `block` unless its block has one vertex (then *moved = 0 and nothing changes). */
int sbp_bm_move_vertex(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int32_t vertex, int32_t block, int32_t *moved);
/* Replaces Blockmodel::merge_block (src/blockmodel/blockmodel.cpp:337-363): every vertex of block `from` moves to
 * `to`; the block count stays (block `from` is left empty). */
int sbp_bm_merge_block(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int32_t from, int32_t to);
/* Copies device state to host; any pointer may be NULL.  row_ptr has B+1 entries; cols/cnts nnz entries
 * (row-major, ascending columns).  Degrees are Blockmodel::degrees_out/_in/degrees. */
int sbp_bm_download(sbp_ctx ctx, sbp_bm bm, int32_t *assignment, int32_t *sizes, int32_t *d_out, int32_t *d_in,
                    int32_t *d_tot, int32_t *row_ptr, int32_t *cols, int32_t *cnts);
/* Replaces Blockmodel::validate (src/blockmodel/blockmodel.cpp:527-632): rebuilds from the assignment on
 * the device and compares every count; *ok = 1 when equal. */
int sbp_bm_validate(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int32_t *ok);

/* ----------------------------------------------------------------- entropy */

/* Replaces entropy::mdl (src/entropy.cpp:597-604; nonparametric::mdl for a nonparametric blockmodel, whose
 * log_posterior is then NaN) and Blockmodel::log_posterior_probability
 * (src/blockmodel/blockmodel.cpp:288-306). */
int sbp_mdl(sbp_ctx ctx, sbp_bm bm, sbp_graph g, double *mdl, double *log_posterior);

/* sbp_mdl for a parametric blockmodel without allocating device memory.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The reduction's work area is a fixed per-device array of 20,500 bytes (the partial sums and error
 * flags of at most 1,024 workgroups), set up by sbp_create; every blockmodel fits, since the reduction grid is capped at 1,024 workgroups.  Calls that overlap on
 * one device race on that array.  SBP_E_ARG for a nonparametric blockmodel. */
int sbp_mdl_noalloc(sbp_ctx ctx, sbp_bm bm, sbp_graph g, double *mdl, double *log_posterior);

/* Sets the blockmodel's entropy model, the reference's global args.nonparametric / args.degreecorrected:
 * 0 = entropy::mdl and its deltas, 1 = entropy::nonparametric (src/entropy.cpp:617-1225; directed), 2 =
 * entropy::nonparametric with degree correction, whose vertex-degree histograms are then built from the
 * current assignment and g (Blockmodel::initialize_edge_counts, src/blockmodel/blockmodel.cpp:234-235) and
 * kept up to date by the moves.  The calls that take sbp_params set it from p->nonparametric and
 * p->degreecorrected; blockmodels made by merges inherit it. */
int sbp_bm_set_model(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int32_t model);
/* Replaces entropy::nonparametric::mdl with its parts sparse_entropy, get_partition_dl, get_edges_dl and
 * get_deg_dl_dist (src/entropy.cpp:623-767); any output may be NULL. */
int sbp_np_description_length(sbp_ctx ctx, sbp_bm bm, sbp_graph g, double *mdl, double *sparse_entropy,
                              double *partition_dl, double *edges_dl, double *degree_dl);

/* ----------------------------------------------------------------- finetune */

/* Evaluates one proposal per vertex against the frozen blockmodel, as propose_gibbs_move_v3 +
 * eval_vertex_move_v3 do (src/finetune.cpp:447-462,871-901).  Host outputs, N entries each:
 * s_out = proposed block, dS_out = delta-MDL (0 when s == current block), accept_out = 0/1. */
int sbp_ft_propose_all(sbp_ctx ctx, sbp_bm bm, sbp_graph g, uint32_t epoch, int32_t greedy, int32_t *s_out,
                       double *dS_out, int32_t *accept_out);
/* Replaces entropy::delta_mdl and entropy::hastings_correction of given vertex moves (src/entropy.cpp:298-403,
 * 497-561) with the Delta of finetune::blockmodel_delta (src/finetune.cpp:201-235): for i < n, the change in
 * description length and the Hastings correction of moving vertices[i] to block targets[i] (0 and 1 when the
 * target is the vertex's block), against the current blockmodel; the proposal kernels' evaluation.
 * hastings_out may be NULL (it is 1 for a nonparametric blockmodel, src/entropy.cpp:499). */
int sbp_ft_evaluate(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int64_t n, const int32_t *vertices, const int32_t *targets,
                    double *dS_out, double *hastings_out);
/* One sweep of finetune::asynchronous_gibbs (src/finetune.cpp:50-86): proposals against the frozen
 * blockmodel per batch, commit in vertex-id order with refusal of moves that would empty a block
 * (src/blockmodel/blockmodel.cpp:437-494). */
int sbp_async_gibbs_sweep(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params *p, uint32_t epoch,
                          sbp_stats *stats);
/* The same sweep followed by the MDL that closes an iteration of finetune::asynchronous_gibbs (src/finetune.cpp:
 * 50-88), run as a replay of a HIP graph recorded on the first call for this blockmodel and parameters (recorded
 * again when they change): one graph launch and one read per sweep.  Results are those of sbp_async_gibbs_sweep
 * then sbp_mdl.  Parametric blockmodels only (SBP_E_ARG otherwise).  The partition driver runs its
 * asynchronous-Gibbs finetune this way unless SBP_AMD_RECORDED_SWEEPS=0. */
int sbp_async_gibbs_sweep_recorded(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params *p, uint32_t epoch,
                                   double *mdl_out, sbp_stats *stats);
/* One sweep of sequential Metropolis-Hastings (src/finetune.cpp:773-820; propose_move :842-869).
 * ordered = 1 visits vertices in id order (finetune_assignment); sum_dS = sum of accepted delta-MDL. */
int sbp_mh_sweep(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params *p, uint32_t epoch, int32_t ordered,
                 double *sum_dS, sbp_stats *stats);
/* Replaces finetune::mcmc (src/finetune.cpp:754-771) incl. early stopping; *epoch is advanced by the
 * number of sweeps.  *mdl_out receives Blockmodel::getOverall_entropy after the phase. */
int sbp_mcmc(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params *p, int32_t golden_ratio_not_reached,
             uint32_t *epoch, double *mdl_out, sbp_stats *stats);
/* Replaces finetune::finetune_assignment (src/finetune.cpp:903-932). */
int sbp_finetune_assignment(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params *p, uint32_t *epoch,
                            double *mdl_out, sbp_stats *stats);

/* ----------------------------------------------------------------- block merge */

/* Replaces the proposal loop of block_merge::merge_blocks (src/block_merge.cpp:189-217): 10 proposals
 * per block, best (lowest delta-MDL) per block.  Host outputs of B entries. */
int sbp_merge_propose(sbp_ctx ctx, sbp_bm bm, uint32_t epoch, int32_t *best_s, double *best_dS);
/* Replaces entropy::block_merge_delta_mdl (src/entropy.cpp:129-182) with the Delta of block_merge::blockmodel_delta
 * (src/block_merge.cpp:14-38): for i < n, the change in description length of merging blocks[i] into targets[i]
 * (distinct blocks). */
int sbp_merge_evaluate(sbp_ctx ctx, sbp_bm bm, int64_t n, const int32_t *blocks, const int32_t *targets,
                       double *dS_out);
/* Replaces the carry-out half of block_merge::merge_blocks (src/block_merge.cpp:218-227): carry_out_best_merges_advanced
 * (:40-117), or Blockmodel::carry_out_best_merges (src/blockmodel/blockmodel.cpp:55-122) when approximate, of given
 * proposals best_s / best_dS (B entries), then label compaction and the rebuild: a new blockmodel with
 * B - num_to_merge blocks in *out. */
int sbp_merge_apply(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const int32_t *best_s, const double *best_dS,
                    int64_t num_to_merge, int32_t approximate, sbp_bm *out);
/* Replaces block_merge::merge_blocks (src/block_merge.cpp:178-228): proposals, carry-out
 * (carry_out_best_merges_advanced, or Blockmodel::carry_out_best_merges when approximate), label
 * compaction and rebuild.  Writes a new blockmodel with B - num_to_merge blocks to *out. */
int sbp_merge_blocks(sbp_ctx ctx, sbp_bm bm, sbp_graph g, uint32_t epoch, int64_t num_to_merge,
                     int32_t approximate, sbp_bm *out, sbp_stats *stats);

/* ----------------------------------------------------------------- divisive step */

/* Replaces divisive::split_communities (src/top_down_sbp.cpp:488-549) with random splits (propose_random_split,
 * :228-240, the default --split) and apply_best_splits (:30-66), in place: blocks of at least 0.001 N vertices
 * each draw 10 random two-way splits, keep the one of lowest nonparametric::mdl(two-block model of the block's
 * induced subgraph) - null_mdl_v1(subgraph), and the best splits are applied in ascending order of that value
 * until the block count reaches target (at least one is applied when any block kept one); the blockmodel is then
 * rebuilt (initialize_edge_counts).  null_mdl_v1 is the nonparametric one for a nonparametric blockmodel
 * (args.nonparametric), the closed form otherwise.  Draws are keyed by (seed; block or vertex, epoch), so a
 * divisive driver passes a new epoch per call.  *num_blocks_out: the new block count.  delta_out (may be NULL):
 * host double[B] of the old block count, each block's kept value (DBL_MAX for a block not split).
 * SBP_E_ARG for a degree-corrected blockmodel. */
int sbp_split_communities(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int32_t target, uint32_t epoch,
                          int32_t *num_blocks_out, double *delta_out);

/* ----------------------------------------------------------------- driver */

/* Replaces sbp::stochastic_block_partition (src/sbp.cpp:16-72) with the golden-section triplet
 * (src/blockmodel/blockmodel_triplet.cpp:4-105).  assignment_out: host int32[N]. */
int sbp_partition(sbp_ctx ctx, sbp_graph g, const sbp_params *p, int32_t *assignment_out,
                  int32_t *num_blocks_out, sbp_stats *stats);

/* ----------------------------------------------------------------- many graphs on one device
 *
 * No counterpart in the reference, which partitions one graph per process (src/main.cpp:60-177).  Each graph
 * runs the driver above (sbp_partition) on a worker context of its own (stream, buffer pool, pinned staging;
 * merge carry-out on one host thread), one host thread per worker; workers overlap on the device.  A graph's
 * result (assignment, block count, description length) is the one sbp_partition gives it alone. */

/* Device bytes a context holds (buffer pool, scratch slots, lgamma table; graphs not counted), now and the
 * largest seen at the end of an sbp_partition call. */
int sbp_ctx_memory(sbp_ctx ctx, uint64_t *current_bytes, uint64_t *peak_bytes);
/* Frees the buffers the context's pool keeps cached for reuse and its scratch slots. */
int sbp_ctx_trim(sbp_ctx ctx);
/* Host threads of the context's merge carry-out (0: SBP_AMD_CARRY_THREADS, else up to 16). */
int sbp_set_carry_threads(sbp_ctx ctx, int32_t threads);
/* Launch shape of the finetune proposal launches, per degree class of k = |out(v)| + |in(v)| (b = 0..4 for
 * k <= 8, 32, 128, 256 and 1,024; the rest take 512-thread workgroups).  lanes[b]: lanes (threads) per vertex,
 * allowed 2, 4, 8, 16, 32 or 64 for class 0, 4, 8, 16, 32 or 64 for class 1, 16, 32 or 64 for class 2, 32 or 64
 * for class 3, 128 or 256 for class 4.
 * waves_per_simd[b] (b = 0..3): the wavefronts per SIMD that the greedy parametric lane-group launch's register
 * budget is sized for, 3 or 4.  SBP_E_ARG (nothing changed) otherwise.  With ctx NULL, sets the default of the
 * contexts created afterwards (lanes 2, 4, 16, 64, 256 at 4 wavefronts per SIMD).  Results do not depend on it.
 * No counterpart in the reference: a tuning knob of the launch shape, which the autotuner
 * (sbp_amd/autotune.py) sets. */
int sbp_set_finetune_shape(sbp_ctx ctx, const int32_t lanes[5], const int32_t waves_per_simd[4]);
int sbp_get_finetune_shape(sbp_ctx ctx, int32_t lanes[5], int32_t waves_per_simd[4]);
/* Whether the partition driver (sbp_partition, sbp_mcmc, batches) runs asynchronous-Gibbs sweeps as replays of a
 * recorded graph: -1 for graphs of at least 100,000 vertices (default; SBP_AMD_RECORDED_SWEEPS=0 or 1 overrides
 * it), 0 never, 1 always.  Batch workers take the calling context's setting.  Results do not depend on it.  No
 * counterpart in the reference. */
int sbp_set_recorded_sweeps(sbp_ctx ctx, int32_t mode);
/* Device bytes of a graph with N vertices and E edges, and an upper bound of the device bytes one partition
 * run of it needs (graph, blockmodels, scratch), given the largest |out(v)| + |in(v)|, the number of vertices
 * where that exceeds 256 and the algorithm (sbp_params.algorithm): the server admits requests by the latter. */
uint64_t sbp_graph_bytes(int32_t N, int64_t E);
uint64_t sbp_footprint_bytes(int32_t N, int64_t E, int64_t max_degree, int64_t n_high_degree, int32_t algorithm);

/* Partitions n device graphs (uploaded on any context of ctx's device) with per-graph parameters params[n]
 * (each its own seed) over `slots` workers (0: SBP_AMD_SLOTS, else min(12, host threads)), largest graphs
 * first.  The workers are contexts owned by ctx, kept for later calls.  Optional outputs per graph:
 * assignments[i] (host int32[N_i]), num_blocks[i], stats[i], finish_ms[i] (milliseconds from the start of the
 * call to the graph's completion).  Returns the first failing graph's status, its index in the message. */
int sbp_partition_batch(sbp_ctx ctx, const sbp_graph *graphs, int32_t n, const sbp_params *params, int32_t slots,
                        int32_t *const *assignments, int32_t *num_blocks, sbp_stats *stats, double *finish_ms);

/* Request server: `slots` workers (0 as above) take submitted partition requests in arrival order, each as soon
 * as a worker is free and its footprint (sbp_footprint_bytes) fits in the budget left by the requests in
 * flight; each worker also keeps up to min(4 GiB, budget / (4 slots)) cached between requests, which is set
 * aside from the budget.  A request uploads its graph on its worker's stream, partitions it and frees it. */
typedef struct sbp_server_s *sbp_server;
int sbp_server_create(int device, int32_t slots, uint64_t hbm_budget_bytes, sbp_server *out);
const char *sbp_server_last_error(sbp_server srv);
/* Queues a request.  The host CSR arrays are read when the request starts and must stay valid until
 * sbp_server_wait returns for it.  SBP_E_OOM when the footprint exceeds the budget. */
int sbp_server_submit(sbp_server srv, int32_t N, int64_t E, const int64_t *out_off, const int32_t *out_adj,
                      const int64_t *in_off, const int32_t *in_adj, const sbp_params *p, int64_t *ticket);
/* Blocks until the request is done and returns its status and results: latency_ms from submission to
 * completion, queue_ms from submission to start.  A ticket is waited for once. */
int sbp_server_wait(sbp_server srv, int64_t ticket, int32_t *assignment_out, int32_t *num_blocks_out,
                    sbp_stats *stats, double *latency_ms, double *queue_ms);
int sbp_server_info(sbp_server srv, int32_t *running, int32_t *queued, int32_t *max_running, uint64_t *budget_bytes,
                    uint64_t *in_use_bytes, uint64_t *peak_in_use_bytes);
/* Lets running requests finish, ends queued ones with SBP_E_STATE, joins the workers.  Call it after every
 * sbp_server_wait has returned. */
int sbp_server_destroy(sbp_server srv);

/* ----------------------------------------------------------------- evaluation (host only) */

/* Scores of a partition against a true one (evaluate::Eval and the values calculate_f1_score prints,
 * src/evaluate.cpp:9-96).  The reference's names are kept: recall = same-pair agreement over the pairs
 * within the table's rows (the algorithm's communities), precision over the pairs within its columns. */
typedef struct {
    double f1_score, nmi, accuracy, rand_index, precision, recall;
    int32_t found_blocks, true_blocks;  /* table rows, columns */
} sbp_evaluation;
/* Replaces calculate_f1_score + calculate_nmi (src/evaluate.cpp:9-96) on a contingency table (row-major
 * int32[nrows * ncols], rows = the algorithm's communities); accuracy = diagonal / num_vertices. */
int sbp_evaluate_table(int64_t num_vertices, int32_t nrows, int32_t ncols, const int32_t *table, sbp_evaluation *out);
/* Replaces evaluate::hungarian + the two above (src/evaluate.cpp:134-210): builds the table from the true
 * assignment (-1 = unknown, left out) and the algorithm's (block ids >= 0), matches communities by maximum
 * weight (not above 4096 communities: then accuracy is NaN) and scores it.  Optional outputs: the matched
 * table, row-major int32[found_blocks * true_blocks] (rows = the algorithm's communities), and its shape. */
int sbp_evaluate(int32_t N, const int32_t *true_assignment, const int32_t *assignment, sbp_evaluation *out,
                 int32_t *table_out, int32_t *table_rows, int32_t *table_cols);

/* ----------------------------------------------------------------- diagnostics */

/* Device logf over float bit patterns [first_bits, first_bits + count) (checks the restated glibc logf). */
int sbp_debug_logf(sbp_ctx ctx, uint32_t first_bits, uint32_t count, float *out);
/* The library's device primitives on host arrays (checks them; no counterpart in the reference): the stable radix
 * sort of n 64-bit keys (key_bits 64) or 32-bit keys (key_bits 32, in the low words of keys) by bits
 * [begin_bit, end_bit), carrying vals when given (NULL: keys only); the int64 prefix sum (inclusive or not);
 * run-length encoding of n sorted 64-bit keys (uniq and counts hold n entries; *num_runs set). */
int sbp_debug_sort(sbp_ctx ctx, int64_t n, int32_t key_bits, const uint64_t *keys, const int32_t *vals,
                   int32_t begin_bit, int32_t end_bit, uint64_t *keys_out, int32_t *vals_out);
int sbp_debug_scan(sbp_ctx ctx, int64_t n, const int64_t *in, int32_t inclusive, int64_t *out);
int sbp_debug_rle(sbp_ctx ctx, int64_t n, const uint64_t *keys, uint64_t *uniq, int32_t *counts, int64_t *num_runs);

#ifdef __cplusplus
}
#endif
#endif /* SBP_AMD_H */
