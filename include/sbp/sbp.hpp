// C++ drop-in surface of the SBP hot path, implemented over the C-ABI in include/sbp_amd.h.
//
// Mirrors the reference's public functions and classes (vtsynergy/SBP):
//   sbp::stochastic_block_partition / done_blockmodeling      include/sbp.hpp:18-21
//   Blockmodel                                                include/blockmodel/blockmodel.hpp:46-219
//   BlockmodelTriplet                                         include/blockmodel/blockmodel_triplet.hpp:12-55
//   block_merge::merge_blocks                                 include/block_merge.hpp:48
//   finetune::mcmc / asynchronous_gibbs / metropolis_hastings /
//            hybrid_mcmc / finetune_assignment                include/finetune.hpp:38-123
//   entropy::mdl / null_mdl_v1 / normalize_mdl_v1             include/entropy.hpp:71-80
//   sample::sample / extend                                   include/sample.hpp:53-71
// Vertex and block ids are `long` at this surface, as in the reference; the device stores int32.
// Errors from the device path are thrown as std::runtime_error (std::invalid_argument for bad input).
#ifndef SBP_AMD_SBP_HPP
#define SBP_AMD_SBP_HPP

#include <memory>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../sbp_amd.h"
#include "args.hpp"

// ------------------------------------------------------------------ the reference's shared types and globals

/// include/typedefs.hpp: MapVector (a robin_map there; ids in any order), EdgeWeights (:16-19), Vertex
/// (:114-118), VertexMove_v3 (:140-147).
template <typename T>
using MapVector = std::unordered_map<long, T>;
struct EdgeWeights {
    std::vector<long> indices;
    std::vector<long> values;
};
struct Vertex {
    long id;
    long out_degree;
    long in_degree;
};
struct VertexMove_v3 {
    double delta_entropy;
    bool did_move;
    Vertex vertex;
    long proposed_block;
    EdgeWeights out_edges;
    EdgeWeights in_edges;
};

/// include/globals.hpp:17: the arguments every entry point below reads (set them, or pass them to
/// sbp::stochastic_block_partition, which stores them here).
extern Args args;
/// include/mpi_data.hpp: the single rank of this build (no MPI).
struct MPI_t {
    int rank = 0;
    int num_processes = 1;
};
extern MPI_t mpi;

/// include/globals.hpp:19-67: phase timers (seconds) and counters, filled by the partition driver, the sampled
/// flow's callers and finetune::finetune_assignment.
namespace timers {
extern double BlockMerge_time, BlockMerge_loop_time, BLOCKMODEL_BUILD_TIME, MCMC_time, MCMC_parallel_time, total_time;
extern double finetune_time, sample_time, sample_extend_time, sample_finetune_time;
extern long MCMC_iterations, total_num_islands;
extern unsigned long MCMC_moves;
}  // namespace timers

/// src/rng.cpp:34-46: creates the device context for args.device and keys the replay stream with args.seed.
namespace rng {
void init_generators();
}

namespace sbp_amd {
/// Process-wide device context (one HIP stream), created on first use from the Args' device and seed.
sbp_ctx context(const Args &args);
void throw_on(int status, sbp_ctx ctx, const char *what);
}  // namespace sbp_amd

/// Directed graph with per-vertex neighbour lists in file order (include/graph.hpp:15-140).
class Graph {
  public:
    Graph() = default;
    Graph(long num_vertices, std::vector<int64_t> out_off, std::vector<int32_t> out_adj, std::vector<int64_t> in_off,
          std::vector<int32_t> in_adj);
    /// Graph::load (src/graph.cpp:33-72): <filepath>.tsv (1-indexed "from to" lines) or .mtx, plus
    /// <filepath>_truePartition.tsv when present.
    static Graph load(const Args &args);
    /// Graph::load() (src/main.cpp:82): Graph::load(args) with the global args.
    static Graph load();
    /// Graph::num_islands (src/graph.cpp:293-300): vertices of degree 0.
    long num_islands() const;
    /// Builds from an edge list in file order (0-indexed).
    static Graph from_edges(long num_vertices, const std::vector<std::pair<long, long>> &edges);
    long num_vertices() const { return N_; }
    long num_edges() const { return (long)out_adj_.size(); }
    std::vector<long> out_neighbors(long v) const;
    std::vector<long> in_neighbors(long v) const;
    /// Graph::degrees (src/graph.cpp:24-31)
    std::vector<long> degrees() const;
    const std::vector<long> &assignment() const { return truth_; }
    void assignment(std::vector<long> a) { truth_ = std::move(a); }
    /// Device copy on the context (uploaded on first use).
    sbp_graph device(sbp_ctx ctx) const;
    /// Takes ownership of a device graph (e.g. a sample) and downloads its CSR.
    static Graph adopt(sbp_ctx ctx, sbp_graph g, std::shared_ptr<struct sbp_graph_s> keep_alive = nullptr);
    const std::vector<int64_t> &out_off() const { return out_off_; }
    const std::vector<int32_t> &out_adj() const { return out_adj_; }
    const std::vector<int64_t> &in_off() const { return in_off_; }
    const std::vector<int32_t> &in_adj() const { return in_adj_; }

  private:
    long N_ = 0;
    std::vector<int64_t> out_off_{0}, in_off_{0};
    std::vector<int32_t> out_adj_, in_adj_;
    std::vector<long> truth_;
    mutable std::shared_ptr<struct sbp_graph_s> dev_;
    std::shared_ptr<struct sbp_graph_s> parent_;  // a sample's device graph refers to its parent's
};

/// Changed cells of a move (include/blockmodel/sparse/delta.hpp:22-149): rows current_block / proposed_block, then
/// columns (routing priority row == current, row == proposed, col == current, col == proposed); zero-net entries
/// stay.  A Delta made by finetune::blockmodel_delta or block_merge::blockmodel_delta also records its move, which
/// entropy::delta_mdl / block_merge_delta_mdl / hastings_correction evaluate on the device.
class Delta {
  public:
    Delta() = default;
    Delta(long current_block, long proposed_block, long buckets = 10);
    void add(long row, long col, long value);
    void sub(long row, long col, long value) { add(row, col, -value); }
    long get(long row, long col) const;
    std::vector<std::tuple<long, long, long>> entries() const;
    long current_block() const { return current_; }
    long proposed_block() const { return proposed_; }
    long self_edge_weight() const { return self_; }
    void self_edge_weight(long w) { self_ = w; }
    long vertex = -1;    // the moved vertex (finetune::blockmodel_delta), or -1
    bool merge = false;  // a block merge of current_block into proposed_block (block_merge::blockmodel_delta)

  private:
    long current_ = -1, proposed_ = -1, self_ = 0;
    MapVector<long> cur_row_, prop_row_, cur_col_, prop_col_;
};

namespace utils {
/// include/utils.hpp:25-30
typedef struct proposal_and_edge_counts_t {
    long proposal;
    long num_out_neighbor_edges;
    long num_in_neighbor_edges;
    long num_neighbor_edges;
} ProposalAndEdgeCounts;
/// src/utils.cpp:258-292: the run's JSON record appended to <args.json>/<args.output_file>, with the
/// reference's keys.
void write_json(const std::vector<long> &block_assignment, double description_length, unsigned long MCMC_moves,
                unsigned long MCMC_iterations, double runtime);
}  // namespace utils

/// A host snapshot of a blockmodel's block matrix (the read side of ISparseMatrix,
/// include/blockmodel/sparse/csparse_matrix.hpp:43-120), taken when Blockmodel::blockmatrix() is called.  It does
/// not write back: change the blockmodel through its methods (the mutators throw std::logic_error).
class ISparseMatrix {
  public:
    ISparseMatrix(long num_blocks, std::vector<int32_t> row_ptr, std::vector<int32_t> cols, std::vector<int32_t> cnts);
    long get(long row, long col) const;
    std::vector<long> getrow(long row) const;
    std::vector<long> getcol(long col) const;
    MapVector<long> getrow_sparse(long row) const;
    MapVector<long> getcol_sparse(long col) const;
    EdgeWeights outgoing_edges(long block) const;
    EdgeWeights incoming_edges(long block) const;
    /// row(block) + col(block), the diagonal once (dict_transpose_matrix.cpp:176-188)
    MapVector<long> neighbors_weights(long block) const;
    std::vector<std::tuple<long, long, long>> entries() const;
    long distinct_edges(long block) const;
    long edges() const;
    long trace() const;
    std::vector<long> shape() const { return {B_, B_}; }
    void add(long, long, long) { read_only(); }
    void sub(long, long, long) { read_only(); }
    void clearrow(long) { read_only(); }
    void clearcol(long) { read_only(); }

  private:
    [[noreturn]] static void read_only();
    long B_;
    std::vector<int32_t> row_ptr_, cols_, cnts_;      // CSR, ascending columns
    std::vector<int32_t> col_ptr_, rows_, col_cnts_;  // CSC, ascending rows
};

/// Device-resident blockmodel (include/blockmodel/blockmodel.hpp:46-219).  Copy-assignment shares the
/// device state, as the reference's copy-assignment shares its block matrix; copy() deep-copies.
class Blockmodel {
  public:
    Blockmodel() = default;  // empty
    /// Blockmodel(num_blocks, graph, rate): every vertex in its own block (num_blocks == N).
    Blockmodel(long num_blocks, const Graph &graph, double block_reduction_rate);
    /// Blockmodel(num_blocks, graph, rate, assignment) + initialize_edge_counts.
    Blockmodel(long num_blocks, const Graph &graph, double block_reduction_rate, std::vector<long> &assignment);
    Blockmodel copy() const;
    long num_blocks() const;
    long num_vertices() const { return N_; }
    std::vector<long> block_assignment() const;
    long block_assignment(long vertex) const { return block_assignment()[vertex]; }
    std::vector<long> block_sizes() const;
    std::vector<long> degrees() const;
    std::vector<long> degrees_in() const;
    std::vector<long> degrees_out() const;
    double getOverall_entropy() const { return overall_entropy_; }
    void setOverall_entropy(double e) { overall_entropy_ = e; }
    long &getNum_blocks_to_merge() { return num_blocks_to_merge_; }
    void setNum_blocks_to_merge(long n) { num_blocks_to_merge_ = n; }
    /// Blockmodel::validate: rebuild from the assignment and compare every count.
    bool validate(const Graph &graph) const;
    /// log_posterior_probability (src/blockmodel/blockmodel.cpp:288-306)
    double log_posterior_probability(const Graph &graph) const;
    /// initialize_edge_counts (src/blockmodel/blockmodel.cpp:193-276): recount from the assignment.
    void initialize_edge_counts(const Graph &graph);
    /// move_vertex(const VertexMove_v3&) (src/blockmodel/blockmodel.cpp:437-494): false (no change) when the
    /// vertex's block has one vertex.
    bool move_vertex(const VertexMove_v3 &move);
    /// move_vertex(Vertex, const Delta&, ProposalAndEdgeCounts&) (src/blockmodel/blockmodel.cpp:408-435): the
    /// blockmodel of the new assignment (not for degree-corrected blockmodels, whose histograms this form keys by
    /// the proposal's edge counts).
    void move_vertex(Vertex vertex, const Delta &delta, utils::ProposalAndEdgeCounts &proposal);
    /// merge_block (src/blockmodel/blockmodel.cpp:337-363): every vertex of merge_from moves to merge_to.
    void merge_block(long merge_from, long merge_to, const Delta &delta, utils::ProposalAndEdgeCounts proposal);
    /// carry_out_best_merges (src/blockmodel/blockmodel.cpp:55-122), the approximate carry-out, of
    /// getNum_blocks_to_merge() merges, with label compaction; the block matrix is recounted as merge_blocks'
    /// initialize_edge_counts would.
    void carry_out_best_merges(const std::vector<double> &delta_entropy_for_each_block,
                               const std::vector<long> &best_merge_for_each_block);
    /// blockmatrix() (include/blockmodel/blockmodel.hpp:150): a snapshot of the block matrix (see ISparseMatrix).
    std::shared_ptr<ISparseMatrix> blockmatrix() const;
    /// Graph the blockmodel was built on (needed to recount after a move); set by the constructors and the
    /// driver.  The graph must outlive the blockmodel's use of it.
    const Graph *graph() const { return graph_; }
    void set_graph(const Graph *g) { graph_ = g; }
    sbp_bm handle() const { return bm_.get(); }
    sbp_ctx ctx() const { return ctx_; }
    /// Takes ownership of a device blockmodel handle.
    static Blockmodel adopt(sbp_ctx ctx, sbp_bm bm, long num_vertices, double overall_entropy,
                            long num_blocks_to_merge);
    bool empty = true;

  private:
    std::shared_ptr<struct sbp_bm_s> bm_;
    sbp_ctx ctx_ = nullptr;
    const Graph *graph_ = nullptr;
    long N_ = 0;
    double overall_entropy_ = 1.7976931348623157e308;
    long num_blocks_to_merge_ = 0;
};

/// Golden-section bracket over block counts (src/blockmodel/blockmodel_triplet.cpp:4-105).
class BlockmodelTriplet {
  public:
    bool optimal_num_blocks_found = false;
    void update(Blockmodel &blockmodel);
    void status() const;
    Blockmodel &get(long i) { return blockmodels_[i]; }
    bool golden_ratio_not_reached() const { return blockmodels_[2].empty; }
    bool is_done();
    Blockmodel get_next_blockmodel(Blockmodel &old_blockmodel);

  private:
    long nb(long i) const { return blockmodels_[i].empty ? 0 : blockmodels_[i].num_blocks(); }
    Blockmodel blockmodels_[3];
};

namespace sbp {
/// src/sbp.cpp:16-72.  The epoch counter of the replay contract starts at 0 for every call.
Blockmodel stochastic_block_partition(Graph &graph, Args &args, bool divide_and_conquer = false);
bool done_blockmodeling(Blockmodel &blockmodel, BlockmodelTriplet &blockmodel_triplet, long min_num_blocks = 0);
/// Statistics of the last stochastic_block_partition call on this thread (timers::* equivalents).
const sbp_stats &last_stats();
/// Block count of the last stochastic_block_partition call's result (the final row of the results CSV).
long last_num_blocks();
}  // namespace sbp

namespace block_merge {
/// src/block_merge.cpp:178-228.  Consumes one epoch of the replay counter.
Blockmodel &merge_blocks(Blockmodel &blockmodel, const Graph &graph, long num_edges);
/// src/block_merge.cpp:14-38: the Delta of merging current_block into proposed_block.
Delta blockmodel_delta(long current_block, long proposed_block, const Blockmodel &blockmodel);
/// src/block_merge.cpp:40-117: the heap carry-out with re-evaluation, of getNum_blocks_to_merge() merges, label
/// compaction and recount.
void carry_out_best_merges_advanced(Blockmodel &blockmodel, const std::vector<double> &delta_entropy_for_each_block,
                                    const std::vector<long> &best_merge_for_each_block, const Graph &graph);
}  // namespace block_merge

namespace divisive {
/// src/top_down_sbp.cpp:488-549 with random splits (the default --split) and apply_best_splits (:30-66): the
/// blockmodel's blocks are split in place, up to target_num_communities blocks, and it is returned.  Consumes one
/// epoch of the replay counter.  std::runtime_error for a degree-corrected blockmodel.
Blockmodel split_communities(Blockmodel &blockmodel, const Graph &graph, int target_num_communities);
}  // namespace divisive

namespace finetune {
/// src/finetune.cpp:754-771.  Dispatches on args.algorithm (and args.asynciterations).
Blockmodel &mcmc(int iteration, const Graph &graph, Blockmodel &blockmodel, BlockmodelTriplet &blockmodel_triplet);
Blockmodel &asynchronous_gibbs(Blockmodel &blockmodel, const Graph &graph, bool golden_ratio_not_reached);
Blockmodel &metropolis_hastings(Blockmodel &blockmodel, const Graph &graph, bool golden_ratio_not_reached);
Blockmodel &hybrid_mcmc(Blockmodel &blockmodel, const Graph &graph, bool golden_ratio_not_reached);
Blockmodel &finetune_assignment(Blockmodel &blockmodel, Graph &graph);
/// src/finetune.cpp:419-430: a vertex's neighbours (weights 1); ignore_self drops the vertex itself.
EdgeWeights edge_weights(const Graph &graph, long vertex, bool in_edges, bool ignore_self);
/// src/finetune.cpp:201-235: the Delta of moving `vertex` from current_block to proposed_block.
Delta blockmodel_delta(long vertex, long current_block, long proposed_block, EdgeWeights &out_edges,
                       EdgeWeights &in_edges, const Blockmodel &blockmodel);
}  // namespace finetune

namespace entropy {
/// src/entropy.cpp:597-604
double mdl(const Blockmodel &blockmodel, const Graph &graph);
/// src/entropy.cpp:298-403: the change in description length of the vertex move a Delta from
/// finetune::blockmodel_delta describes (evaluated on the device; std::invalid_argument for other Deltas).
double delta_mdl(const Blockmodel &blockmodel, const Delta &delta, const utils::ProposalAndEdgeCounts &proposal);
/// src/entropy.cpp:129-182: the same for a block merge (a Delta from block_merge::blockmodel_delta).
double block_merge_delta_mdl(long current_block, utils::ProposalAndEdgeCounts proposal, const Blockmodel &blockmodel,
                             const Delta &delta);
/// src/entropy.cpp:497-561: the Hastings correction of that vertex move (1 when greedy or nonparametric).
double hastings_correction(long vertex, const Graph &graph, const Blockmodel &blockmodel, const Delta &delta,
                           long current_block, const utils::ProposalAndEdgeCounts &proposal);
/// src/entropy.cpp:571-585
double null_mdl_v1(const Graph &graph);
/// src/entropy.cpp:563-565
double normalize_mdl_v1(double mdl, const Graph &graph);
}  // namespace entropy

namespace Hungarian {
/// The contingency-table type of the reference's evaluation (extern/hungarian): rows x columns of counts.
typedef std::vector<std::vector<int>> Matrix;
}  // namespace Hungarian

namespace evaluate {
/// include/evaluate.hpp:16-20
struct Eval {
    double f1_score;
    double nmi;
    double true_mdl;
};
/// src/evaluate.cpp:9-59 (on the host; prints Accuracy, Rand index, Recall, Precision and F1 Score)
double calculate_f1_score(long num_vertices, Hungarian::Matrix &contingency_table);
/// src/evaluate.cpp:61-96
double calculate_nmi(long num_vertices, Hungarian::Matrix &contingency_table);
/// src/evaluate.cpp:134-210: the table matched by maximum weight, rows = the blockmodel's communities
Hungarian::Matrix hungarian(const Graph &graph, Blockmodel &blockmodel);
/// src/evaluate.cpp:98-112: F1, NMI and the description length of the true partition
Eval evaluate_blockmodel(const Graph &graph, Blockmodel &blockmodel);
/// src/evaluate.cpp:205-326: appends the run's row to <args.csv><args.numvertices>/<args.type>.csv (header first
/// when the file is new), with the reference's columns; the run is one profile row (iteration -1, totals).
void write_results(const Graph &graph, const Eval &eval, double runtime);
}  // namespace evaluate

namespace sample {
/// include/sample.hpp:27-30: the sampled subgraph and mapping[v] = sample id of v, or -1.
struct Sample {
    Graph graph;
    std::vector<long> mapping;
};
/// src/sample.cpp:352-365: args.samplingalg in {max_degree, random, degree_product, degree_weighted},
/// (long)(args.samplesize * N) vertices, induced subgraph (from_vertices) built on the device.
Sample sample(const Graph &graph);
/// src/sample.cpp:157-200: block of every vertex of `graph` from the sample's blockmodel.
std::vector<long> extend(const Graph &graph, const Blockmodel &sample_blockmodel, const Sample &sample);
/// src/sample.cpp:12-26: the subgraph of the vertices of degree > 1 (--detach).
Sample detach(const Graph &graph);
/// src/sample.cpp:367-405: the full graph's blockmodel from the detached graph's (random draws keyed by
/// args.seed).
Blockmodel reattach(const Graph &graph, const Blockmodel &sample_blockmodel, const Sample &sample);
}  // namespace sample

namespace sbp_amd {
/// Arguments used by the namespaced entry points above (the reference reads the global `args`).
Args &args();
/// Replay epoch counter shared by merge_blocks and the finetune sweeps.
uint32_t &epoch();
/// Translates the flags into the C-ABI parameter block.
sbp_params to_params(const Args &a);
/// The agglomerative driver over a context and a device graph (shared by sbp::stochastic_block_partition
/// and the C-ABI's sbp_partition).
Blockmodel drive(sbp_ctx ctx, sbp_graph graph, int32_t num_vertices, const sbp_params &p, sbp_stats *stats);
}  // namespace sbp_amd

#endif  // SBP_AMD_SBP_HPP
