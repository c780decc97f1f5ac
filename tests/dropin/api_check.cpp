// Exercises the drop-in C++ surface the way the reference's own callers use it (src/finetune.cpp:842-869,
// src/block_merge.cpp:40-117 and src/top_down_sbp.cpp:488-549), on a graph given as a TSV path: prints one line
// per checked value, which the GPU test compares with the oracle.
//   api_check <graph base path> <num_blocks> <seed>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "args.hpp"
#include "blockmodel/blockmodel.hpp"
#include "blockmodel/sparse/delta.hpp"
#include "block_merge.hpp"
#include "entropy.hpp"
#include "finetune.hpp"
#include "globals.hpp"
#include "graph.hpp"
#include "rng.hpp"
#include "top_down_sbp.hpp"
#include "utils.hpp"

int main(int argc, char *argv[]) {
    if (argc < 4) return 2;
    args.filepath = argv[1];
    args.greedy = false;  // so hastings_correction computes the correction
    args.seed = std::strtoull(argv[3], nullptr, 10);
    const long B = std::atol(argv[2]);
    rng::init_generators();
    Graph graph = Graph::load();
    std::vector<long> assignment(graph.num_vertices());
    for (long v = 0; v < graph.num_vertices(); ++v) assignment[v] = (v * 7919 + 13) % B;
    Blockmodel blockmodel(B, graph, 0.5, assignment);
    printf("mdl %.17g\n", entropy::mdl(blockmodel, graph));
    // vertex moves: the Delta of propose_move / move_vertex, its delta-MDL and Hastings correction
    for (long v = 0; v < graph.num_vertices(); v += 37) {
        const long r = blockmodel.block_assignment(v), s = (r + 1 + v % (B - 1)) % B;
        EdgeWeights out_edges = finetune::edge_weights(graph, v, false, false);
        EdgeWeights in_edges = finetune::edge_weights(graph, v, true, true);
        Delta delta = finetune::blockmodel_delta(v, r, s, out_edges, in_edges, blockmodel);
        long k_out = 0, k_in = 0;
        for (long w : out_edges.values) k_out += w;
        for (long w : in_edges.values) k_in += w;
        utils::ProposalAndEdgeCounts proposal{s, k_out, k_in, k_out + k_in};
        const double dS = entropy::delta_mdl(blockmodel, delta, proposal);
        const double H = entropy::hastings_correction(v, graph, blockmodel, delta, r, proposal);
        long delta_sum = 0;
        for (auto &e : delta.entries()) delta_sum += std::get<2>(e);
        printf("move %ld %ld %.17g %.17g %ld\n", v, s, dS, H, delta_sum);
    }
    // merges: the Delta of block_merge::blockmodel_delta and block_merge_delta_mdl
    std::shared_ptr<ISparseMatrix> matrix = blockmodel.blockmatrix();
    for (long r = 0; r < B; r += 3) {
        const long s = (r + 1) % B;
        Delta delta = block_merge::blockmodel_delta(r, s, blockmodel);
        EdgeWeights out_blocks = matrix->outgoing_edges(r), in_blocks = matrix->incoming_edges(r);
        long k_out = 0, k_in = 0;
        for (long w : out_blocks.values) k_out += w;
        for (long w : in_blocks.values) k_in += w;
        utils::ProposalAndEdgeCounts proposal{s, k_out, k_in, k_out + k_in};
        printf("merge %ld %ld %.17g\n", r, s, entropy::block_merge_delta_mdl(r, proposal, blockmodel, delta));
    }
    // committed moves (move_vertex(const VertexMove_v3&), with its refusal) and a merged block
    Blockmodel moved = blockmodel.copy();
    for (long v = 0; v < graph.num_vertices(); v += 53) {
        VertexMove_v3 move{0.0, true, {v, 0, 0}, (v * 31 + 5) % B, {}, {}};
        printf("commit %ld %ld %d\n", v, move.proposed_block, moved.move_vertex(move) ? 1 : 0);
    }
    printf("after_moves %.17g\n", entropy::mdl(moved, graph));
    utils::ProposalAndEdgeCounts none{0, 0, 0, 0};
    moved.merge_block(0, 1, Delta(0, 1), none);
    printf("after_merge %.17g %d\n", entropy::mdl(moved, graph), moved.validate(graph) ? 1 : 0);
    std::vector<long> a = moved.block_assignment();
    printf("assignment");
    for (long x : a) printf(" %ld", x);
    printf("\n");
    // the divisive step on the first blockmodel (src/top_down_sbp.cpp:488-549), the first replay epoch
    Blockmodel split = divisive::split_communities(blockmodel, graph, (int)(B + B / 2));
    std::vector<long> sa = split.block_assignment();
    printf("split %ld", split.num_blocks());
    for (long x : sa) printf(" %ld", x);
    printf("\n");
    return 0;
}
