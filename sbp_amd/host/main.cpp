// Command-line driver: the reference's single-rank flow (src/main.cpp:55-164) written against the drop-in API
// with the reference's own header names and signatures -- the global `args`, rng::init_generators(),
// Graph::load(), sbp::stochastic_block_partition, sample::sample / extend / detach / reattach,
// Blockmodel(num_blocks, graph, 0.5, assignment), finetune::finetune_assignment, utils::write_json and
// evaluate::write_results.  MPI is not part of this build: the run is rank 0 of 1.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "args.hpp"
#include "blockmodel/blockmodel.hpp"
#include "entropy.hpp"
#include "evaluate.hpp"
#include "finetune.hpp"
#include "globals.hpp"
#include "graph.hpp"
#include "rng.hpp"
#include "sample.hpp"
#include "sbp.hpp"
#include "utils.hpp"

namespace {

struct Partition {
    Graph graph;
    Blockmodel blockmodel;
};

double seconds_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

// The JSON record, and with --evaluate the scores and the CSV row.
void report(Graph &graph, Blockmodel &blockmodel, double runtime) {
    if (mpi.rank != 0) return;
    utils::write_json(blockmodel.block_assignment(), blockmodel.getOverall_entropy(), timers::MCMC_moves,
                      timers::MCMC_iterations, runtime);
    if (!args.evaluate) return;
    evaluate::Eval result = evaluate::evaluate_blockmodel(graph, blockmodel);
    printf("Final F1 score = %g\nCommunity detection runtime = %gs\n", result.f1_score, runtime);
    if (!std::isfinite(result.nmi)) result.nmi = 0.0;
    evaluate::write_results(graph, result, runtime);
}

void partition_graph(Partition &p) {
    timers::total_num_islands = p.graph.num_islands();
    p.blockmodel = sbp::stochastic_block_partition(p.graph, args);
}

}  // namespace

int main(int argc, char *argv[]) {
    try {
        args = Args(argc, argv);
        if (args.filepath.empty()) throw std::invalid_argument("-f/--filepath is required");
        rng::init_generators();
        Graph graph = Graph::load();
        fprintf(stderr, "V: %ld E: %ld\n", graph.num_vertices(), graph.num_edges());
        sample::Sample detached;
        Partition partition;
        const auto start = std::chrono::steady_clock::now();
        if (args.detach) {  // partition the vertices of degree > 1, reattach the others at the end
            detached = sample::detach(graph);
            partition.graph = detached.graph;
            fprintf(stderr, "detached num vertices: %ld E: %ld\n", partition.graph.num_vertices(),
                    partition.graph.num_edges());
        } else {
            partition.graph = graph;
        }
        if (args.samplesize <= 0.0) {
            fprintf(stderr, "ERROR Sample size of %g is too low. Must be greater than 0.0\n", args.samplesize);
            return -5;
        }
        if (args.samplesize < 1.0) {  // partition a sample, extend its partition, finetune on the whole graph
            const auto t_sample = std::chrono::steady_clock::now();
            sample::Sample s = sample::sample(partition.graph);
            Partition sample_partition;
            sample_partition.graph = s.graph;
            timers::sample_time = seconds_since(t_sample);
            partition_graph(sample_partition);
            const auto t_extend = std::chrono::steady_clock::now();
            std::vector<long> assignment = sample::extend(partition.graph, sample_partition.blockmodel, s);
            const auto t_finetune = std::chrono::steady_clock::now();
            partition.blockmodel = Blockmodel(sample_partition.blockmodel.num_blocks(), partition.graph, 0.5,
                                              assignment);
            partition.blockmodel = finetune::finetune_assignment(partition.blockmodel, partition.graph);
            timers::sample_extend_time = std::chrono::duration<double>(t_finetune - t_extend).count();
            timers::sample_finetune_time = seconds_since(t_finetune);
        } else {
            partition_graph(partition);
        }
        if (args.detach) partition.blockmodel = sample::reattach(graph, partition.blockmodel, detached);
        const double runtime = seconds_since(start);
        report(graph, partition.blockmodel, runtime);
        printf("blocks=%ld mdl=%.6f runtime=%.3fs\n", partition.blockmodel.num_blocks(),
               partition.blockmodel.getOverall_entropy(), runtime);
        return 0;
    } catch (const std::exception &e) {
        fprintf(stderr, "ERROR %s\n", e.what());
        return 1;
    }
}
