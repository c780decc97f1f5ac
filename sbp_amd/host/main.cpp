// Command-line driver with the reference's flags (src/main.cpp:64-164, single-rank path): load the graph,
// partition it on the device, write a JSON result and, when a true partition is given, evaluate it
// (evaluate::evaluate_blockmodel: F1, NMI and the other pair-counting scores).
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sbp/sbp.hpp"

int main(int argc, char **argv) {
    try {
        Args args(argc, argv);
        if (args.filepath.empty()) throw std::invalid_argument("-f/--filepath is required");
        Graph graph = Graph::load(args);
        fprintf(stderr, "V: %ld E: %ld\n", graph.num_vertices(), graph.num_edges());
        sbp_amd::args() = args;
        sample::Sample detached;
        Graph *work = &graph;
        if (args.detach) {  // src/main.cpp:86-92: partition the vertices of degree > 1, reattach the rest
            detached = sample::detach(graph);
            work = &detached.graph;
            fprintf(stderr, "detached num vertices: %ld E: %ld\n", work->num_vertices(), work->num_edges());
        }
        if (args.samplesize <= 0.0) {  // src/main.cpp:93-97
            fprintf(stderr, "ERROR Sample size of %g is too low. Must be greater than 0.0\n", args.samplesize);
            return -5;
        }
        auto t0 = std::chrono::steady_clock::now();
        Blockmodel bm;
        if (args.samplesize < 1.0) {  // src/main.cpp:98-146: partition a sample, extend, finetune the full graph
            sample::Sample s = sample::sample(*work);
            fprintf(stderr, "sample V: %ld E: %ld\n", s.graph.num_vertices(), s.graph.num_edges());
            Blockmodel sample_bm = sbp::stochastic_block_partition(s.graph, args);
            std::vector<long> assignment = sample::extend(*work, sample_bm, s);
            bm = Blockmodel(sample_bm.num_blocks(), *work, 0.5, assignment);
            bm = finetune::finetune_assignment(bm, *work);
        } else {
            bm = sbp::stochastic_block_partition(*work, args);
        }
        if (args.detach) {  // src/main.cpp:153-156; the written description length is the full graph's
            bm = sample::reattach(graph, bm, detached);
            bm.setOverall_entropy(entropy::mdl(bm, graph));
        }
        double runtime = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const sbp_stats &st = sbp::last_stats();
        std::vector<long> a = bm.block_assignment();
        // evaluate_partition (src/main.cpp:41-53) with --evaluate, and whenever a true partition with more than
        // one community was loaded (Graph::load fills a missing one with zeros)
        const std::vector<long> &truth = graph.assignment();
        const bool has_truth = !truth.empty() && *std::max_element(truth.begin(), truth.end()) > 0;
        double score = -1.0, f1 = -1.0;
        if (!truth.empty() && (args.evaluate || has_truth)) {
            evaluate::Eval e = evaluate::evaluate_blockmodel(graph, bm);
            score = e.nmi;
            f1 = e.f1_score;
        }
        std::string out = args.output_file.empty() ? args.json + ".json" : args.output_file;
        FILE *f = fopen(out.c_str(), "w");
        if (!f) throw std::runtime_error("cannot write " + out);
        auto num = [](double x, const char *fmt) {  // JSON has no NaN or infinity
            char b[64];
            if (!std::isfinite(x)) return std::string("null");
            snprintf(b, sizeof(b), fmt, x);
            return std::string(b);
        };
        fprintf(f, "{\"num_blocks\": %ld, \"mdl\": %s, \"normalized_mdl_v1\": %s, \"nmi\": %s, \"f1_score\": %s, "
                   "\"runtime\": %.6f, \"mcmc_proposals\": %ld, \"mcmc_moves\": %ld, \"sweeps\": %ld, "
                   "\"merge_proposals\": %ld, \"assignment\": [",
                bm.num_blocks(), num(bm.getOverall_entropy(), "%.17g").c_str(),
                num(entropy::normalize_mdl_v1(bm.getOverall_entropy(), graph), "%.17g").c_str(),
                num(score, "%.6f").c_str(), num(f1, "%.6f").c_str(), runtime, (long)st.proposals, (long)st.moves, (long)st.sweeps, (long)st.merge_proposals);
        for (size_t i = 0; i < a.size(); ++i) fprintf(f, "%s%ld", i ? ", " : "", a[i]);
        fprintf(f, "]}\n");
        fclose(f);
        printf("blocks=%ld mdl=%.6f nmi=%.4f runtime=%.3fs\n", bm.num_blocks(), bm.getOverall_entropy(), score,
               runtime);
        return 0;
    } catch (const std::exception &e) {
        fprintf(stderr, "ERROR %s\n", e.what());
        return 1;
    }
}
