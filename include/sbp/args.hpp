// Command-line flags of the SBP hot path, with the reference's names and defaults
// (vtsynergy/SBP include/args.hpp:14-213).  The parser is this project's own; only flags that change the
// hot path's behaviour are acted on, the rest are accepted for command-line compatibility.
#ifndef SBP_AMD_ARGS_HPP
#define SBP_AMD_ARGS_HPP

#include <climits>
#include <cstdint>
#include <string>

class Args {
  public:
    std::string algorithm = "hybrid_mcmc";  // -a/--algorithm (args.hpp:65-68)
    bool approximate = false;               // --approximate (:69-71)
    int asynciterations = INT_MAX;          // --asynciterations (:72-74)
    int batches = 1;                        // --batches (:75-78)
    int cachesize = 100;                    // --cachesize (:81); logf caching changes nothing numerically
    bool degreecorrected = false;           // --degreecorrected (:87-88)
    bool degreeproductsort = false;         // --degreeproductsort (:89-90)
    bool detach = false;                    // --detach (:93-94)
    bool evaluate = false;                  // --evaluate (:107-108)
    std::string filepath;                   // -f/--filepath (:109-110), required by the CLI
    bool greedy = true;                     // --greedy (:111-112): a switch whose default is true
    std::string json = "output";            // -j/--json (:113-114)
    double mh_percent = 0.075;              // -m/--mh_percent (:115-116)
    bool modularity = false;                // --modularity (:118-119)
    bool nodelta = false;                   // --nodelta (:120-121)
    bool noduplicates = false;              // --noduplicates (:122-124)
    bool nonparametric = false;             // --nonparametric (:126-127)
    bool ordered = false;                   // --ordered (:130)
    std::string output_file;                // --output_file (:131-132)
    double samplesize = 1.0;                // --samplesize (:135-136)
    std::string samplingalg = "random";     // --samplingalg (:137-138)
    std::string tag = "default tag";        // --tag (:148-150)
    int threads = 1;                        // --threads (:151-152); the device path ignores it
    bool no_transpose = false;              // --no_transpose (:153-154); one CSR+CSC layout on the device
    bool undirected = false;                // --undirected (:157-158)
    // named in the reference's JSON/CSV records only (:21-50; the graph-tool generator's naming):
    std::string blocksizevar = "low";       // -b/--blocksizevar
    std::string csv = "./eval/test";        // -c/--csv: results CSV prefix (evaluate::write_results)
    std::string distribute = "none";        // --distribute (MPI data distribution; single rank here)
    int numvertices = 1000;                 // -n/--numvertices
    std::string overlap = "low";            // -o/--overlap
    int subgraphs = 0;                      // --subgraphs
    std::string subgraphpartition = "round_robin";  // --subgraphpartition
    std::string type = "static";            // -t/--type
    uint64_t seed = 0;                      // --seed: replay seed (new; the reference seeds from random_device)
    int device = 0;                         // --device: HIP device ordinal (new)

    Args() = default;
    /// Parses argv; throws std::invalid_argument on an unknown flag or a missing value.
    Args(int argc, char **argv);
};

#endif  // SBP_AMD_ARGS_HPP
