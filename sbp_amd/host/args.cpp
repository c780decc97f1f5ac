// Flag parser for the reference's command line (include/args.hpp:61-212 lists the flags and defaults).
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <stdexcept>
#include <string>

#include "../../include/sbp/args.hpp"

Args::Args(int argc, char **argv) {
    auto need = [&](int &i) -> std::string {
        if (i + 1 >= argc) throw std::invalid_argument(std::string("missing value for ") + argv[i]);
        return std::string(argv[++i]);
    };
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "-a" || a == "--algorithm") algorithm = need(i);
        else if (a == "--approximate") approximate = true;
        else if (a == "--asynciterations") asynciterations = std::stoi(need(i));
        else if (a == "--batches") batches = std::stoi(need(i));
        else if (a == "--cachesize") cachesize = std::stoi(need(i));
        else if (a == "--degreecorrected") degreecorrected = true;
        else if (a == "--degreeproductsort") degreeproductsort = true;
        else if (a == "--detach") detach = true;
        else if (a == "--evaluate") evaluate = true;
        else if (a == "-f" || a == "--filepath") filepath = need(i);
        else if (a == "--greedy") greedy = false;  // TCLAP SwitchArg whose default is true: passing it flips it
        else if (a == "-j" || a == "--json") json = need(i);
        else if (a == "-m" || a == "--mh_percent") mh_percent = std::stod(need(i));
        else if (a == "--modularity") modularity = true;
        else if (a == "--nodelta") nodelta = true;
        else if (a == "--noduplicates") noduplicates = true;
        else if (a == "--nonparametric") {
            nonparametric = true;
            greedy = false;  // args.hpp:204-207
        } else if (a == "--ordered") ordered = true;
        else if (a == "--output_file") output_file = need(i);
        else if (a == "--samplesize") samplesize = std::stod(need(i));
        else if (a == "--samplingalg") samplingalg = need(i);
        else if (a == "--tag") tag = need(i);
        else if (a == "--threads") threads = std::stoi(need(i));
        else if (a == "--no_transpose") no_transpose = true;
        else if (a == "--undirected") undirected = true;
        else if (a == "--seed") seed = std::strtoull(need(i).c_str(), nullptr, 10);
        else if (a == "--device") device = std::stoi(need(i));
        // named in the JSON and CSV records
        else if (a == "-c" || a == "--csv") csv = need(i);
        else if (a == "-o" || a == "--overlap") overlap = need(i);
        else if (a == "-b" || a == "--blocksizevar") blocksizevar = need(i);
        else if (a == "-t" || a == "--type") type = need(i);
        else if (a == "-n" || a == "--numvertices") numvertices = std::stoi(need(i));
        else if (a == "--subgraphs") subgraphs = std::stoi(need(i));
        else if (a == "--subgraphpartition") subgraphpartition = need(i);
        else if (a == "--distribute") distribute = need(i);
        // flags of modes outside this path (MPI, divisive) are accepted
        else if (a == "-d" || a == "--directory" || a == "--delimiter" || a == "--splitinit")
            (void)need(i);
        else if (a == "--nonblocking" || a == "--split" || a == "--mix") {
        } else throw std::invalid_argument("unknown flag " + a);
    }
    if (output_file.empty()) {  // args.hpp:187-191
        char b[64];
        snprintf(b, sizeof(b), "%ld_%ld.json", (long)getpid(), (long)time(nullptr));
        output_file = b;
    }
    if (batches < 1) throw std::invalid_argument("--batches must be >= 1");
    if (threads <= 0) threads = 1;  // threads <= 0 creates no RNG generators in the reference (src/rng.cpp:37-38)
}
