// Prints the graph that Graph::load() reads for the global args: "N E", then each vertex's out- and
// in-neighbour lists in stored order, then the true assignment.  Touches no device.
//   load_check <graph base path> [--undirected] [--noduplicates]
#include <cstdio>
#include <cstring>

#include "args.hpp"
#include "globals.hpp"
#include "graph.hpp"

int main(int argc, char *argv[]) {
    if (argc < 2) return 2;
    args.filepath = argv[1];
    for (int i = 2; i < argc; ++i) {
        if (!strcmp(argv[i], "--undirected")) args.undirected = true;
        else if (!strcmp(argv[i], "--noduplicates")) args.noduplicates = true;
        else return 2;
    }
    Graph graph = Graph::load();
    printf("%ld %ld\n", graph.num_vertices(), graph.num_edges());
    for (long v = 0; v < graph.num_vertices(); ++v) {
        printf("out %ld:", v);
        for (long w : graph.out_neighbors(v)) printf(" %ld", w);
        printf("\nin %ld:", v);
        for (long w : graph.in_neighbors(v)) printf(" %ld", w);
        printf("\n");
    }
    printf("truth:");
    for (long c : graph.assignment()) printf(" %ld", c);
    printf("\n");
    return 0;
}
