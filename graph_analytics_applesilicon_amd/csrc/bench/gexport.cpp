// gexport: binary CSR (.bin) -> GAP Benchmark Suite serialized graph (.sg),
// the reference's tool (src/graph/export_sg.cpp) over the host C-ABI.
//   gexport <in.bin> <out.sg> [--sym]
// --sym writes the symmetrized graph as undirected (what BFS traverses);
// otherwise a directed graph carries its transpose as the in-CSR (what
// PageRank pulls over). Exit codes: 0 ok, 1 failure, 2 usage.
#include <cstdio>
#include <string>
#include <vector>

#include "ggpu_graph.h"

int main(int argc, char** argv) {
    std::vector<std::string> pos;
    bool sym = false, bad = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--sym") sym = true;
        else if (a.rfind("--", 0) == 0) bad = true;
        else pos.push_back(a);
    }
    if (pos.size() != 2 || bad) {
        std::fprintf(stderr, "usage: gexport <in.bin> <out.sg> [--sym]\n");
        return 2;
    }
    ggpu_graph* g = nullptr;
    if (ggpu_graph_load(pos[0].c_str(), &g) != GGPU_OK || ggpu_graph_save_sg(g, pos[1].c_str(), sym) != GGPU_OK) {
        std::fprintf(stderr, "gexport: %s\n", ggpu_host_last_error());
        ggpu_graph_free(g);
        return 1;
    }
    const bool directed = !sym && (ggpu_graph_flags(g) & GGPU_FLAG_DIRECTED);
    unsigned long long m = ggpu_graph_m(g);
    if (sym && (ggpu_graph_flags(g) & GGPU_FLAG_DIRECTED)) {
        ggpu_graph* s = nullptr;
        if (ggpu_graph_symmetrize(g, &s) == GGPU_OK) m = ggpu_graph_m(s);
        ggpu_graph_free(s);
    }
    std::printf("%s -> %s: %s, n %llu, %llu adjacency entries\n", pos[0].c_str(), pos[1].c_str(),
                directed ? "directed (+ in-CSR)" : "undirected", (unsigned long long)ggpu_graph_n(g), m);
    ggpu_graph_free(g);
    return 0;
}
