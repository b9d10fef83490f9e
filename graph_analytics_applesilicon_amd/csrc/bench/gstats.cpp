// gstats: dataset statistics, the reference's tool (src/graph/stats.cpp) over
// the host C-ABI.
//   gstats <graph.bin> [--csv results/dataset_stats.csv] [--hist results/degree_hist.csv]
// Loads and validates the graph (and its symmetrization), then reports n, m,
// degree figures, isolated vertices, connected components, an approximate
// diameter and a log2 degree histogram; each CSV keeps one block of rows per
// dataset. Exit codes: 0 ok, 1 failure, 2 usage.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "cli.hpp"
#include "ggpu_graph.h"

namespace {

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

void hist_rows(std::vector<std::string>& rows, const std::string& name, const char* kind, const ggpu_graph* g,
               int which) {
    uint64_t h[34];
    uint32_t k = 0;
    if (ggpu_degree_histogram(g, which, h, &k) != GGPU_OK) return;
    for (uint32_t b = 0; b < k; ++b) {
        const unsigned long long lo = b ? 1ull << (b - 1) : 0, hi = b ? (1ull << b) - 1 : 0;
        char row[256];
        std::snprintf(row, sizeof row, "%s,%s,%u,%llu,%llu,%llu", name.c_str(), kind, b, lo, hi,
                      (unsigned long long)h[b]);
        rows.push_back(row);
    }
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> pos;
    std::string csv = "results/dataset_stats.csv", hist = "results/degree_hist.csv";
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--csv" && i + 1 < argc) csv = argv[++i];
        else if (a == "--hist" && i + 1 < argc) hist = argv[++i];
        else pos.push_back(a);
    }
    if (pos.size() != 1 || pos[0].rfind("--", 0) == 0) {
        std::fprintf(stderr, "usage: gstats <graph.bin> [--csv path] [--hist path]\n");
        return 2;
    }
    const std::string path = pos[0], name = cli::stem(path);
    auto t = std::chrono::steady_clock::now();
    ggpu_graph* g = nullptr;
    if (ggpu_graph_load(path.c_str(), &g) != GGPU_OK) {
        std::fprintf(stderr, "gstats: %s\n", ggpu_host_last_error());
        return 1;
    }
    char err[512];
    if (ggpu_graph_validate(g, err, sizeof err) != GGPU_OK) {
        std::fprintf(stderr, "gstats: %s INVALID: %s\n", path.c_str(), err);
        ggpu_graph_free(g);
        return 1;
    }
    const double load_ms = ms_since(t);
    t = std::chrono::steady_clock::now();
    ggpu_graph_stats s{};
    if (ggpu_graph_stats_compute(g, &s) != GGPU_OK) {
        std::fprintf(stderr, "gstats: %s\n", ggpu_host_last_error());
        ggpu_graph_free(g);
        return 1;
    }
    const double stats_ms = ms_since(t);
    const bool directed = ggpu_graph_flags(g) & GGPU_FLAG_DIRECTED;
    struct stat fst;
    const double file_mb = stat(path.c_str(), &fst) == 0 ? fst.st_size / (1024.0 * 1024.0) : 0;
    std::printf("%s (%s)\n", name.c_str(), directed ? "directed" : "undirected");
    std::printf("  n %llu, m %llu (stored entries), file %.1f MB\n", (unsigned long long)s.n, (unsigned long long)s.m,
                file_mb);
    std::printf("  symmetrized: %llu entries = %llu undirected edges\n", (unsigned long long)(2 * s.undirected_edges),
                (unsigned long long)s.undirected_edges);
    std::printf("  %sdegree min/mean/max %u / %.2f / %u", directed ? "out-" : "", s.min_deg, s.mean_deg, s.max_deg);
    if (directed) std::printf(", max in-degree %u", s.max_in_deg);
    std::printf("\n  isolated vertices (symmetrized) %llu\n", (unsigned long long)s.isolated);
    std::printf("  connected components %llu, largest %llu (%.1f%%)\n", (unsigned long long)s.num_cc,
                (unsigned long long)s.largest_cc, s.n ? 100.0 * s.largest_cc / s.n : 0.0);
    std::printf("  approx diameter (10 BFS) %u\n", s.approx_diameter);
    std::printf("  timings: load+validate %.0f ms, statistics %.0f ms\n", load_ms, stats_ms);
    if (!csv.empty()) {
        char row[512];
        std::snprintf(row, sizeof row, "%s,%llu,%llu,%d,%llu,%u,%.3f,%u,%u,%llu,%llu,%llu,%u,%.1f", name.c_str(),
                      (unsigned long long)s.n, (unsigned long long)s.m, directed ? 1 : 0,
                      (unsigned long long)s.undirected_edges, s.min_deg, s.mean_deg, s.max_deg, s.max_in_deg,
                      (unsigned long long)s.isolated, (unsigned long long)s.num_cc, (unsigned long long)s.largest_cc,
                      s.approx_diameter, file_mb);
        cli::csv_upsert(csv,
                        "dataset,n,m,directed,undirected_edges,min_deg,mean_deg,max_deg,max_in_deg,isolated,num_cc,"
                        "largest_cc,approx_diameter,file_mb",
                        name, {row});
    }
    if (!hist.empty()) {
        std::vector<std::string> rows;
        hist_rows(rows, name, directed ? "out" : "deg", g, 0);
        if (directed) {
            hist_rows(rows, name, "in", g, 1);
            hist_rows(rows, name, "sym", g, 2);
        }
        cli::csv_upsert(hist, "dataset,kind,bucket,lo,hi,count", name, rows);
    }
    ggpu_graph_free(g);
    return 0;
}
