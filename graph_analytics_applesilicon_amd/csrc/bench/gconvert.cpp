// gconvert: SNAP edge list (.txt or .txt.gz) -> binary CSR (.bin), the
// reference's tool (src/graph/convert.cpp) over the host C-ABI.
//   gconvert <in.txt[.gz]> <out.bin> --directed|--undirected [--csv results/convert_stats.csv]
// Prints the conversion figures and replaces the dataset's row in the CSV
// (dataset,directed,lines,n,m_before,m,seconds,peak_rss_mb). Exit codes: 0 ok,
// 1 failure, 2 usage.
#include <sys/resource.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cli.hpp"
#include "ggpu_graph.h"

int main(int argc, char** argv) {
    std::vector<std::string> pos;
    std::string csv = "results/convert_stats.csv";
    int directed = -1;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--directed") directed = directed == -1 ? 1 : 2;
        else if (a == "--undirected") directed = directed == -1 ? 0 : 2;
        else if (a == "--csv" && i + 1 < argc) csv = argv[++i];
        else if (a.rfind("--", 0) == 0) directed = 2;
        else pos.push_back(a);
    }
    if (pos.size() != 2 || directed < 0 || directed > 1) {
        std::fprintf(stderr, "usage: gconvert <in.txt.gz> <out.bin> --directed|--undirected [--csv path]\n");
        return 2;
    }
    const auto t0 = std::chrono::steady_clock::now();
    ggpu_graph* g = nullptr;
    ggpu_snap_stats st{};
    if (ggpu_graph_from_snap(pos[0].c_str(), directed, &g, &st) != GGPU_OK ||
        ggpu_graph_save(g, pos[1].c_str()) != GGPU_OK) {
        std::fprintf(stderr, "gconvert: %s\n", ggpu_host_last_error());
        ggpu_graph_free(g);
        return 1;
    }
    ggpu_graph_free(g);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    struct rusage ru;
    getrusage(RUSAGE_SELF, &ru);
    const double rss_mb = ru.ru_maxrss / 1024.0;  // Linux reports KiB
    std::printf("%s -> %s (%s)\n", pos[0].c_str(), pos[1].c_str(), directed ? "directed" : "undirected");
    std::printf("  lines %llu, raw max id %llu, n %llu\n", (unsigned long long)st.lines,
                (unsigned long long)st.raw_max_id, (unsigned long long)st.n);
    std::printf("  entries %llu before dedup/self-loops, %llu after\n", (unsigned long long)st.m_before,
                (unsigned long long)st.m);
    std::printf("  %.2f s, peak RSS %.1f MB\n", secs, rss_mb);
    if (!csv.empty()) {
        const std::string name = cli::stem(pos[1]);
        char row[512];
        std::snprintf(row, sizeof row, "%s,%d,%llu,%llu,%llu,%llu,%.2f,%.1f", name.c_str(), directed,
                      (unsigned long long)st.lines, (unsigned long long)st.n, (unsigned long long)st.m_before,
                      (unsigned long long)st.m, secs, rss_mb);
        cli::csv_upsert(csv, "dataset,directed,lines,n,m_before,m,seconds,peak_rss_mb", name, {row});
    }
    return 0;
}
