/* Host graph library: the GGPUCSR1 .bin format, generators, transforms and
 * the CPU BFS / PageRank variants that gbench runs beside the GPU.
 *
 * Replaces, for C callers:
 *   graph::CSR, load/save/validate/equal      src/graph/csr.hpp:29-70
 *   transpose/symmetrize/permute              src/graph/transform.hpp:11-19
 *   gen::path/star/.../rmat/suite             src/graph/generators.hpp:12-24
 *   reorder::compute(..., "random", seed)     src/graph/reorder.cpp:20-25, :94-100
 *   cpu::bfsSerial/bfsOmp/bfsDO               src/cpu/bfs_cpu.hpp:29-39
 *   cpu::pagerankSerial/pagerankOmp           src/cpu/pagerank_cpu.hpp:32-33
 *   gbench pickSources/largestComponent       src/bench/main.cpp:130-169
 * Graphs are opaque host objects; every function returns a GGPU_* status and
 * ggpu_host_last_error() describes the last failure on the calling thread. */
#ifndef GGPU_GRAPH_H
#define GGPU_GRAPH_H

#include "ggpu_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ggpu_graph ggpu_graph;

const char* ggpu_host_last_error(void);

/* ---- construction / format (src/graph/csr.hpp:47-70, csr.cpp:30-163) ---- */
int ggpu_graph_load(const char* path, ggpu_graph** out);
int ggpu_graph_save(const ggpu_graph* g, const char* path);
/* Copies the arrays. row_ptr has n+1 entries, col_idx m entries. */
int ggpu_graph_from_arrays(uint64_t n, uint64_t m, const uint32_t* row_ptr, const uint32_t* col_idx,
                           uint32_t flags, ggpu_graph** out);
/* Edge list -> CSR; undirected edges are stored both ways; lists sorted,
 * deduplicated, self-loops dropped (graph::fromEdgeList, csr.cpp:146-163). */
int ggpu_graph_from_edges(uint64_t n, uint64_t num_edges, const uint32_t* src, const uint32_t* dst,
                          int directed, ggpu_graph** out);
void ggpu_graph_free(ggpu_graph* g);
uint64_t ggpu_graph_n(const ggpu_graph* g);
uint64_t ggpu_graph_m(const ggpu_graph* g);
uint32_t ggpu_graph_flags(const ggpu_graph* g);
const uint32_t* ggpu_graph_row_ptr(const ggpu_graph* g);
const uint32_t* ggpu_graph_col_idx(const ggpu_graph* g);
/* "" (returns GGPU_OK) if structurally valid, else a description in buf. */
int ggpu_graph_validate(const ggpu_graph* g, char* buf, size_t buflen);
int ggpu_graph_equal(const ggpu_graph* a, const ggpu_graph* b);

/* ---- transforms (src/graph/transform.cpp:9-76) ---- */
int ggpu_graph_transpose(const ggpu_graph* g, ggpu_graph** out);
int ggpu_graph_symmetrize(const ggpu_graph* g, ggpu_graph** out);
/* vertex v becomes perm[v] */
int ggpu_graph_permute(const ggpu_graph* g, const uint32_t* perm, ggpu_graph** out);
/* Fisher-Yates permutation with the SplitMix64 generator (reorder.cpp:20-25). */
int ggpu_random_perm(uint64_t n, uint64_t seed, uint32_t* perm_out);
/* graph::reorder::compute(sym, ordering, seed) (src/graph/reorder.cpp:97-103):
 * "original" | "random" | "degree" (descending degree, ties by id) | "rcm"
 * (reverse Cuthill-McKee per component from a pseudo-peripheral start).
 * perm_out[old] = new, n entries; computed on a symmetric graph. */
int ggpu_order_perm(const ggpu_graph* sym, const char* ordering, uint64_t seed, uint32_t* perm_out);
/* graph::reorder::locality (reorder.cpp:113-125): mean |u - v| over the
 * stored arcs and the share of arcs within one 64-vertex block. */
int ggpu_graph_locality(const ggpu_graph* g, double* mean_gap, double* same_block64);

/* graph::convertSnap (src/graph/snap.cpp:85-148) without the file write: a
 * SNAP edge list (plain or gzip text, "u v" lines, '#'/'%' comments, extra
 * columns ignored) as a CSR; the raw IDs that occur are numbered densely in
 * increasing order; directed = 0 stores both directions of each line; lists
 * sorted, deduplicated, self-loops dropped. Save with ggpu_graph_save. */
typedef struct ggpu_snap_stats {
    uint64_t lines;      /* edge lines parsed (comments excluded) */
    uint64_t raw_max_id; /* largest raw vertex ID */
    uint64_t n;          /* vertices after relabeling */
    uint64_t m_before;   /* list entries before dedup / self-loop removal */
    uint64_t m;          /* list entries kept */
} ggpu_snap_stats;
int ggpu_graph_from_snap(const char* path, int directed, ggpu_graph** out, ggpu_snap_stats* stats);

/* The figures gstats prints (src/graph/stats.cpp:40-120): stored-graph
 * degrees (out-degrees when directed), isolated vertices, connected
 * components and an approximate diameter (deepest BFS level from 10
 * SplitMix64(1) sources of nonzero degree) of the symmetrized graph. */
typedef struct ggpu_graph_stats {
    uint64_t n, m, undirected_edges;
    uint32_t min_deg, max_deg, max_in_deg, approx_diameter;
    double mean_deg;
    uint64_t isolated, num_cc, largest_cc;
} ggpu_graph_stats;
int ggpu_graph_stats_compute(const ggpu_graph* g, ggpu_graph_stats* out);
/* log2 degree histogram: bucket 0 = degree 0, bucket k = [2^(k-1), 2^k);
 * kind 0 stored (out-) degrees, 1 in-degrees, 2 symmetrized degrees.
 * hist has room for 34 buckets; *buckets = used (trailing zeros dropped). */
int ggpu_degree_histogram(const ggpu_graph* g, int kind, uint64_t* hist, uint32_t* buckets);
/* gexport (src/graph/export_sg.cpp): the GAP Benchmark Suite .sg layout;
 * sym != 0 writes the symmetrized graph as undirected, else a directed graph
 * carries its transpose as the in-CSR. n must be below 2^31. */
int ggpu_graph_save_sg(const ggpu_graph* g, const char* path, int sym);
/* Largest connected component of a symmetric graph, relabeled in increasing
 * original-id order (row-major for grids); old_ids_out (optional, n entries)
 * receives the original id of each new vertex. */
int ggpu_graph_largest_component(const ggpu_graph* g, ggpu_graph** out, uint32_t* old_ids_out,
                                 uint64_t* n_out);

/* ---- generators (src/graph/generators.cpp:8-88, extended) ----
 * name: path1000 star100k complete200 two_components empty1000 grid100x100
 *       rmat12 rmat12_directed (the reference suite) and the workload graphs
 *       rmat16_c1 = permute(rmat(16, 16, .57, .19, .19, seed 1), random_perm(seed 1)),
 *       rmat22_c2 = permute(rmat(22, 8, .57, .19, .19, seed 2), random_perm(seed 2)),
 *       grid_c3   = largest component of grid(1024, 2048, keep 0.7, seed 3),
 *       rmat23_c5 = rmat(23, 16, .60, .17, .17, undirected, seed 5), ids not permuted. */
int ggpu_gen_named(const char* name, ggpu_graph** out);
int ggpu_gen_rmat(uint32_t scale, uint32_t edgefactor, double a, double b, double c, int directed,
                  uint64_t seed, ggpu_graph** out);
/* rows x cols 4-neighbour grid; each undirected edge kept with probability
 * keep (keep >= 1: all kept, the reference grid2d). */
int ggpu_gen_grid(uint32_t rows, uint32_t cols, double keep, uint64_t seed, ggpu_graph** out);

/* ---- gbench helpers (src/bench/main.cpp:130-169) ---- */
/* count distinct seeded sources with degree > 0 in the largest component
 * (anywhere if any_component). *n_out <= count. */
int ggpu_pick_sources(const ggpu_graph* g, uint32_t count, uint64_t seed, int any_component,
                      uint32_t* out, uint32_t* n_out);

/* ---- CPU BFS (src/cpu/bfs_cpu.cpp) ---- */
typedef struct ggpu_cpu_bfs_opts {
    int threads;      /* 0 = all cores */
    int schedule;     /* cpu_omp: 0 static, 1 dynamic,64, 2 guided */
    double alpha;     /* cpu_do */
    double beta;
} ggpu_cpu_bfs_opts;
void ggpu_cpu_bfs_opts_default(ggpu_cpu_bfs_opts* o);
/* variant: "cpu_serial" | "cpu_omp" | "cpu_do". trace may be NULL. */
int ggpu_cpu_bfs(const ggpu_graph* g, const char* variant, uint32_t source, const ggpu_cpu_bfs_opts* opt,
                 uint32_t* dist, ggpu_level* trace, size_t trace_cap, size_t* trace_len);

/* ---- CPU PageRank (src/cpu/pagerank_cpu.cpp:20-63) ---- */
/* in = transpose (in-neighbour lists); outdeg = out-degrees of the original
 * graph. parallel = 0: cpu_serial, 1: cpu_omp. r_out: n doubles. */
int ggpu_cpu_pagerank(const ggpu_graph* in, const uint32_t* outdeg, double damping, int iters, double tol,
                      int parallel, int threads, double* r_out, int* iters_out, double* last_delta_out);
/* acceptance (src/cpu/pagerank_cpu.hpp:37-60): L1, max rel over ref's top-1000, sum */
void ggpu_check_ranks(uint64_t n, const double* ref, const float* got, double* l1, double* max_rel_top,
                      double* sum);

#ifdef __cplusplus
}
#endif
#endif
