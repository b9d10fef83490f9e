// Host-side graph layer: the GGPUCSR1 CSR format, transforms, generators,
// the CPU BFS / PageRank variants and the hybrid BFS's CPU level steps.
// Everything here is plain C++17 + OpenMP (compiled by g++ into
// libggpu_host.so); the HIP backend links it for hybrid CPU levels and for
// the Gauss-Seidel ordering preprocessing.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "ggpu_types.h"

namespace ggpu {

constexpr uint32_t kInf = GGPU_INF;

// Owning CSR. row_ptr has n+1 entries; col_idx has exactly m entries.
// Format and flag bits follow src/graph/csr.hpp:14-45.
struct Csr {
    uint64_t n = 0, m = 0;
    uint32_t flags = 0;
    std::vector<uint32_t> rp, ci;

    bool directed() const { return flags & GGPU_FLAG_DIRECTED; }
    uint32_t deg(uint64_t v) const { return rp[v + 1] - rp[v]; }
    const uint32_t* nbr_begin(uint64_t v) const { return ci.data() + rp[v]; }
    const uint32_t* nbr_end(uint64_t v) const { return ci.data() + rp[v + 1]; }
};

// ---- SplitMix64 (the generator of src/util/rng.hpp:8-26: same constants,
// same Lemire bound and 53-bit double, so seeded streams match). Counter
// based: draw k of a stream seeded with s is mix(s + (k+1)*golden), which
// lets the generators below draw in parallel and still produce the serial
// sequence.
struct SplitMix64 {
    static constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
    uint64_t state;
    explicit SplitMix64(uint64_t seed) : state(seed) {}
    static uint64_t mix(uint64_t z) {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    uint64_t next() { state += kGolden; return mix(state); }
    uint64_t below(uint64_t bound) { return (uint64_t)(((unsigned __int128)next() * bound) >> 64); }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // The k-th (0-based) uniform draw of a stream seeded with `seed`.
    static double uniform_at(uint64_t seed, uint64_t k) {
        return (double)(mix(seed + (k + 1) * kGolden) >> 11) * (1.0 / 9007199254740992.0);
    }
};

// ---- format (csr.cpp:30-104)
void save_bin(const Csr& g, const std::string& path);
Csr load_bin(const std::string& path);
std::string validate(const Csr& g);
bool equal(const Csr& a, const Csr& b);

// ---- construction (csr.cpp:112-163). Lists come out sorted, deduplicated and
// without self-loops; undirected edges are stored in both directions.
Csr from_edges(uint64_t n, const std::vector<std::pair<uint32_t, uint32_t>>& edges, bool directed);
Csr from_arrays(uint64_t n, uint64_t m, const uint32_t* rp, const uint32_t* ci, uint32_t flags);

// ---- transforms (transform.cpp:9-76)
Csr transpose(const Csr& g);
Csr symmetrize(const Csr& g);
Csr permute(const Csr& g, const std::vector<uint32_t>& perm);
std::vector<uint32_t> random_perm(uint64_t n, uint64_t seed);
// reorder.cpp: --ordering original|random|degree|rcm on a symmetric graph
std::vector<uint32_t> order_perm(const Csr& sym, const std::string& ordering, uint64_t seed);
struct Locality {
    double mean_gap = 0;      // mean |u - v| over the stored arcs
    double same_block64 = 0;  // share of arcs with u / 64 == v / 64
};
Locality locality(const Csr& g);
// sorts each list, drops duplicates and self-loops, packs col_idx
void canonicalize_lists(uint64_t n, std::vector<uint32_t>& rp, std::vector<uint32_t>& ci);

// ---- snap.cpp: edge-list ingest (snap.cpp:85-148) and gstats figures (stats.cpp:40-120)
struct SnapStats {
    uint64_t lines = 0, raw_max_id = 0, n = 0, m_before = 0, m = 0;
};
Csr load_snap(const std::string& path, bool directed, SnapStats* st);
struct GraphStats {
    uint64_t n = 0, m = 0, undirected_edges = 0;
    uint32_t min_deg = 0, max_deg = 0, max_in_deg = 0;
    double mean_deg = 0;
    uint64_t isolated = 0, num_cc = 0, largest_cc = 0;
    uint32_t approx_diameter = 0;
};
GraphStats graph_stats(const Csr& g);
// GAP serialized graph (.sg) of g (export_sg.cpp); sym: the symmetrization, undirected
void save_sg(const Csr& g, const std::string& path, bool sym);
std::vector<uint64_t> log2_histogram(const std::vector<uint32_t>& deg);
std::vector<uint32_t> out_degrees(const Csr& g);
// Component id per vertex (BFS labeling in increasing root order) and the
// id of the largest component (first one found among equal sizes).
std::vector<uint32_t> components(const Csr& g, uint32_t* largest, uint64_t* largest_size);
Csr largest_component(const Csr& g, std::vector<uint32_t>* old_ids);

// ---- generators (generators.cpp:8-88 plus the parametric workloads)
Csr gen_path(uint32_t n);
Csr gen_star(uint32_t leaves);
Csr gen_complete(uint32_t n);
Csr gen_two_components();
Csr gen_empty(uint32_t n);
Csr gen_grid(uint32_t rows, uint32_t cols, double keep, uint64_t seed);
Csr gen_rmat(uint32_t scale, uint32_t edgefactor, double a, double b, double c, bool directed, uint64_t seed);
Csr gen_named(const std::string& name);
std::vector<std::string> suite_names();

// ---- gbench sources (main.cpp:130-169)
std::vector<uint32_t> pick_sources(const Csr& g, uint32_t count, uint64_t seed, bool any_component);

// ---- CPU BFS (bfs_cpu.cpp)
enum class Schedule { Static = 0, Dynamic64 = 1, Guided = 2 };
struct CpuBfsOpts {
    int threads = 0;
    Schedule schedule = Schedule::Dynamic64;
    double alpha = 15.0, beta = 18.0;
};
void bfs_serial(const Csr& g, uint32_t src, uint32_t* dist);
void bfs_omp(const Csr& g, uint32_t src, uint32_t* dist, const CpuBfsOpts& o, std::vector<ggpu_level>* trace);
void bfs_do(const Csr& g, uint32_t src, uint32_t* dist, const CpuBfsOpts& o, std::vector<ggpu_level>* trace);

// ---- hybrid BFS CPU level steps (bfs_steps.cpp:27-143), on the GPU layout:
// queue + degree arrays for top-down, bitmaps of u64 words (bit v%64 of
// word v/64) for bottom-up.
namespace steps {
struct Out { uint64_t found = 0, examined = 0, next_edges = 0; };
constexpr uint64_t kSerialEdges = 8192;
void init_dist(uint32_t* dist, uint64_t n, uint32_t src, int threads);
Out top_down(const Csr& g, uint32_t* dist, const uint32_t* q, uint64_t size, uint64_t edges, uint32_t* q_next,
             uint32_t* deg_next, uint32_t level, int threads);
Out bottom_up(const Csr& g, uint32_t* dist, const uint64_t* front, uint64_t* next, uint32_t level, int threads);
void queue_to_bitmap(const uint32_t* q, uint64_t size, uint64_t* bm, uint64_t n, int threads);
uint64_t bitmap_to_queue(const Csr& g, const uint64_t* bm, uint32_t* q, uint32_t* deg, int threads);
}  // namespace steps

// ---- CPU PageRank (pagerank_cpu.cpp:20-63) and acceptance (pagerank_cpu.hpp:37-84)
struct CpuPrResult { int iters = 0; double last_delta = 0; };
CpuPrResult pagerank(const Csr& in, const uint32_t* outdeg, double* r, double damping, int iters, double tol,
                     bool parallel, int threads);
struct PrCheck {
    double l1 = 0, max_rel_top = 0, sum = 0;
    bool ok() const;
};
PrCheck check_ranks(const std::vector<double>& ref, const float* got);
PrCheck check_ranks(const std::vector<double>& ref, const double* got);
bool gs_no_worse_than_jacobi(const PrCheck& gs, const PrCheck& jacobi);

// ---- Gauss-Seidel sweep order (pagerank_gpu.mm:88-192), host preprocessing
struct GsOrder {
    bool colored = false;
    // colored: sets[i] are contiguous ranges of NEW ids: [set_begin[i], set_begin[i+1])
    std::vector<uint32_t> set_begin;
    std::vector<uint32_t> new_id;   // new_id[old]
    Csr relabeled_in;               // in-CSR in new ids (colored mode only)
    std::vector<uint32_t> relabeled_outdeg;
    // natural order with pair deferral: main list then deferred list (old ids)
    std::vector<uint32_t> main_list, deferred_list;
};
// allow_colored = false skips the colouring and returns the natural order.
GsOrder gs_order(const Csr& in, const uint32_t* outdeg, uint32_t pair_max_outdeg, bool allow_colored = true);

}  // namespace ggpu
