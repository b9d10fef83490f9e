// Synthetic graphs.
//   path/star/complete/two_components/empty/grid2d/rmat/suite
//       src/graph/generators.cpp:8-88 (same edge lists, so the same CSRs)
//   gen_rmat's quadrant recursion follows generators.cpp:51-75 with the
//       probabilities as parameters; draws come from the counter form of
//       SplitMix64 so edges are generated in parallel yet equal the serial
//       stream of the reference.
//   gen_grid with keep < 1 and the *_c* workload graphs are new: they are the
//       inputs the benchmark configurations name (recipes in gen_named below).
#include "graph.hpp"

#include <omp.h>

#include <stdexcept>

namespace ggpu {

using EdgeList = std::vector<std::pair<uint32_t, uint32_t>>;

Csr gen_path(uint32_t n) {
    EdgeList e;
    for (uint32_t i = 0; i + 1 < n; ++i) e.push_back({i, i + 1});
    return from_edges(n, e, false);
}

Csr gen_star(uint32_t leaves) {
    EdgeList e;
    for (uint32_t i = 1; i <= leaves; ++i) e.push_back({0, i});
    return from_edges((uint64_t)leaves + 1, e, false);
}

Csr gen_complete(uint32_t n) {
    EdgeList e;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = i + 1; j < n; ++j) e.push_back({i, j});
    return from_edges(n, e, false);
}

namespace {
// Candidate edges of a rows x cols mesh in row-major order, right neighbour
// before lower neighbour (the order of generators.cpp:27-34). With keep < 1
// each candidate consumes one uniform draw and survives if the draw < keep.
void add_grid(EdgeList& e, uint32_t base, uint32_t rows, uint32_t cols, double keep, SplitMix64* rng) {
    for (uint32_t r = 0; r < rows; ++r)
        for (uint32_t c = 0; c < cols; ++c) {
            const uint32_t v = base + r * cols + c;
            if (c + 1 < cols && (!rng || rng->uniform() < keep)) e.push_back({v, v + 1});
            if (r + 1 < rows && (!rng || rng->uniform() < keep)) e.push_back({v, v + cols});
        }
}
}  // namespace

Csr gen_two_components() {
    EdgeList e;
    for (uint32_t i = 0; i + 1 < 300; ++i) e.push_back({i, i + 1});
    add_grid(e, 300, 10, 10, 1.0, nullptr);
    return from_edges(400, e, false);
}

Csr gen_empty(uint32_t n) { return from_edges(n, {}, false); }

Csr gen_grid(uint32_t rows, uint32_t cols, double keep, uint64_t seed) {
    EdgeList e;
    e.reserve(2ull * rows * cols);
    if (keep >= 1.0) {
        add_grid(e, 0, rows, cols, 1.0, nullptr);
    } else {
        SplitMix64 rng(seed);
        add_grid(e, 0, rows, cols, keep, &rng);
    }
    return from_edges((uint64_t)rows * cols, e, false);
}

Csr gen_rmat(uint32_t scale, uint32_t edgefactor, double a, double b, double c, bool directed, uint64_t seed) {
    if (scale == 0 || scale > 31) throw std::runtime_error("rmat: scale must be in [1, 31]");
    const uint64_t n = 1ull << scale, m = (uint64_t)edgefactor << scale;
    // Cumulative thresholds evaluated exactly as the reference does: a, a+b, a+b+c.
    const double t0 = a, t1 = a + b, t2 = a + b + c;
    EdgeList e(m);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)m; ++i) {
        uint32_t u = 0, v = 0;
        const uint64_t k0 = (uint64_t)i * scale;
        for (uint32_t bit = 0; bit < scale; ++bit) {
            const double r = SplitMix64::uniform_at(seed, k0 + bit);
            uint32_t ub = 0, vb = 0;
            if (r < t0) {
            } else if (r < t1) {
                vb = 1;
            } else if (r < t2) {
                ub = 1;
            } else {
                ub = 1;
                vb = 1;
            }
            u = (u << 1) | ub;
            v = (v << 1) | vb;
        }
        e[i] = {u, v};
    }
    return from_edges(n, e, directed);
}

std::vector<std::string> suite_names() {
    return {"path1000", "star100k", "complete200", "two_components", "empty1000", "grid100x100", "rmat12",
            "rmat12_directed"};
}

Csr gen_named(const std::string& name) {
    // The reference suite (generators.cpp:77-88).
    if (name == "path1000") return gen_path(1000);
    if (name == "star100k") return gen_star(100000);
    if (name == "complete200") return gen_complete(200);
    if (name == "two_components") return gen_two_components();
    if (name == "empty1000") return gen_empty(1000);
    if (name == "grid100x100") return gen_grid(100, 100, 1.0, 0);
    if (name == "rmat12") return gen_rmat(12, 16, 0.57, 0.19, 0.19, false, 42);
    if (name == "rmat12_directed") return gen_rmat(12, 16, 0.57, 0.19, 0.19, true, 43);
    // Workload graphs. RMAT: edges from seed s, then ids permuted by the
    // Fisher-Yates permutation of the same seed where the workload says so.
    if (name == "rmat16_c1") {
        Csr g = gen_rmat(16, 16, 0.57, 0.19, 0.19, false, 1);
        return permute(g, random_perm(g.n, 1));
    }
    if (name == "rmat22_c2") {
        Csr g = gen_rmat(22, 8, 0.57, 0.19, 0.19, false, 2);
        return permute(g, random_perm(g.n, 2));
    }
    if (name == "grid_c3") return largest_component(gen_grid(1024, 2048, 0.7, 3), nullptr);
    if (name == "rmat23_c5") return gen_rmat(23, 16, 0.60, 0.17, 0.17, false, 5);
    throw std::runtime_error("unknown synthetic graph '" + name + "'");
}

}  // namespace ggpu
