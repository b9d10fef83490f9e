// Vertex orderings for --ordering (src/graph/reorder.cpp:10-125): each returns
// perm with perm[old] = new, computed on the symmetrized graph.
//   original  identity
//   random    Fisher-Yates with the SplitMix64 stream (random_perm)
//   degree    descending degree, ties by id (hubs first)
//   rcm       reverse Cuthill-McKee per component from a pseudo-peripheral
//             start, reversed at the end
// and the locality figures of a labeling (mean |u - v| over the stored arcs,
// the share of arcs inside one 64-vertex block).
#include <algorithm>
#include <numeric>
#include <stdexcept>

#include "graph.hpp"

namespace ggpu {

namespace {

// Breadth-first order from `root` over vertices not yet in `seen`, appended
// to `order`; the vertices found from one vertex are appended in ascending
// (degree, id). Returns the position in `order` where the last level starts.
size_t cuthill_mckee(const Csr& g, uint32_t root, std::vector<uint8_t>& seen, std::vector<uint32_t>& order,
                     std::vector<uint32_t>& found) {
    const auto by_degree_then_id = [&](uint32_t a, uint32_t b) {
        const uint32_t da = g.deg(a), db = g.deg(b);
        return da != db ? da < db : a < b;
    };
    size_t level_begin = order.size();
    seen[root] = 1;
    order.push_back(root);
    size_t last_level = level_begin;
    while (level_begin < order.size()) {
        const size_t level_end = order.size();
        last_level = level_begin;
        for (size_t i = level_begin; i < level_end; ++i) {
            found.clear();
            for (const uint32_t* p = g.nbr_begin(order[i]); p != g.nbr_end(order[i]); ++p)
                if (!seen[*p]) {
                    seen[*p] = 1;
                    found.push_back(*p);
                }
            std::sort(found.begin(), found.end(), by_degree_then_id);
            order.insert(order.end(), found.begin(), found.end());
        }
        level_begin = level_end;
    }
    return last_level;
}

std::vector<uint32_t> degree_perm(const Csr& g) {
    std::vector<uint32_t> by(g.n);
    std::iota(by.begin(), by.end(), 0u);
    std::stable_sort(by.begin(), by.end(), [&](uint32_t a, uint32_t b) { return g.deg(a) > g.deg(b); });
    std::vector<uint32_t> perm(g.n);
    for (uint32_t i = 0; i < g.n; ++i) perm[by[i]] = i;
    return perm;
}

std::vector<uint32_t> rcm_perm(const Csr& g) {
    const uint32_t n = (uint32_t)g.n;
    // components are entered from their lowest-degree vertex (stable in id)
    std::vector<uint32_t> roots(n);
    std::iota(roots.begin(), roots.end(), 0u);
    std::stable_sort(roots.begin(), roots.end(), [&](uint32_t a, uint32_t b) { return g.deg(a) < g.deg(b); });
    std::vector<uint8_t> seen(n, 0), probed(n, 0);
    std::vector<uint32_t> order, probe, found;
    order.reserve(n);
    for (uint32_t r : roots) {
        if (seen[r]) continue;
        // one probing traversal: restart from the first lowest-degree vertex
        // of its farthest level (a pseudo-peripheral vertex). Components are
        // disjoint, so `probed` never needs clearing.
        probe.clear();
        const size_t last = cuthill_mckee(g, r, probed, probe, found);
        uint32_t start = probe[last];
        for (size_t i = last + 1; i < probe.size(); ++i)
            if (g.deg(probe[i]) < g.deg(start)) start = probe[i];
        cuthill_mckee(g, start, seen, order, found);
    }
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; ++i) perm[order[i]] = n - 1 - i;
    return perm;
}

}  // namespace

std::vector<uint32_t> order_perm(const Csr& sym, const std::string& ordering, uint64_t seed) {
    if (ordering == "original") {
        std::vector<uint32_t> p(sym.n);
        std::iota(p.begin(), p.end(), 0u);
        return p;
    }
    if (ordering == "random") return random_perm(sym.n, seed);
    if (ordering == "degree") return degree_perm(sym);
    if (ordering == "rcm") return rcm_perm(sym);
    throw std::invalid_argument("unknown ordering '" + ordering + "' (original|random|degree|rcm)");
}

Locality locality(const Csr& g) {
    Locality l;
    if (!g.m) return l;
    double gap = 0;
    uint64_t same = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : gap, same)
    for (int64_t u = 0; u < (int64_t)g.n; ++u)
        for (const uint32_t* p = g.nbr_begin(u); p != g.nbr_end(u); ++p) {
            gap += (double)(u > (int64_t)*p ? (uint64_t)u - *p : (uint64_t)*p - u);
            same += ((uint64_t)u >> 6) == (*p >> 6);
        }
    l.mean_gap = gap / (double)g.m;
    l.same_block64 = (double)same / (double)g.m;
    return l;
}

}  // namespace ggpu
