// CPU PageRank, acceptance checks and the Gauss-Seidel sweep order.
//   pagerank        src/cpu/pagerank_cpu.cpp:20-53 (definition: pagerank_cpu.hpp:4-8)
//   check_ranks     src/cpu/pagerank_cpu.hpp:37-60
//   gs_no_worse...  src/cpu/pagerank_cpu.hpp:80-84
//   gs_order        src/metal/pagerank_gpu.mm:88-192 (Jones-Plassmann colouring with the same
//                   hash priority, 8 rounds, 95% coverage rule, colour-major relabeling; else
//                   natural order with greedy mutual-pair deferral)
#include "graph.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <numeric>

namespace ggpu {

CpuPrResult pagerank(const Csr& in, const uint32_t* outdeg, double* r, double damping, int iters, double tol,
                     bool parallel, int threads) {
    const uint64_t n = in.n;
    CpuPrResult res;
    if (n == 0) return res;
    const int nt = parallel ? (threads > 0 ? threads : omp_get_max_threads()) : 1;
    const double base = (1.0 - damping) / (double)n;
    std::vector<double> contrib(n), next(n);
    for (uint64_t v = 0; v < n; ++v) r[v] = 1.0 / (double)n;
    for (int it = 0; it < iters; ++it) {
        double dangling = 0;
#pragma omp parallel for num_threads(nt) schedule(static) reduction(+ : dangling) if (parallel)
        for (int64_t u = 0; u < (int64_t)n; ++u) {
            if (outdeg[u]) {
                contrib[u] = r[u] / outdeg[u];
            } else {
                contrib[u] = 0;
                dangling += r[u];
            }
        }
        const double add = base + damping * dangling / (double)n;
        double delta = 0;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 256) reduction(+ : delta) if (parallel)
        for (int64_t v = 0; v < (int64_t)n; ++v) {
            double s = 0;
            for (const uint32_t* p = in.nbr_begin(v); p != in.nbr_end(v); ++p) s += contrib[*p];
            next[v] = add + damping * s;
            delta += std::fabs(next[v] - r[v]);
        }
        std::copy(next.begin(), next.end(), r);
        res.iters = it + 1;
        res.last_delta = delta;
        if (tol > 0 && delta < tol) break;
    }
    return res;
}

bool PrCheck::ok() const { return l1 < 1e-4 && max_rel_top < 1e-3 && std::fabs(sum - 1.0) < 1e-4; }

namespace {
template <class T>
PrCheck check_impl(const std::vector<double>& ref, const T* got) {
    PrCheck c;
    const size_t n = ref.size();
    for (size_t v = 0; v < n; ++v) {
        c.l1 += std::fabs(ref[v] - (double)got[v]);
        c.sum += (double)got[v];
    }
    std::vector<uint32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    const size_t k = std::min<size_t>(1000, n);
    std::partial_sort(idx.begin(), idx.begin() + k, idx.end(), [&](uint32_t a, uint32_t b) { return ref[a] > ref[b]; });
    for (size_t i = 0; i < k; ++i) {
        const uint32_t v = idx[i];
        c.max_rel_top = std::max(c.max_rel_top, std::fabs(ref[v] - (double)got[v]) / ref[v]);
    }
    return c;
}
}  // namespace

PrCheck check_ranks(const std::vector<double>& ref, const float* got) { return check_impl(ref, got); }
PrCheck check_ranks(const std::vector<double>& ref, const double* got) { return check_impl(ref, got); }

bool gs_no_worse_than_jacobi(const PrCheck& gs, const PrCheck& jacobi) {
    const double budget = std::max(1e-4, jacobi.l1);
    return gs.l1 <= budget && std::fabs(gs.sum - 1.0) <= budget &&
           gs.max_rel_top <= std::max(1e-3, jacobi.max_rel_top * (1 + 1e-4));
}

GsOrder gs_order(const Csr& in, const uint32_t* outdeg, uint32_t pair_max_outdeg, bool allow_colored) {
    constexpr int kRounds = 8;
    constexpr double kCoverage = 0.95;
    const uint32_t n = (uint32_t)in.n;
    GsOrder o;
    // Hash priority, ties impossible: the id is the low word.
    auto prio = [](uint32_t v) -> uint64_t {
        uint64_t x = (uint64_t)v * 0x9E3779B97F4A7C15ull;
        x ^= x >> 31;
        return (x << 32) | v;
    };
    std::vector<uint8_t> color(n, 0xFF);
    std::vector<std::vector<uint32_t>> sets;
    uint64_t covered = 0;
    for (int c = 0; allow_colored && c < kRounds && covered < n; ++c) {
        std::vector<uint8_t> pick(n, 0);
        // A vertex joins this round's set if it beats every uncoloured in-neighbour
        // (decided on the colouring state at the start of the round).
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t v = 0; v < (int64_t)n; ++v) {
            if (color[v] != 0xFF) continue;
            bool best = true;
            for (uint32_t e = in.rp[v]; e < in.rp[v + 1] && best; ++e) {
                const uint32_t u = in.ci[e];
                best = u == (uint32_t)v || color[u] != 0xFF || prio(u) < prio((uint32_t)v);
            }
            pick[v] = best;
        }
        std::vector<uint32_t> set;
        for (uint32_t v = 0; v < n; ++v)
            if (pick[v]) set.push_back(v);
        for (uint32_t v : set) color[v] = (uint8_t)c;
        covered += set.size();
        sets.push_back(std::move(set));
    }
    o.colored = n > 0 && (double)covered >= kCoverage * n;
    if (o.colored) {
        std::vector<uint32_t> rest;
        for (uint32_t v = 0; v < n; ++v)
            if (color[v] == 0xFF) rest.push_back(v);
        if (!rest.empty()) sets.push_back(std::move(rest));
        o.new_id.assign(n, 0);
        std::vector<uint32_t> order;
        order.reserve(n);
        o.set_begin.push_back(0);
        for (auto& s : sets) {
            if (s.empty()) continue;
            for (uint32_t v : s) {
                o.new_id[v] = (uint32_t)order.size();
                order.push_back(v);
            }
            o.set_begin.push_back((uint32_t)order.size());
        }
        Csr& r = o.relabeled_in;
        r.n = n;
        r.m = in.m;
        r.flags = in.flags;
        r.rp.assign(n + 1, 0);
        for (uint32_t i = 0; i < n; ++i) r.rp[i + 1] = r.rp[i] + in.deg(order[i]);
        r.ci.resize(in.m);
        o.relabeled_outdeg.resize(n);
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            const uint32_t v = order[i];
            o.relabeled_outdeg[i] = outdeg[v];
            uint32_t* w = r.ci.data() + r.rp[i];
            for (uint32_t e = in.rp[v]; e < in.rp[v + 1]; ++e) *w++ = o.new_id[in.ci[e]];
            std::sort(r.ci.data() + r.rp[i], w);  // ascending gathers, as in the original lists
        }
        return o;
    }
    // Natural order: defer one vertex of each mutual pair whose out-degrees are
    // both <= K, greedily in id order, never both partners.
    std::vector<uint8_t> deferred(n, 0);
    if (pair_max_outdeg > 0) {
        for (uint32_t v = 0; v < n; ++v) {
            if (outdeg[v] == 0 || outdeg[v] > pair_max_outdeg) continue;
            bool partner = false, partner_deferred = false;
            for (uint32_t e = in.rp[v]; e < in.rp[v + 1] && !partner_deferred; ++e) {
                const uint32_t u = in.ci[e];  // u -> v
                if (u >= v || outdeg[u] == 0 || outdeg[u] > pair_max_outdeg) continue;
                if (std::binary_search(in.nbr_begin(u), in.nbr_end(u), v)) {  // v -> u as well
                    partner = true;
                    partner_deferred = deferred[u];
                }
            }
            deferred[v] = partner && !partner_deferred;
        }
    }
    for (uint32_t v = 0; v < n; ++v) (deferred[v] ? o.deferred_list : o.main_list).push_back(v);
    return o;
}

}  // namespace ggpu
