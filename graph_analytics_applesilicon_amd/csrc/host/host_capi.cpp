// extern "C" surface of the host graph library (include/ggpu_graph.h).
// Nothing throws across the boundary: exceptions become status codes and a
// thread-local message.
#include <cstring>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "ggpu_graph.h"
#include "graph.hpp"

struct ggpu_graph {
    ggpu::Csr g;
};

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return GGPU_OK;
    } catch (const std::bad_alloc&) {
        return fail(GGPU_ENOMEM, "out of host memory");
    } catch (const std::invalid_argument& e) {
        return fail(GGPU_EINVAL, e.what());
    } catch (const std::exception& e) {
        const std::string m = e.what();
        const bool io = m.find("cannot") != std::string::npos || m.find("magic") != std::string::npos ||
                        m.find("size does not match") != std::string::npos || m.find("short") != std::string::npos ||
                        m.find("too small") != std::string::npos || m.find("version") != std::string::npos;
        return fail(io ? GGPU_EIO : GGPU_EINVAL, m);
    }
}

int wrap(ggpu::Csr&& g, ggpu_graph** out) {
    *out = new ggpu_graph{std::move(g)};
    return GGPU_OK;
}
}  // namespace

extern "C" {

const char* ggpu_host_last_error(void) { return g_err.c_str(); }

int ggpu_graph_load(const char* path, ggpu_graph** out) {
    if (!path || !out) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] { wrap(ggpu::load_bin(path), out); });
}

int ggpu_graph_save(const ggpu_graph* g, const char* path) {
    if (!g || !path) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] { ggpu::save_bin(g->g, path); });
}

int ggpu_graph_from_arrays(uint64_t n, uint64_t m, const uint32_t* rp, const uint32_t* ci, uint32_t flags,
                           ggpu_graph** out) {
    if (!rp || (m && !ci) || !out) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] { wrap(ggpu::from_arrays(n, m, rp, ci, flags), out); });
}

int ggpu_graph_from_edges(uint64_t n, uint64_t num_edges, const uint32_t* src, const uint32_t* dst, int directed,
                          ggpu_graph** out) {
    if ((num_edges && (!src || !dst)) || !out) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] {
        std::vector<std::pair<uint32_t, uint32_t>> e(num_edges);
        for (uint64_t i = 0; i < num_edges; ++i) {
            if (src[i] >= n || dst[i] >= n) throw std::invalid_argument("edge endpoint out of range");
            e[i] = {src[i], dst[i]};
        }
        wrap(ggpu::from_edges(n, e, directed != 0), out);
    });
}

void ggpu_graph_free(ggpu_graph* g) { delete g; }
uint64_t ggpu_graph_n(const ggpu_graph* g) { return g ? g->g.n : 0; }
uint64_t ggpu_graph_m(const ggpu_graph* g) { return g ? g->g.m : 0; }
uint32_t ggpu_graph_flags(const ggpu_graph* g) { return g ? g->g.flags : 0; }
const uint32_t* ggpu_graph_row_ptr(const ggpu_graph* g) { return g ? g->g.rp.data() : nullptr; }
const uint32_t* ggpu_graph_col_idx(const ggpu_graph* g) { return g ? g->g.ci.data() : nullptr; }

int ggpu_graph_validate(const ggpu_graph* g, char* buf, size_t buflen) {
    if (!g) return fail(GGPU_EINVAL, "null graph");
    const std::string e = ggpu::validate(g->g);
    if (buf && buflen) {
        std::strncpy(buf, e.c_str(), buflen - 1);
        buf[buflen - 1] = 0;
    }
    return e.empty() ? GGPU_OK : fail(GGPU_EINVAL, e);
}

int ggpu_graph_equal(const ggpu_graph* a, const ggpu_graph* b) { return a && b && ggpu::equal(a->g, b->g); }

int ggpu_graph_transpose(const ggpu_graph* g, ggpu_graph** out) {
    if (!g || !out) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] { wrap(ggpu::transpose(g->g), out); });
}

int ggpu_graph_symmetrize(const ggpu_graph* g, ggpu_graph** out) {
    if (!g || !out) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] { wrap(ggpu::symmetrize(g->g), out); });
}

int ggpu_graph_permute(const ggpu_graph* g, const uint32_t* perm, ggpu_graph** out) {
    if (!g || !perm || !out) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] {
        std::vector<uint32_t> p(perm, perm + g->g.n);
        std::vector<uint8_t> seen(g->g.n, 0);
        for (uint32_t x : p) {
            if (x >= g->g.n || seen[x]) throw std::invalid_argument("not a permutation");
            seen[x] = 1;
        }
        wrap(ggpu::permute(g->g, p), out);
    });
}

int ggpu_random_perm(uint64_t n, uint64_t seed, uint32_t* perm_out) {
    if (!perm_out && n) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] {
        std::vector<uint32_t> p = ggpu::random_perm(n, seed);
        std::memcpy(perm_out, p.data(), 4 * n);
    });
}

int ggpu_order_perm(const ggpu_graph* sym, const char* ordering, uint64_t seed, uint32_t* perm_out) {
    if (!sym || !ordering || (!perm_out && sym->g.n)) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] {
        std::vector<uint32_t> p = ggpu::order_perm(sym->g, ordering, seed);
        if (!p.empty()) std::memcpy(perm_out, p.data(), 4 * p.size());
    });
}

int ggpu_graph_from_snap(const char* path, int directed, ggpu_graph** out, ggpu_snap_stats* stats) {
    if (!path || !out) return fail(GGPU_EINVAL, "null argument");
    *out = nullptr;
    return guarded([&] {
        ggpu::SnapStats st;
        auto* g = new ggpu_graph{ggpu::load_snap(path, directed != 0, &st)};
        *out = g;
        if (stats) *stats = ggpu_snap_stats{st.lines, st.raw_max_id, st.n, st.m_before, st.m};
    });
}

int ggpu_graph_stats_compute(const ggpu_graph* g, ggpu_graph_stats* out) {
    if (!g || !out) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] {
        const ggpu::GraphStats t = ggpu::graph_stats(g->g);
        ggpu_graph_stats o{};
        o.n = t.n;
        o.m = t.m;
        o.undirected_edges = t.undirected_edges;
        o.min_deg = t.min_deg;
        o.max_deg = t.max_deg;
        o.max_in_deg = t.max_in_deg;
        o.approx_diameter = t.approx_diameter;
        o.mean_deg = t.mean_deg;
        o.isolated = t.isolated;
        o.num_cc = t.num_cc;
        o.largest_cc = t.largest_cc;
        *out = o;
    });
}

int ggpu_degree_histogram(const ggpu_graph* g, int kind, uint64_t* hist, uint32_t* buckets) {
    if (!g || !hist || !buckets) return fail(GGPU_EINVAL, "null argument");
    if (kind < 0 || kind > 2) return fail(GGPU_EINVAL, "kind must be 0 (stored), 1 (in) or 2 (symmetrized)");
    return guarded([&] {
        const ggpu::Csr& c = g->g;
        std::vector<uint32_t> deg(c.n, 0);
        if (kind == 0) {
            for (uint64_t v = 0; v < c.n; ++v) deg[v] = c.deg(v);
        } else if (kind == 1) {
            for (uint64_t e = 0; e < c.m; ++e) ++deg[c.ci[e]];
        } else {
            const ggpu::Csr s = ggpu::symmetrize(c);
            for (uint64_t v = 0; v < c.n; ++v) deg[v] = s.deg(v);
        }
        const std::vector<uint64_t> h = ggpu::log2_histogram(deg);
        std::memcpy(hist, h.data(), h.size() * sizeof(uint64_t));
        *buckets = (uint32_t)h.size();
    });
}

int ggpu_graph_save_sg(const ggpu_graph* g, const char* path, int sym) {
    if (!g || !path) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] { ggpu::save_sg(g->g, path, sym != 0); });
}

int ggpu_graph_locality(const ggpu_graph* g, double* mean_gap, double* same_block64) {
    if (!g || !mean_gap || !same_block64) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] {
        const ggpu::Locality l = ggpu::locality(g->g);
        *mean_gap = l.mean_gap;
        *same_block64 = l.same_block64;
    });
}

int ggpu_graph_largest_component(const ggpu_graph* g, ggpu_graph** out, uint32_t* old_ids_out, uint64_t* n_out) {
    if (!g || !out) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] {
        std::vector<uint32_t> old;
        ggpu::Csr s = ggpu::largest_component(g->g, &old);
        if (old_ids_out) std::memcpy(old_ids_out, old.data(), 4 * old.size());
        if (n_out) *n_out = s.n;
        wrap(std::move(s), out);
    });
}

int ggpu_gen_named(const char* name, ggpu_graph** out) {
    if (!name || !out) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] {
        try {
            wrap(ggpu::gen_named(name), out);
        } catch (const std::runtime_error& e) {
            throw std::invalid_argument(e.what());
        }
    });
}

int ggpu_gen_rmat(uint32_t scale, uint32_t edgefactor, double a, double b, double c, int directed, uint64_t seed,
                  ggpu_graph** out) {
    if (!out) return fail(GGPU_EINVAL, "null argument");
    if (scale == 0 || scale > 30) return fail(GGPU_EINVAL, "rmat scale must be in [1, 30]");
    if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) return fail(GGPU_EINVAL, "rmat probabilities invalid");
    return guarded([&] { wrap(ggpu::gen_rmat(scale, edgefactor, a, b, c, directed != 0, seed), out); });
}

int ggpu_gen_grid(uint32_t rows, uint32_t cols, double keep, uint64_t seed, ggpu_graph** out) {
    if (!out) return fail(GGPU_EINVAL, "null argument");
    if ((uint64_t)rows * cols >= (1ull << 31)) return fail(GGPU_EINVAL, "grid too large");
    return guarded([&] { wrap(ggpu::gen_grid(rows, cols, keep, seed), out); });
}

int ggpu_pick_sources(const ggpu_graph* g, uint32_t count, uint64_t seed, int any_component, uint32_t* out,
                      uint32_t* n_out) {
    if (!g || (!out && count) || !n_out) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] {
        std::vector<uint32_t> s = ggpu::pick_sources(g->g, count, seed, any_component != 0);
        if (!s.empty()) std::memcpy(out, s.data(), 4 * s.size());
        *n_out = (uint32_t)s.size();
    });
}

void ggpu_cpu_bfs_opts_default(ggpu_cpu_bfs_opts* o) {
    if (!o) return;
    o->threads = 0;
    o->schedule = 1;
    o->alpha = 15.0;
    o->beta = 18.0;
}

int ggpu_cpu_bfs(const ggpu_graph* g, const char* variant, uint32_t source, const ggpu_cpu_bfs_opts* opt,
                 uint32_t* dist, ggpu_level* trace, size_t trace_cap, size_t* trace_len) {
    if (!g || !variant || !dist) return fail(GGPU_EINVAL, "null argument");
    if (source >= g->g.n) return fail(GGPU_EINVAL, "source out of range");
    ggpu_cpu_bfs_opts d;
    ggpu_cpu_bfs_opts_default(&d);
    const ggpu_cpu_bfs_opts& o = opt ? *opt : d;
    if (o.schedule < 0 || o.schedule > 2) return fail(GGPU_EINVAL, "schedule must be 0, 1 or 2");
    return guarded([&] {
        ggpu::CpuBfsOpts co;
        co.threads = o.threads;
        co.schedule = (ggpu::Schedule)o.schedule;
        co.alpha = o.alpha;
        co.beta = o.beta;
        std::vector<ggpu_level> tr;
        const std::string v = variant;
        if (v == "cpu_serial") ggpu::bfs_serial(g->g, source, dist);
        else if (v == "cpu_omp") ggpu::bfs_omp(g->g, source, dist, co, trace ? &tr : nullptr);
        else if (v == "cpu_do") ggpu::bfs_do(g->g, source, dist, co, trace ? &tr : nullptr);
        else throw std::invalid_argument("unknown CPU BFS variant '" + v + "'");
        if (trace) {
            const size_t k = std::min(tr.size(), trace_cap);
            std::memcpy(trace, tr.data(), k * sizeof(ggpu_level));
        }
        if (trace_len) *trace_len = tr.size();
    });
}

int ggpu_cpu_pagerank(const ggpu_graph* in, const uint32_t* outdeg, double damping, int iters, double tol,
                      int parallel, int threads, double* r_out, int* iters_out, double* last_delta_out) {
    if (!in || (!outdeg && in->g.n) || (!r_out && in->g.n)) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] {
        ggpu::CpuPrResult r = ggpu::pagerank(in->g, outdeg, r_out, damping, iters, tol, parallel != 0, threads);
        if (iters_out) *iters_out = r.iters;
        if (last_delta_out) *last_delta_out = r.last_delta;
    });
}

void ggpu_check_ranks(uint64_t n, const double* ref, const float* got, double* l1, double* max_rel_top,
                      double* sum) {
    std::vector<double> r(ref, ref + n);
    ggpu::PrCheck c = ggpu::check_ranks(r, got);
    if (l1) *l1 = c.l1;
    if (max_rel_top) *max_rel_top = c.max_rel_top;
    if (sum) *sum = c.sum;
}

}  // extern "C"

// ---- test hooks (include/ggpu_testing.h)
#include "ggpu_testing.h"
#include "pr_tail_plan.hpp"
#include "pr_tiling.hpp"

extern "C" int ggpu_test_pr_tiled_emulate(uint64_t n, uint64_t m, const uint32_t* in_row_ptr, const uint32_t* in_col_idx,
                                          const uint32_t* params, const float* contrib, double* sums, uint64_t* stats) {
    if (!in_row_ptr || (m && !in_col_idx) || !params || !contrib || !sums || !stats) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] {
        const ggpu::Csr in = ggpu::from_arrays(n, m, in_row_ptr, in_col_idx, GGPU_FLAG_SORTED);
        ggpu::TiledParams p;
        p.min_indeg = params[0];
        p.tile = params[1];
        p.nnz_cap = params[2];
        p.rows_cap = params[3];
        p.slots = params[4];
        p.dense_min_arcs = params[5];
        p.task_arcs = params[6];
        p.task_items = params[7];
        p.groups = params[8];
        const ggpu::TiledPlan plan = ggpu::plan_tiled(in, p);
        ggpu::emulate_tiled(in, p, plan, contrib, sums);
        stats[0] = plan.R;
        stats[1] = plan.K;
        stats[2] = plan.dense_arcs;
        stats[3] = plan.cold_arcs;
        stats[4] = plan.untiled_arcs;
        stats[5] = plan.items.size();
        stats[6] = plan.tasks.size();
        stats[7] = plan.segs.size();
    });
}

extern "C" int ggpu_test_pr_tail_emulate(uint64_t n, uint64_t m, const uint32_t* in_row_ptr, const uint32_t* in_col_idx,
                                         const uint32_t* params, const float* contrib, double* sums, uint64_t* stats) {
    if (!in_row_ptr || (m && !in_col_idx) || !params || !contrib || !sums || !stats) return fail(GGPU_EINVAL, "null argument");
    return guarded([&] {
        const ggpu::Csr in = ggpu::from_arrays(n, m, in_row_ptr, in_col_idx, GGPU_FLAG_SORTED);
        ggpu::TailParams p;
        p.min_indeg = params[0];
        p.tail_begin = params[1];
        p.tile = params[2];
        p.nnz_cap = params[3];
        p.pieces_cap = params[4];
        p.band_end = params[5];
        p.band_tile = params[6];
        const ggpu::TailPlan plan = ggpu::plan_tail(in, p);
        ggpu::emulate_tail(in, p, plan, contrib, sums);
        stats[0] = plan.R;
        stats[1] = plan.T;
        stats[2] = plan.tci.size();
        stats[3] = plan.head_arcs;
        stats[4] = plan.pdst.size();
        stats[5] = plan.items.size();
        stats[6] = plan.ovptr.back();
        stats[7] = plan.T ? ggpu::tail_band_tiles(p, n) : 0;
    });
}
