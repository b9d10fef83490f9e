// ref_capi.cpp -- a C entry-point shim over the REFERENCE's own CPU code.
//
// TEST INFRASTRUCTURE ONLY. oracle/Makefile compiles this file together with
// the reference's sources where they lie (src/graph/{csr,transform,generators,
// reorder,snap}.cpp and src/cpu/{bfs_cpu,bfs_steps,pagerank_cpu}.cpp under the
// reference checkout) into oracle/_ref/libggpu_ref.so. Nothing from the
// reference is copied here: this file only calls its public functions
// (graph::gen::*, graph::transpose/symmetrize/permute, graph::reorder::compute,
// graph::convertSnap,
// cpu::bfsSerial/bfsDO/bfsOmp, cpu::pagerankSerial/pagerankOmp) so tests can
// pin the oracle restatement against the real thing, and bench.py can time
// the reference's cpu_do / cpu_omp as the CPU baseline.
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "cpu/bfs_cpu.hpp"
#include "cpu/pagerank_cpu.hpp"
#include "graph/csr.hpp"
#include "graph/generators.hpp"
#include "graph/reorder.hpp"
#include "graph/snap.hpp"
#include "graph/transform.hpp"

namespace {

graph::CSR view(uint64_t n, uint64_t m, uint32_t flags, const uint32_t* rp, const uint32_t* ci) {
    graph::CSR g;
    g.n = n;
    g.m = m;
    g.flags = flags;
    g.row_ptr = rp;
    g.col_idx = ci;
    return g;
}

// Hands a CSR to the caller as malloc'ed arrays (freed with ref_free).
int export_csr(const graph::CSR& g, uint64_t* n, uint64_t* m, uint32_t* flags, uint32_t** rp, uint32_t** ci) {
    *n = g.n;
    *m = g.m;
    *flags = g.flags;
    *rp = (uint32_t*)std::malloc(4 * (g.n + 1));
    *ci = (uint32_t*)std::malloc(4 * (g.m ? g.m : 1));
    if (!*rp || !*ci) return 4;
    std::memcpy(*rp, g.row_ptr, 4 * (g.n + 1));
    if (g.m) std::memcpy(*ci, g.col_idx, 4 * g.m);
    return 0;
}

}  // namespace

extern "C" {

void ref_free(void* p) { std::free(p); }

int ref_gen_suite(const char* name, uint64_t* n, uint64_t* m, uint32_t* flags, uint32_t** rp, uint32_t** ci) {
    try {
        for (auto& [nm, g] : graph::gen::suite())
            if (nm == name) return export_csr(g, n, m, flags, rp, ci);
        return 1;
    } catch (const std::exception&) {
        return 2;
    }
}

int ref_gen_rmat(uint32_t scale, uint32_t ef, int directed, uint64_t seed, uint64_t* n, uint64_t* m, uint32_t* flags,
                 uint32_t** rp, uint32_t** ci) {
    try {
        return export_csr(graph::gen::rmat(scale, ef, directed != 0, seed), n, m, flags, rp, ci);
    } catch (const std::exception&) {
        return 2;
    }
}

int ref_symmetrize(uint64_t n, uint64_t m, uint32_t fl, const uint32_t* rp, const uint32_t* ci, uint64_t* on,
                   uint64_t* om, uint32_t* oflags, uint32_t** orp, uint32_t** oci) {
    return export_csr(graph::symmetrize(view(n, m, fl, rp, ci)), on, om, oflags, orp, oci);
}

int ref_transpose(uint64_t n, uint64_t m, uint32_t fl, const uint32_t* rp, const uint32_t* ci, uint64_t* on,
                  uint64_t* om, uint32_t* oflags, uint32_t** orp, uint32_t** oci) {
    return export_csr(graph::transpose(view(n, m, fl, rp, ci)), on, om, oflags, orp, oci);
}

int ref_random_perm(uint64_t n, uint64_t seed, uint32_t* perm) {
    graph::CSR g;  // compute() reads only n for "random"
    g.n = n;
    std::vector<uint32_t> p = graph::reorder::compute(g, "random", seed);
    std::memcpy(perm, p.data(), 4 * n);
    return 0;
}

int ref_order_perm(uint64_t n, uint64_t m, uint32_t fl, const uint32_t* rp, const uint32_t* ci, const char* ordering,
                   uint64_t seed, uint32_t* perm) {
    try {
        std::vector<uint32_t> p = graph::reorder::compute(view(n, m, fl, rp, ci), ordering, seed);
        if (n) std::memcpy(perm, p.data(), 4 * n);
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

void ref_locality(uint64_t n, uint64_t m, uint32_t fl, const uint32_t* rp, const uint32_t* ci, double* gap,
                  double* same) {
    const graph::reorder::Locality l = graph::reorder::locality(view(n, m, fl, rp, ci));
    *gap = l.meanGap;
    *same = l.sameBlock64;
}

// graph::convertSnap: edge list -> .bin file; stats[5] = lines, raw max id,
// n, entries before and after compaction. Returns 0, or 1 with the message in
// err (err_len bytes) on an error the converter reports.
int ref_convert_snap(const char* in, const char* out, int directed, uint64_t* stats, char* err, size_t err_len) {
    try {
        const graph::ConvertStats st = graph::convertSnap(in, out, directed != 0);
        stats[0] = st.lines;
        stats[1] = st.raw_max_id;
        stats[2] = st.n;
        stats[3] = st.m_before;
        stats[4] = st.m;
        return 0;
    } catch (const std::exception& e) {
        std::strncpy(err, e.what(), err_len - 1);
        err[err_len - 1] = 0;
        return 1;
    }
}

int ref_permute(uint64_t n, uint64_t m, uint32_t fl, const uint32_t* rp, const uint32_t* ci, const uint32_t* perm,
                uint64_t* on, uint64_t* om, uint32_t* oflags, uint32_t** orp, uint32_t** oci) {
    std::vector<uint32_t> p(perm, perm + n);
    return export_csr(graph::permute(view(n, m, fl, rp, ci), p), on, om, oflags, orp, oci);
}

void ref_bfs_serial(uint64_t n, uint64_t m, const uint32_t* rp, const uint32_t* ci, uint32_t src, uint32_t* dist) {
    cpu::bfsSerial(view(n, m, graph::kFlagSorted | graph::kFlagDedup, rp, ci), src, dist);
}

// cpu_do (variant 2) or cpu_omp (variant 1); trace arrays may be NULL.
uint64_t ref_bfs_cpu(int variant, uint64_t n, uint64_t m, const uint32_t* rp, const uint32_t* ci, uint32_t src,
                     double alpha, double beta, int threads, uint32_t* dist, char* dir, uint64_t* frontier,
                     uint64_t* edges, uint64_t cap) {
    cpu::BfsOptions o;
    o.alpha = alpha;
    o.beta = beta;
    o.threads = threads;
    std::vector<util::BfsLevel> tr;
    graph::CSR g = view(n, m, graph::kFlagSorted | graph::kFlagDedup, rp, ci);
    if (variant == 2) cpu::bfsDO(g, src, dist, o, dir ? &tr : nullptr);
    else cpu::bfsOmp(g, src, dist, o, dir ? &tr : nullptr);
    for (size_t i = 0; dir && i < tr.size() && i < cap; ++i) {
        dir[i] = tr[i].direction;
        frontier[i] = tr[i].frontier;
        if (edges) edges[i] = tr[i].edges;
    }
    return tr.size();
}

int ref_pagerank(uint64_t n, uint64_t m, const uint32_t* in_rp, const uint32_t* in_ci, const uint32_t* outdeg,
                 double damping, int iters, double tol, int parallel, int threads, double* r, double* last_delta) {
    cpu::PrOptions o;
    o.damping = damping;
    o.iters = iters;
    o.tol = tol;
    o.threads = threads;
    graph::CSR in = view(n, m, 0, in_rp, in_ci);
    cpu::PrResult res = parallel ? cpu::pagerankOmp(in, outdeg, r, o) : cpu::pagerankSerial(in, outdeg, r, o);
    if (last_delta) *last_delta = res.lastDelta;
    return res.iters;
}

void ref_check_ranks(uint64_t n, const double* ref, const float* got, double* l1, double* max_rel_top, double* sum) {
    std::vector<double> r(ref, ref + n);
    cpu::PrCheck c = cpu::checkRanks(r, got);
    *l1 = c.l1;
    *max_rel_top = c.maxRelTop;
    *sum = c.sum;
}

int ref_ranks_ok(double l1, double rel, double sum) {
    cpu::PrCheck c;
    c.l1 = l1;
    c.maxRelTop = rel;
    c.sum = sum;
    return c.ok() ? 1 : 0;
}

int ref_gs_no_worse(double gl1, double grel, double gsum, double jl1, double jrel, double jsum) {
    cpu::PrCheck g, j;
    g.l1 = gl1;
    g.maxRelTop = grel;
    g.sum = gsum;
    j.l1 = jl1;
    j.maxRelTop = jrel;
    j.sum = jsum;
    return cpu::gsNoWorseThanJacobi(g, j) ? 1 : 0;
}

}  // extern "C"
