// CSR format, construction and transforms.
//   format      src/graph/csr.hpp:20-27, csr.cpp:30-84 (header layout kept bit for bit)
//   validate    csr.cpp:86-104
//   from_edges  csr.cpp:112-163 (same canonical result; built in parallel here)
//   transforms  transform.cpp:9-76
//   random_perm reorder.cpp:20-25 (Fisher-Yates with SplitMix64::below)
#include "graph.hpp"

#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>

namespace ggpu {

namespace {

struct BinHeader {
    char magic[8];
    uint32_t version;
    uint32_t flags;
    uint64_t n;
    uint64_t m;
};
static_assert(sizeof(BinHeader) == 32, "GGPUCSR1 header is 32 bytes");
const char kMagic[8] = {'G', 'G', 'P', 'U', 'C', 'S', 'R', '1'};

// Exclusive prefix sum of counts[0..n) into rp[0..n], rp[n] = total.
void prefix_into(const std::vector<uint32_t>& counts, std::vector<uint32_t>& rp) {
    const uint64_t n = counts.size();
    rp.assign(n + 1, 0);
    uint64_t acc = 0;
    for (uint64_t v = 0; v < n; ++v) {
        rp[v] = (uint32_t)acc;
        acc += counts[v];
    }
    if (acc >= (1ull << 32)) throw std::runtime_error("graph has 2^32 or more arcs; ids are 32-bit");
    rp[n] = (uint32_t)acc;
}

// Sorts each list, drops duplicates and self-loops, and packs the result into
// a fresh col_idx (the compaction of csr.cpp:112-144, parallel in both phases).
void canonicalize(uint64_t n, std::vector<uint32_t>& rp, std::vector<uint32_t>& ci) {
    std::vector<uint32_t> keep(n);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
        uint32_t* b = ci.data() + rp[v];
        uint32_t* e = ci.data() + rp[v + 1];
        std::sort(b, e);
        uint32_t* w = b;
        for (uint32_t* p = b; p != e; ++p) {
            if (*p == (uint32_t)v) continue;
            if (w != b && w[-1] == *p) continue;
            *w++ = *p;
        }
        keep[v] = (uint32_t)(w - b);
    }
    std::vector<uint32_t> nrp;
    prefix_into(keep, nrp);
    std::vector<uint32_t> nci(nrp[n]);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t v = 0; v < (int64_t)n; ++v)
        if (keep[v]) std::memcpy(nci.data() + nrp[v], ci.data() + rp[v], 4ull * keep[v]);
    rp.swap(nrp);
    ci.swap(nci);
}

}  // namespace

void canonicalize_lists(uint64_t n, std::vector<uint32_t>& rp, std::vector<uint32_t>& ci) { canonicalize(n, rp, ci); }

void save_bin(const Csr& g, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    BinHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = 1;
    h.flags = g.flags;
    h.n = g.n;
    h.m = g.m;
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && std::fwrite(g.rp.data(), 4, g.n + 1, f) == g.n + 1 &&
              (g.m == 0 || std::fwrite(g.ci.data(), 4, g.m, f) == g.m);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw std::runtime_error("short write to " + path);
}

Csr load_bin(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::fseek(f, 0, SEEK_END);
    long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    BinHeader h{};
    if (len < (long)sizeof h || std::fread(&h, sizeof h, 1, f) != 1) {
        std::fclose(f);
        throw std::runtime_error(path + ": too small");
    }
    if (std::memcmp(h.magic, kMagic, 8) != 0) { std::fclose(f); throw std::runtime_error(path + ": bad magic"); }
    if (h.version != 1) { std::fclose(f); throw std::runtime_error(path + ": unsupported version"); }
    const uint64_t expect = sizeof h + 4 * (h.n + 1) + 4 * h.m;
    if (expect != (uint64_t)len) { std::fclose(f); throw std::runtime_error(path + ": size does not match header"); }
    Csr g;
    g.n = h.n;
    g.m = h.m;
    g.flags = h.flags;
    g.rp.resize(g.n + 1);
    g.ci.resize(g.m);
    bool ok = std::fread(g.rp.data(), 4, g.n + 1, f) == g.n + 1 && (g.m == 0 || std::fread(g.ci.data(), 4, g.m, f) == g.m);
    std::fclose(f);
    if (!ok) throw std::runtime_error(path + ": short read");
    return g;
}

std::string validate(const Csr& g) {
    if (g.m >= (1ull << 32)) return "m does not fit in uint32";
    if (g.rp.size() != g.n + 1) return "row_ptr has the wrong length";
    if (g.rp[0] != 0) return "row_ptr[0] != 0";
    if (g.rp[g.n] != g.m) return "row_ptr[n] != m";
    const bool strict = (g.flags & GGPU_FLAG_SORTED) && (g.flags & GGPU_FLAG_DEDUP);
    for (uint64_t v = 0; v < g.n; ++v) {
        const uint32_t b = g.rp[v], e = g.rp[v + 1];
        if (e < b) return "row_ptr decreases at v=" + std::to_string(v);
        for (uint32_t i = b; i < e; ++i) {
            const uint32_t u = g.ci[i];
            if (u >= g.n) return "col_idx out of range at v=" + std::to_string(v);
            if (strict) {
                if (u == v) return "self-loop at v=" + std::to_string(v);
                if (i > b && g.ci[i - 1] >= u) return "list not strictly sorted at v=" + std::to_string(v);
            }
        }
    }
    return "";
}

bool equal(const Csr& a, const Csr& b) {
    return a.n == b.n && a.m == b.m && a.flags == b.flags && a.rp == b.rp && a.ci == b.ci;
}

Csr from_arrays(uint64_t n, uint64_t m, const uint32_t* rp, const uint32_t* ci, uint32_t flags) {
    Csr g;
    g.n = n;
    g.m = m;
    g.flags = flags;
    g.rp.assign(rp, rp + n + 1);
    g.ci.assign(ci, ci + m);
    return g;
}

Csr from_edges(uint64_t n, const std::vector<std::pair<uint32_t, uint32_t>>& edges, bool directed) {
    if (n >= (1ull << 32) - 1) throw std::runtime_error("too many vertices for 32-bit ids");
    const int64_t ne = (int64_t)edges.size();
    std::vector<uint32_t> cnt(n, 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ne; ++i) {
        __atomic_fetch_add(&cnt[edges[i].first], 1u, __ATOMIC_RELAXED);
        if (!directed) __atomic_fetch_add(&cnt[edges[i].second], 1u, __ATOMIC_RELAXED);
    }
    Csr g;
    g.n = n;
    prefix_into(cnt, g.rp);
    g.ci.resize(g.rp[n]);
    std::vector<uint32_t> cur(g.rp.begin(), g.rp.end() - 1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ne; ++i) {
        const uint32_t u = edges[i].first, v = edges[i].second;
        g.ci[__atomic_fetch_add(&cur[u], 1u, __ATOMIC_RELAXED)] = v;
        if (!directed) g.ci[__atomic_fetch_add(&cur[v], 1u, __ATOMIC_RELAXED)] = u;
    }
    canonicalize(n, g.rp, g.ci);
    g.m = g.rp[n];
    g.flags = (directed ? GGPU_FLAG_DIRECTED : 0u) | GGPU_FLAG_SORTED | GGPU_FLAG_DEDUP;
    return g;
}

std::vector<uint32_t> out_degrees(const Csr& g) {
    std::vector<uint32_t> d(g.n);
    for (uint64_t v = 0; v < g.n; ++v) d[v] = g.deg(v);
    return d;
}

Csr transpose(const Csr& g) {
    // Scatter in increasing source order: every in-list comes out sorted.
    Csr t;
    t.n = g.n;
    t.m = g.m;
    t.flags = g.flags;
    std::vector<uint32_t> cnt(g.n, 0);
    for (uint64_t i = 0; i < g.m; ++i) cnt[g.ci[i]]++;
    prefix_into(cnt, t.rp);
    t.ci.resize(g.m);
    std::vector<uint32_t> cur(t.rp.begin(), t.rp.end() - 1);
    for (uint64_t u = 0; u < g.n; ++u)
        for (uint32_t e = g.rp[u]; e < g.rp[u + 1]; ++e) t.ci[cur[g.ci[e]]++] = (uint32_t)u;
    return t;
}

Csr symmetrize(const Csr& g) {
    if (!g.directed()) return g;
    const Csr t = transpose(g);
    const int64_t n = (int64_t)g.n;
    // Union of the sorted out- and in-list of each vertex, each id once.
    auto merged = [&](int64_t v, uint32_t* out) -> uint32_t {
        const uint32_t *a = g.nbr_begin(v), *ae = g.nbr_end(v), *b = t.nbr_begin(v), *be = t.nbr_end(v);
        uint32_t k = 0;
        while (a != ae || b != be) {
            uint32_t x;
            if (b == be || (a != ae && *a < *b)) x = *a++;
            else if (a == ae || *b < *a) x = *b++;
            else { x = *a; ++a; ++b; }
            if (out) out[k] = x;
            ++k;
        }
        return k;
    };
    std::vector<uint32_t> cnt(n);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < n; ++v) cnt[v] = merged(v, nullptr);
    Csr s;
    s.n = g.n;
    prefix_into(cnt, s.rp);
    s.ci.resize(s.rp[n]);
    s.m = s.rp[n];
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < n; ++v) merged(v, s.ci.data() + s.rp[v]);
    s.flags = GGPU_FLAG_SORTED | GGPU_FLAG_DEDUP;
    return s;
}

Csr permute(const Csr& g, const std::vector<uint32_t>& perm) {
    if (perm.size() != g.n) throw std::runtime_error("permute: perm size != n");
    std::vector<uint32_t> inv(g.n);
    for (uint64_t v = 0; v < g.n; ++v) inv[perm[v]] = (uint32_t)v;
    std::vector<uint32_t> cnt(g.n);
    for (uint64_t nv = 0; nv < g.n; ++nv) cnt[nv] = g.deg(inv[nv]);
    Csr p;
    p.n = g.n;
    p.m = g.m;
    p.flags = g.flags;
    prefix_into(cnt, p.rp);
    p.ci.resize(g.m);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t nv = 0; nv < (int64_t)g.n; ++nv) {
        uint32_t* out = p.ci.data() + p.rp[nv];
        uint32_t k = 0;
        for (const uint32_t* q = g.nbr_begin(inv[nv]); q != g.nbr_end(inv[nv]); ++q) out[k++] = perm[*q];
        std::sort(out, out + k);
    }
    return p;
}

std::vector<uint32_t> random_perm(uint64_t n, uint64_t seed) {
    std::vector<uint32_t> p(n);
    std::iota(p.begin(), p.end(), 0u);
    SplitMix64 rng(seed);
    for (uint64_t i = n; i > 1; --i) std::swap(p[i - 1], p[rng.below(i)]);
    return p;
}

std::vector<uint32_t> components(const Csr& g, uint32_t* largest, uint64_t* largest_size) {
    std::vector<uint32_t> comp(g.n, kInf), queue(g.n);
    uint32_t id = 0, best = 0;
    uint64_t best_size = 0;
    for (uint64_t r = 0; r < g.n; ++r) {
        if (comp[r] != kInf) continue;
        uint64_t head = 0, tail = 0;
        comp[r] = id;
        queue[tail++] = (uint32_t)r;
        while (head < tail) {
            const uint32_t u = queue[head++];
            for (const uint32_t* p = g.nbr_begin(u); p != g.nbr_end(u); ++p)
                if (comp[*p] == kInf) { comp[*p] = id; queue[tail++] = *p; }
        }
        if (tail > best_size) { best_size = tail; best = id; }
        ++id;
    }
    if (largest) *largest = best;
    if (largest_size) *largest_size = best_size;
    return comp;
}

Csr largest_component(const Csr& g, std::vector<uint32_t>* old_ids) {
    uint32_t best = 0;
    std::vector<uint32_t> comp = components(g, &best, nullptr);
    std::vector<uint32_t> new_id(g.n, kInf), old;
    for (uint64_t v = 0; v < g.n; ++v)
        if (comp[v] == best) { new_id[v] = (uint32_t)old.size(); old.push_back((uint32_t)v); }
    Csr s;
    s.n = old.size();
    s.flags = g.flags;
    std::vector<uint32_t> cnt(s.n);
    for (uint64_t i = 0; i < s.n; ++i) cnt[i] = g.deg(old[i]);
    prefix_into(cnt, s.rp);
    s.m = s.rp[s.n];
    s.ci.resize(s.m);
    // A component is closed under adjacency and ids keep their order, so the
    // relabeled lists stay sorted.
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < (int64_t)s.n; ++i) {
        uint32_t* out = s.ci.data() + s.rp[i];
        for (const uint32_t* p = g.nbr_begin(old[i]); p != g.nbr_end(old[i]); ++p) *out++ = new_id[*p];
    }
    if (old_ids) old_ids->swap(old);
    return s;
}

std::vector<uint32_t> pick_sources(const Csr& g, uint32_t count, uint64_t seed, bool any_component) {
    std::vector<uint8_t> ok(g.n, 1);
    if (!any_component) {
        uint32_t best = 0;
        std::vector<uint32_t> comp = components(g, &best, nullptr);
        for (uint64_t v = 0; v < g.n; ++v) ok[v] = comp[v] == best;
    }
    uint64_t candidates = 0;
    for (uint64_t v = 0; v < g.n; ++v) candidates += ok[v] && g.deg(v) > 0;
    const uint64_t want = std::min<uint64_t>(count, candidates);
    std::vector<uint32_t> out;
    std::vector<uint8_t> used(g.n, 0);
    SplitMix64 rng(seed);
    while (out.size() < want) {
        const uint32_t v = (uint32_t)rng.below(g.n);
        if (ok[v] && g.deg(v) > 0 && !used[v]) {
            used[v] = 1;
            out.push_back(v);
        }
    }
    return out;
}

}  // namespace ggpu
