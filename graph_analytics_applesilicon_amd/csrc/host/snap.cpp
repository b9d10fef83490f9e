// SNAP edge-list ingest (src/graph/snap.cpp:85-148) and dataset statistics
// (src/graph/stats.cpp:40-120).
//
// Ingest: a text file of "u v" lines, plain or gzip (zlib reads both), '#' and
// '%' lines are comments, blank lines are skipped and columns after the second
// are ignored. The vertices are the raw IDs that occur, numbered densely in
// increasing raw-ID order (the file's own order is the "original" ordering);
// an undirected file stores both directions of every line; lists are sorted
// and deduplicated and self-loops dropped. Two streaming passes (count per raw
// ID, then scatter), so no edge list is ever held.
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "graph.hpp"

namespace ggpu {

namespace {

constexpr uint32_t kAbsent = 0xFFFFFFFFu;

// Streams the (u, v) pairs of an edge-list file to fn; returns the number of
// edge lines. Lines are parsed from 4 MiB chunks; a line that does not fit
// one chunk is an error.
template <class Fn>
uint64_t scan_edges(const std::string& path, Fn&& fn) {
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    struct Closer {
        gzFile f;
        ~Closer() { gzclose(f); }
    } closer{f};
    gzbuffer(f, 1 << 20);
    const size_t kChunk = 4u << 20;
    std::vector<char> buf(kChunk + 1);
    size_t carry = 0;
    uint64_t edges = 0, line = 0;
    for (bool eof = false; !eof;) {
        const int got = gzread(f, buf.data() + carry, (unsigned)(kChunk - carry));
        if (got < 0) throw std::runtime_error("read error in " + path);
        eof = got == 0;
        size_t len = carry + (size_t)got;
        if (eof && len && buf[len - 1] != '\n') buf[len++] = '\n';  // a last line without newline
        size_t end = len;
        while (end && buf[end - 1] != '\n') --end;
        if (!end && !eof && len == kChunk) throw std::runtime_error("line longer than 4 MiB in " + path);
        const char* p = buf.data();
        const char* const stop = buf.data() + end;
        while (p < stop) {
            ++line;
            while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
            if (*p == '#' || *p == '%' || *p == '\n') {  // comment or blank
                while (*p != '\n') ++p;
                ++p;
                continue;
            }
            uint32_t id[2];
            for (uint32_t& x : id) {
                while (*p == ' ' || *p == '\t') ++p;
                if (*p < '0' || *p > '9') throw std::runtime_error(path + ": malformed line " + std::to_string(line));
                uint64_t v = 0;
                while (*p >= '0' && *p <= '9') {
                    v = v * 10 + (uint64_t)(*p++ - '0');
                    if (v >= kAbsent) throw std::runtime_error(path + ": vertex ID too large on line " + std::to_string(line));
                }
                x = (uint32_t)v;
            }
            while (*p != '\n') ++p;
            ++p;
            fn(id[0], id[1]);
            ++edges;
        }
        carry = len - end;
        std::memmove(buf.data(), buf.data() + end, carry);
    }
    return edges;
}

}  // namespace

Csr load_snap(const std::string& path, bool directed, SnapStats* st) {
    SnapStats s;
    // pass 1: which raw IDs occur and how many list entries each owns
    std::vector<uint32_t> count;  // by raw ID; kAbsent = never seen
    auto see = [&](uint32_t id) {
        if (id >= count.size()) count.resize(std::max<size_t>((size_t)id + 1, count.size() * 3 / 2), kAbsent);
        if (count[id] == kAbsent) count[id] = 0;
    };
    uint64_t entries = 0;
    s.lines = scan_edges(path, [&](uint32_t u, uint32_t v) {
        see(u);
        see(v);
        ++count[u];
        if (!directed) ++count[v];
        entries += directed ? 1 : 2;
    });
    if (entries >= kAbsent) throw std::runtime_error("too many edges for 32-bit offsets");
    s.m_before = entries;
    // dense ids in raw-ID order; count[] becomes the raw -> dense map
    std::vector<uint32_t> rp(1, 0);
    for (size_t id = 0; id < count.size(); ++id) {
        if (count[id] == kAbsent) continue;
        rp.push_back(rp.back() + count[id]);
        count[id] = (uint32_t)(rp.size() - 2);
        s.raw_max_id = id;
    }
    const uint64_t n = rp.size() - 1;
    s.n = n;
    // pass 2: scatter through per-vertex cursors
    std::vector<uint32_t> ci(entries), at(rp.begin(), rp.end() - 1);
    scan_edges(path, [&](uint32_t u, uint32_t v) {
        const uint32_t a = count[u], b = count[v];
        ci[at[a]++] = b;
        if (!directed) ci[at[b]++] = a;
    });
    count = {};
    at = {};
    canonicalize_lists(n, rp, ci);
    Csr g;
    g.n = n;
    g.m = rp[n];
    g.flags = (directed ? GGPU_FLAG_DIRECTED : 0u) | GGPU_FLAG_SORTED | GGPU_FLAG_DEDUP;
    g.rp = std::move(rp);
    g.ci = std::move(ci);
    g.ci.resize(g.m);
    s.m = g.m;
    if (st) *st = s;
    return g;
}

// Degrees, components and an approximate diameter as gstats reports them.
GraphStats graph_stats(const Csr& g) {
    GraphStats t;
    const Csr s = symmetrize(g);
    t.n = g.n;
    t.m = g.m;
    t.undirected_edges = s.m / 2;
    if (!g.n) return t;
    std::vector<uint32_t> in;
    if (g.directed()) {
        in.assign(g.n, 0);
        for (uint64_t e = 0; e < g.m; ++e) ++in[g.ci[e]];
    }
    t.min_deg = UINT32_MAX;
    for (uint64_t v = 0; v < g.n; ++v) {
        t.min_deg = std::min(t.min_deg, g.deg(v));
        t.max_deg = std::max(t.max_deg, g.deg(v));
        t.isolated += s.deg(v) == 0;
    }
    t.max_in_deg = in.empty() ? t.max_deg : *std::max_element(in.begin(), in.end());
    t.mean_deg = (double)g.m / (double)g.n;
    uint32_t largest = 0;
    uint64_t largest_size = 0;
    const std::vector<uint32_t> comp = components(s, &largest, &largest_size);
    t.num_cc = comp.empty() ? 0 : (uint64_t)*std::max_element(comp.begin(), comp.end()) + 1;
    t.largest_cc = largest_size;
    // the deepest level reached from 10 random non-isolated sources (SplitMix64, seed 1)
    if (t.isolated < g.n) {
        SplitMix64 rng(1);
        std::vector<uint32_t> dist(g.n), q(g.n);
        for (int k = 0; k < 10; ++k) {
            uint32_t src;
            do src = (uint32_t)rng.below(g.n);
            while (s.deg(src) == 0);
            std::fill(dist.begin(), dist.end(), kAbsent);
            size_t head = 0, tail = 0;
            dist[src] = 0;
            q[tail++] = src;
            while (head < tail) {
                const uint32_t u = q[head++];
                t.approx_diameter = std::max(t.approx_diameter, dist[u]);
                for (const uint32_t* p = s.nbr_begin(u); p != s.nbr_end(u); ++p)
                    if (dist[*p] == kAbsent) {
                        dist[*p] = dist[u] + 1;
                        q[tail++] = *p;
                    }
            }
        }
    }
    return t;
}

// GAP Benchmark Suite serialized graph (src/graph/export_sg.cpp:1-12, the
// layout of gapbs' WriteSerializedGraph): bool directed, int64 adjacency
// entries, int64 vertices, int64 offsets[n+1], int32 neighbours[m], and for a
// directed graph the in-CSR (its transpose) the same way. sym: write the
// symmetrized graph as undirected (what BFS traverses).
void save_sg(const Csr& g_in, const std::string& path, bool sym) {
    if (g_in.n >= (1ull << 31)) throw std::invalid_argument("vertex ids do not fit GAP's int32 ids");
    const Csr* g = &g_in;
    Csr s;
    if (sym && g_in.directed()) {
        s = symmetrize(g_in);
        g = &s;
    }
    const bool directed = !sym && g->directed();
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    auto put_csr = [&](const Csr& c) {
        std::vector<int64_t> wide;
        for (uint64_t i = 0; i <= c.n; i += 1u << 20) {
            const uint64_t e = std::min<uint64_t>(c.n + 1, i + (1u << 20));
            wide.assign(c.rp.begin() + i, c.rp.begin() + e);
            if (std::fwrite(wide.data(), 8, wide.size(), f) != wide.size()) return false;
        }
        return c.m == 0 || std::fwrite(c.ci.data(), 4, c.m, f) == c.m;
    };
    const int64_t m = (int64_t)g->m, n = (int64_t)g->n;
    const uint8_t d = directed ? 1 : 0;
    bool ok = std::fwrite(&d, 1, 1, f) == 1 && std::fwrite(&m, 8, 1, f) == 1 && std::fwrite(&n, 8, 1, f) == 1 &&
              put_csr(*g);
    if (ok && directed) ok = put_csr(transpose(*g));
    ok = std::fclose(f) == 0 && ok;
    if (!ok) throw std::runtime_error("short write to " + path);
}

// Bucket 0: degree 0; bucket k >= 1: degrees in [2^(k-1), 2^k). Trailing empty
// buckets dropped (one bucket at least).
std::vector<uint64_t> log2_histogram(const std::vector<uint32_t>& deg) {
    std::vector<uint64_t> h(34, 0);
    for (uint32_t d : deg) ++h[d ? 32 - __builtin_clz(d) : 0];
    while (h.size() > 1 && !h.back()) h.pop_back();
    return h;
}

}  // namespace ggpu
