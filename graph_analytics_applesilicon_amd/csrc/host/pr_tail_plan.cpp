// Planner of the HOT_TAIL PageRank layout (see pr_tail_plan.hpp and
// csrc/hip/pr_tail.hpp). No reference counterpart: the reference pulls every
// row from the whole contribution vector (src/shaders/pagerank.metal:306-381).
#include "pr_tail_plan.hpp"

#include <omp.h>

#include <algorithm>
#include <stdexcept>
#include <string>

namespace ggpu {

namespace {

void fail(const std::string& what) { throw std::runtime_error("tail PageRank plan: " + what); }

constexpr uint32_t kNoSplit = 0xFFFFFFFFu;

}  // namespace

TailPlan plan_tail(const Csr& in, const TailParams& p) {
    if (p.tile == 0) fail("tile must be >= 1");
    if (p.band_tile && p.band_end <= p.tail_begin) fail("a band must end after the first tail source");
    if (p.nnz_cap < 64 || p.nnz_cap % 64 || p.pieces_cap == 0) fail("bad item limits");
    TailPlan P;
    P.px.push_back(0);
    P.ovptr.push_back(0);
    const uint64_t n = in.n;
    if (n <= p.tail_begin) return P;  // no tail sources
    const uint64_t T = tail_tiles(p, n);

    // split rows: in-degree >= min_indeg, a sorted in-list (so that its tail is a
    // suffix) and at least one tail arc
    std::vector<uint32_t> sp(n, kNoSplit);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
        const uint32_t d = in.deg(v);
        if (d == 0 || d < p.min_indeg) continue;
        const uint32_t* b = in.nbr_begin(v);
        const uint32_t* e = in.nbr_end(v);
        if (!std::is_sorted(b, e) || e[-1] < p.tail_begin) continue;
        sp[v] = (uint32_t)(std::lower_bound(b, e, p.tail_begin) - in.ci.data());
    }
    for (uint64_t v = 0; v < n; ++v)
        if (sp[v] != kNoSplit) {
            P.srow.push_back((uint32_t)v);
            P.split.push_back(sp[v]);
            P.head_arcs += sp[v] - in.rp[v];
        }
    const uint64_t R = P.srow.size();
    if (R == 0) return P;
    P.R = (uint32_t)R;
    P.T = (uint32_t)T;

    // tile-t segment length of every split row, [T][R]
    std::vector<uint32_t> len((size_t)T * R, 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t j = 0; j < (int64_t)R; ++j) {
        const uint32_t* q = in.ci.data() + P.split[j];
        const uint32_t* e = in.nbr_end(P.srow[j]);
        for (; q != e; ++q) len[(size_t)tail_tile_of(p, *q) * R + j]++;
    }
    // overflow slots: every piece after the first of a segment
    P.ovptr.assign(R + 1, 0);
    for (uint64_t j = 0; j < R; ++j) {
        uint64_t o = 0;
        for (uint64_t t = 0; t < T; ++t) {
            const uint32_t L = len[t * R + j];
            if (L > p.nnz_cap) o += (L + p.nnz_cap - 1) / p.nnz_cap - 1;
        }
        const uint64_t next = P.ovptr[j] + o;
        if (next >= (1ull << 32)) fail("too many overflow pieces");
        P.ovptr[j + 1] = (uint32_t)next;
    }
    if (T * R + P.ovptr[R] >= (1ull << 32)) fail("partial index beyond 32 bits");
    // tile-major offsets of the segments
    std::vector<uint64_t> off((size_t)T * R);
    uint64_t total = 0;
    for (size_t i = 0; i < off.size(); ++i) {
        off[i] = total;
        total += len[i];
    }
    if (total >= (1ull << 32)) fail("2^32 or more tail arcs");
    P.tci.resize(total);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t j = 0; j < (int64_t)R; ++j) {
        const uint32_t* q = in.ci.data() + P.split[j];
        const uint32_t* e = in.nbr_end(P.srow[j]);
        while (q != e) {
            const uint64_t t = tail_tile_of(p, *q);
            const uint32_t* r = q;
            while (r != e && tail_tile_of(p, *r) == t) ++r;
            std::copy(q, r, P.tci.begin() + off[t * R + j]);
            q = r;
        }
    }

    // pieces (<= nnz_cap arcs) and items (consecutive pieces of one tile)
    std::vector<uint32_t> ovcur(P.ovptr.begin(), P.ovptr.end() - 1);
    for (uint64_t t = 0; t < T; ++t) {
        uint32_t first = (uint32_t)P.pdst.size(), arcs = 0;
        auto close = [&] {
            const uint32_t end = (uint32_t)P.pdst.size();
            if (end > first) P.items.push_back({first, end});
            first = end;
            arcs = 0;
        };
        for (uint64_t j = 0; j < R; ++j) {
            const uint32_t L = len[t * R + j];
            for (uint32_t c = 0; c * p.nnz_cap < L; ++c) {
                const uint32_t a = std::min(p.nnz_cap, L - c * p.nnz_cap);
                if (arcs + a > p.nnz_cap || P.pdst.size() - first == p.pieces_cap) close();
                P.pdst.push_back(c == 0 ? (uint32_t)(t * R + j) : (uint32_t)(T * R + ovcur[j]++));
                P.px.push_back(P.px.back() + a);
                arcs += a;
            }
        }
        close();
    }
    return P;
}

void emulate_tail(const Csr& in, const TailParams& p, const TailPlan& P, const float* contrib, double* sums) {
    const uint64_t n = in.n, R = P.R, T = P.T;
    if (P.srow.size() != R || P.split.size() != R || P.ovptr.size() != R + 1 || P.px.size() != P.pdst.size() + 1)
        fail("array sizes disagree");
    if (P.px.front() != 0 || P.px.back() != P.tci.size()) fail("pieces do not cover the tail arcs");
    std::vector<int64_t> row_of(n, -1);
    for (uint64_t j = 0; j < R; ++j) {
        const uint32_t v = P.srow[j];
        if (v >= n || (j && v <= P.srow[j - 1])) fail("split rows not ascending");
        if (in.deg(v) < p.min_indeg || P.split[j] < in.rp[v] || P.split[j] >= in.rp[v + 1]) fail("bad split row");
        for (uint32_t x = in.rp[v]; x < in.rp[v + 1]; ++x)
            if ((x < P.split[j]) != (in.ci[x] < p.tail_begin)) fail("a split is not at the first tail arc");
        row_of[v] = (int64_t)j;
    }
    const uint64_t slots = T * R + P.ovptr[R];
    std::vector<double> part(slots, 0.0);
    std::vector<uint8_t> written(slots, 0);
    std::vector<std::vector<uint32_t>> got(R);
    // items: consecutive, in tile order, within the limits
    uint32_t next_piece = 0;
    int64_t last_tile = -1;
    for (const auto& it : P.items) {
        if (it[0] != next_piece || it[1] <= it[0]) fail("items do not cover the pieces in order");
        if (it[1] - it[0] > p.pieces_cap || P.px[it[1]] - P.px[it[0]] > p.nnz_cap) fail("item over its limits");
        next_piece = it[1];
        int64_t tile = -1;
        for (uint32_t q = it[0]; q < it[1]; ++q) {
            const uint32_t a = P.px[q], b = P.px[q + 1];
            if (b <= a || b - a > p.nnz_cap) fail("bad piece size");
            const int64_t t = P.tci[a] < p.tail_begin ? -1 : (int64_t)tail_tile_of(p, P.tci[a]);
            double s = 0.0;
            for (uint32_t x = a; x < b; ++x) {
                const uint32_t u = P.tci[x];
                if (u < p.tail_begin || u >= n || (int64_t)tail_tile_of(p, u) != t)
                    fail("a piece leaves its tile");
                s += (double)contrib[u];
            }
            if (tile < 0) tile = t;
            if (t != tile) fail("an item spans two tiles");
            const uint32_t d = P.pdst[q];
            if (d >= slots || written[d]) fail("a partial is written twice or out of range");
            written[d] = 1;
            part[d] = s;
            uint64_t j;
            if (d < T * R) {
                if ((int64_t)(d / R) != t) fail("a first piece in another tile's partial");
                j = d % R;
            } else {
                j = (uint64_t)(std::upper_bound(P.ovptr.begin(), P.ovptr.end(), d - T * R) - P.ovptr.begin()) - 1;
            }
            got[j].insert(got[j].end(), P.tci.begin() + a, P.tci.begin() + b);
        }
        if (tile < last_tile) fail("items not in tile order");
        last_tile = tile;
    }
    if (next_piece != P.pdst.size()) fail("pieces outside every item");
    for (uint64_t o = T * R; o < slots; ++o)
        if (!written[o]) fail("an overflow partial is never written");
    for (uint64_t v = 0; v < n; ++v) {
        const int64_t j = row_of[v];
        if (j < 0) {
            double s = 0.0;
            for (const uint32_t* q = in.nbr_begin(v); q != in.nbr_end(v); ++q) s += (double)contrib[*q];
            sums[v] = s;
            continue;
        }
        std::vector<uint32_t>& g = got[j];
        std::sort(g.begin(), g.end());
        if (!std::equal(g.begin(), g.end(), in.ci.begin() + P.split[j], in.ci.begin() + in.rp[v + 1]) ||
            g.size() != in.rp[v + 1] - P.split[j])
            fail("a split row's tail arcs are not covered exactly once");
        double s = 0.0;  // the merge's order: head, tiles in order, overflow slots in order
        for (uint32_t x = in.rp[v]; x < P.split[j]; ++x) s += (double)contrib[in.ci[x]];
        for (uint64_t t = 0; t < T; ++t) s += part[t * R + j];
        for (uint32_t o = P.ovptr[j]; o < P.ovptr[j + 1]; ++o) s += part[T * R + o];
        sums[v] = s;
    }
}

}  // namespace ggpu
