// Planner of the TILED PageRank layout (see pr_tiling.hpp and
// csrc/hip/pr_tiled.hpp). No reference counterpart: the reference pulls every
// row from the whole contribution vector (src/shaders/pagerank.metal:306-381).
#include "pr_tiling.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <string>

namespace ggpu {

namespace {

void fail(const std::string& what) { throw std::runtime_error("tiled PageRank plan: " + what); }

struct Task {
    std::array<uint32_t, 8> w;
    double cost;
};

// Lanes per segment of a run of r segments (the largest power of two <= 64 / r,
// as the kernel assigns them) times the arcs one lane may take.
uint32_t lane_arcs(const TiledParams& p, uint32_t r) {
    uint32_t g = 64;
    while (g > 1 && g * r > 64) g >>= 1;
    return g * (p.nnz_cap / 64);
}

}  // namespace

TiledPlan plan_tiled(const Csr& in, const TiledParams& p) {
    if (p.tile == 0 || p.tile > 65536) fail("tile must be in 1..65536");
    if (p.nnz_cap < 64 || p.nnz_cap > 65535 || p.nnz_cap % 64 || p.rows_cap == 0 || p.rows_cap > 64 || p.slots == 0 ||
        p.groups == 0)
        fail("bad item, run, slot or group limits");
    if (p.min_indeg > p.nnz_cap) fail("min_indeg above the item size: an untiled row would not fit one item");
    TiledPlan P;
    const uint32_t n = (uint32_t)in.n;
    const uint32_t nt = n ? (n - 1) / p.tile + 1 : 0;
    std::vector<uint8_t> tiled(n, 0);
    for (uint32_t v = 0; v < n; ++v)
        if (in.deg(v) >= p.min_indeg && in.deg(v) > 0) {
            tiled[v] = 1;
            P.trow.push_back(v);
        }
    const uint32_t R = P.R = (uint32_t)P.trow.size();

    // arcs of each source tile into the tiled rows
    std::vector<uint64_t> tile_arcs(nt, 0);
#pragma omp parallel
    {
        std::vector<uint64_t> mine(nt, 0);
#pragma omp for schedule(dynamic, 256) nowait
        for (int64_t j = 0; j < (int64_t)R; ++j) {
            const uint32_t v = P.trow[j];
            for (const uint32_t* q = in.nbr_begin(v); q != in.nbr_end(v); ++q) mine[*q / p.tile]++;
        }
#pragma omp critical
        for (uint32_t t = 0; t < nt; ++t) tile_arcs[t] += mine[t];
    }
    // a tile is worth its LDS copy and its partial array (8 bytes per tiled row,
    // written and read) once it saves a few gathers per tiled row
    const uint64_t dense_min = p.dense_min_arcs ? p.dense_min_arcs : std::max<uint64_t>(p.tile / 16, R / 8);
    std::vector<uint32_t> dense_of(nt, kTiNoTile);
    for (uint32_t t = 0; t < nt; ++t)
        if (tile_arcs[t] >= dense_min && tile_arcs[t] > 0) {
            dense_of[t] = (uint32_t)P.dense_tile.size();
            P.dense_tile.push_back(t);
        }
    const uint32_t K = P.K = (uint32_t)P.dense_tile.size();

    // segment lengths [K + 1][R] (index K: the cold pseudo-tile)
    std::vector<uint32_t> len((size_t)(K + 1) * R, 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t j = 0; j < (int64_t)R; ++j) {
        const uint32_t v = P.trow[j];
        for (const uint32_t* q = in.nbr_begin(v); q != in.nbr_end(v); ++q) {
            const uint32_t k = dense_of[*q / p.tile];
            len[(size_t)(k == kTiNoTile ? K : k) * R + j]++;
        }
    }
    // arc offsets in tile-major order (dense arcs in tci, cold arcs in cci)
    std::vector<uint64_t> off((size_t)(K + 1) * R, 0);
    uint64_t acc = 0;
    for (size_t i = 0; i < (size_t)K * R; ++i) {
        off[i] = acc;
        acc += len[i];
    }
    P.dense_arcs = acc;
    acc = 0;
    for (size_t i = (size_t)K * R; i < (size_t)(K + 1) * R; ++i) {
        off[i] = acc;
        acc += len[i];
    }
    P.cold_arcs = acc;
    if (P.dense_arcs >= (1ull << 32) || P.cold_arcs >= (1ull << 32)) fail("too many tiled arcs");
    P.tci.resize(P.dense_arcs);
    P.cci.resize(P.cold_arcs);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t j = 0; j < (int64_t)R; ++j) {
        const uint32_t v = P.trow[j];
        std::vector<uint64_t> at(K + 1);
        for (uint32_t k = 0; k <= K; ++k) at[k] = off[(size_t)k * R + j];
        for (const uint32_t* q = in.nbr_begin(v); q != in.nbr_end(v); ++q) {
            const uint32_t t = *q / p.tile, k = dense_of[t];
            if (k == kTiNoTile) P.cci[at[K]++] = *q;
            else P.tci[at[k]++] = (uint16_t)(*q - t * p.tile);
        }
    }

    // items and tasks of every partial array, in (k, row) order
    std::vector<Task> tasks;
    for (uint32_t k = 0; k <= K; ++k) {
        const bool cold = k == K;
        Task cur{};
        uint64_t cur_arcs = 0;
        uint32_t cur_slots = 0;
        auto open = [&]() {
            cur = Task{};
            cur.w[0] = cur.w[1] = (uint32_t)P.items.size();
            cur.w[2] = cold ? kTiNoTile : P.dense_tile[k];
            cur.w[3] = k;
            cur.w[4] = cur.w[5] = (uint32_t)P.segs.size();
            cur_arcs = 0;
            cur_slots = 0;
        };
        auto close = [&]() {
            cur.w[1] = (uint32_t)P.items.size();
            cur.w[5] = (uint32_t)P.segs.size();
            if (cur.w[1] > cur.w[0]) {
                // LDS gathers cost ~1, global gathers (cold) several times that, plus the tile copy
                cur.cost = cold ? 8.0 * (double)cur_arcs : (double)cur_arcs + p.tile / 16.0;
                cur.cost += 32.0 * (cur.w[1] - cur.w[0]);
                tasks.push_back(cur);
            }
            open();
        };
        open();
        uint32_t run_e0 = 0, run_n = 0, run_arcs = 0, run_max = 0;
        uint64_t run_a0 = 0;
        auto flush_run = [&]() {
            if (run_n) P.items.push_back({kTiRun | (k << 2), run_e0, run_e0 + run_n, (uint32_t)run_a0});
            cur_arcs += run_arcs;
            run_n = 0;
            run_arcs = 0;
            run_max = 0;
        };
        for (uint32_t j = 0; j < R; ++j) {
            const uint32_t L = len[(size_t)k * R + j];
            if (L == 0) continue;
            const uint64_t a = off[(size_t)k * R + j];
            if (L <= p.nnz_cap) {
                // a run of r segments gives each segment lanes_for(r) lanes; no lane takes more
                // than nnz_cap / 64 arcs, so its index loads are issued at once
                if (run_n && (run_n == p.rows_cap || run_arcs + L > p.nnz_cap ||
                              std::max(run_max, L) > lane_arcs(p, run_n + 1)))
                    flush_run();
                if (!run_n) {
                    run_e0 = (uint32_t)P.ent_row.size();
                    run_a0 = a;
                }
                P.ent_row.push_back(j);
                P.ent_len.push_back((uint16_t)L);
                ++run_n;
                run_arcs += L;
                run_max = std::max(run_max, L);
            } else {
                flush_run();
                // pieces of <= slots chunks; piece 0 writes P[k][j], the others an overflow partial
                const uint64_t piece_arcs = (uint64_t)p.slots * p.nnz_cap;
                for (uint64_t pa = 0; pa < L; pa += piece_arcs) {
                    const uint32_t pl = (uint32_t)std::min<uint64_t>(piece_arcs, L - pa);
                    const uint32_t chunks = (pl + p.nnz_cap - 1) / p.nnz_cap;
                    if (cur_slots + chunks > p.slots) close();
                    uint32_t dest = 0;
                    if (pa) {
                        dest = (uint32_t)P.ovf_row.size() + 1;
                        P.ovf_row.push_back(j);
                    }
                    P.segs.push_back({j, cur_slots, chunks, dest});
                    for (uint32_t c = 0; c < chunks; ++c) {
                        const uint64_t b = a + pa + (uint64_t)c * p.nnz_cap, e = std::min<uint64_t>(a + pa + pl, b + p.nnz_cap);
                        P.items.push_back({kTiChunk | (k << 2), cur_slots + c, (uint32_t)b, (uint32_t)e});
                    }
                    cur_slots += chunks;
                    cur_arcs += pl;
                }
            }
            if (cur_arcs + run_arcs >= p.task_arcs || P.items.size() - cur.w[0] >= p.task_items) {
                flush_run();
                close();
            }
        }
        flush_run();
        close();
    }
    // overflow partials of each tiled row, in (k, piece) order
    P.optr.assign(R + 1, 0);
    for (uint32_t r : P.ovf_row) P.optr[r + 1]++;
    for (uint32_t j = 0; j < R; ++j) P.optr[j + 1] += P.optr[j];
    {
        std::vector<uint32_t> at(P.optr.begin(), P.optr.end() - 1);
        P.oidx.resize(P.ovf_row.size());
        for (uint32_t o = 0; o < (uint32_t)P.ovf_row.size(); ++o) P.oidx[at[P.ovf_row[o]]++] = o;
    }
    // untiled rows: runs of consecutive untiled rows, tile 0 in LDS
    {
        Task cur{};
        double cost = 0;
        uint64_t arcs = 0;
        auto open = [&]() {
            cur = Task{};
            cur.w[0] = cur.w[1] = (uint32_t)P.items.size();
            cur.w[2] = n ? 0u : kTiNoTile;
            cur.w[4] = cur.w[5] = (uint32_t)P.segs.size();
            cost = 0;
            arcs = 0;
        };
        auto close = [&]() {
            cur.w[1] = (uint32_t)P.items.size();
            cur.w[5] = cur.w[4];
            if (cur.w[1] > cur.w[0]) {
                cur.cost = cost + 32.0 * (cur.w[1] - cur.w[0]);
                tasks.push_back(cur);
            }
            open();
        };
        open();
        uint32_t v = 0;
        while (v < n) {
            if (tiled[v]) {
                ++v;
                continue;
            }
            const uint32_t r0 = v;
            uint32_t run_arcs = 0, run_max = 0;
            double run_cost = 0;
            while (v < n && !tiled[v] && v - r0 < p.rows_cap &&
                   (v == r0 || (run_arcs + in.deg(v) <= p.nnz_cap &&
                                std::max(run_max, in.deg(v)) <= lane_arcs(p, v - r0 + 1)))) {
                run_arcs += in.deg(v);
                run_max = std::max(run_max, in.deg(v));
                for (const uint32_t* q = in.nbr_begin(v); q != in.nbr_end(v); ++q) run_cost += *q < p.tile ? 1.0 : 8.0;
                ++v;
            }
            P.items.push_back({kTiUntiled, r0, v, 0});
            arcs += run_arcs;
            cost += run_cost;
            P.untiled_arcs += run_arcs;
            if (arcs >= p.task_arcs || P.items.size() - cur.w[0] >= p.task_items) close();
        }
        close();
    }
    if (P.items.size() >= (1ull << 32)) fail("too many items");

    // longest-processing-time assignment to the workgroups, then each
    // workgroup's tasks grouped by the tile they copy to LDS
    std::vector<uint32_t> order(tasks.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return tasks[a].cost > tasks[b].cost; });
    using Load = std::pair<double, uint32_t>;
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (uint32_t g = 0; g < p.groups; ++g) heap.push({0.0, g});
    std::vector<std::vector<uint32_t>> per(p.groups);
    for (uint32_t t : order) {
        Load l = heap.top();
        heap.pop();
        per[l.second].push_back(t);
        l.first += tasks[t].cost;
        heap.push(l);
    }
    P.wg.resize(p.groups);
    for (uint32_t g = 0; g < p.groups; ++g) {
        auto& L = per[g];
        std::stable_sort(L.begin(), L.end(), [&](uint32_t a, uint32_t b) {
            return tasks[a].w[2] != tasks[b].w[2] ? tasks[a].w[2] < tasks[b].w[2] : a < b;
        });
        P.wg[g] = {(uint32_t)P.tasks.size(), (uint32_t)(P.tasks.size() + L.size())};
        for (uint32_t t : L) P.tasks.push_back(tasks[t].w);
    }
    return P;
}

void emulate_tiled(const Csr& in, const TiledParams& p, const TiledPlan& P, const float* contrib, double* sums) {
    const uint32_t n = (uint32_t)in.n, R = P.R, K = P.K;
    std::vector<double> part((size_t)(K + 1) * R, 0.0), ovf(P.ovf_row.size(), 0.0);
    std::vector<uint8_t> part_set((size_t)(K + 1) * R, 0), ovf_set(P.ovf_row.size(), 0), stored(n, 0);
    std::vector<double> slots(p.slots);
    std::vector<uint8_t> task_seen(P.tasks.size(), 0), item_seen(P.items.size(), 0);
    for (uint32_t v = 0; v < n; ++v) sums[v] = 0;
    auto set_part = [&](uint32_t k, uint32_t j, double s) {
        if (k > K || j >= R) fail("partial index out of range");
        if (part_set[(size_t)k * R + j]++) fail("partial written twice");
        part[(size_t)k * R + j] = s;
    };
    if (P.wg.size() != p.groups) fail("workgroup count");
    for (const auto& g : P.wg) {
        if (g[0] > g[1] || g[1] > P.tasks.size()) fail("workgroup task range");
        for (uint32_t t = g[0]; t < g[1]; ++t) {
            if (task_seen[t]++) fail("task assigned twice");
            const auto& T = P.tasks[t];
            const uint32_t lds = T[2], k = T[3];
            if (T[0] > T[1] || T[1] > P.items.size() || T[4] > T[5] || T[5] > P.segs.size()) fail("task ranges");
            std::fill(slots.begin(), slots.end(), std::nan(""));
            for (uint32_t it = T[0]; it < T[1]; ++it) {
                if (item_seen[it]++) fail("item in two tasks");
                const auto& I = P.items[it];
                const uint32_t kind = I[0] & 3u, ik = I[0] >> 2;
                if (kind != kTiUntiled && ik != k) fail("item of another partial array");
                auto src_of = [&](uint32_t kk, uint64_t a) -> uint32_t {
                    if (kk == K) {
                        if (a >= P.cci.size()) fail("cold arc out of range");
                        return P.cci[a];
                    }
                    if (a >= P.tci.size()) fail("dense arc out of range");
                    if (lds != P.dense_tile[kk]) fail("dense item without its tile in LDS");
                    return lds * p.tile + P.tci[a];
                };
                if (kind == kTiRun) {
                    if (I[2] <= I[1] || I[2] - I[1] > p.rows_cap || I[2] > P.ent_row.size()) fail("run entries");
                    uint64_t a = I[3], arcs = 0;
                    for (uint32_t e = I[1]; e < I[2]; ++e) {
                        double s = 0;
                        for (uint32_t x = 0; x < P.ent_len[e]; ++x) s += contrib[src_of(ik, a + x)];
                        a += P.ent_len[e];
                        arcs += P.ent_len[e];
                        set_part(ik, P.ent_row[e], s);
                    }
                    if (arcs > p.nnz_cap) fail("run over the arc cap");
                    for (uint32_t e = I[1]; e < I[2]; ++e)
                        if (P.ent_len[e] > lane_arcs(p, I[2] - I[1])) fail("a lane of a run takes too many arcs");
                } else if (kind == kTiChunk) {
                    if (I[1] >= p.slots || I[3] <= I[2] || I[3] - I[2] > p.nnz_cap) fail("chunk");
                    double s = 0;
                    for (uint64_t a = I[2]; a < I[3]; ++a) s += contrib[src_of(ik, a)];
                    slots[I[1]] = s;
                } else if (kind == kTiUntiled) {
                    if (lds != 0 || I[2] <= I[1] || I[2] > n || I[2] - I[1] > p.rows_cap) fail("untiled run");
                    for (uint32_t v = I[1]; v < I[2]; ++v)
                        if (in.deg(v) > lane_arcs(p, I[2] - I[1])) fail("a lane of an untiled run takes too many arcs");
                    for (uint32_t v = I[1]; v < I[2]; ++v) {
                        if (stored[v]++) fail("row stored twice");
                        double s = 0;
                        for (const uint32_t* q = in.nbr_begin(v); q != in.nbr_end(v); ++q) s += contrib[*q];
                        sums[v] = s;
                    }
                } else {
                    fail("unknown item kind");
                }
            }
            for (uint32_t s = T[4]; s < T[5]; ++s) {
                const auto& S = P.segs[s];
                if (S[1] + S[2] > p.slots) fail("segment slots");
                double x = 0;
                for (uint32_t c = 0; c < S[2]; ++c) {
                    if (std::isnan(slots[S[1] + c])) fail("segment slot not written by its task");
                    x += slots[S[1] + c];
                }
                if (S[3] == 0) {
                    set_part(k, S[0], x);
                } else {
                    if (S[3] > ovf.size() || P.ovf_row[S[3] - 1] != S[0] || ovf_set[S[3] - 1]++) fail("overflow partial");
                    ovf[S[3] - 1] = x;
                }
            }
        }
    }
    for (size_t t = 0; t < P.tasks.size(); ++t)
        if (!task_seen[t]) fail("task not assigned");
    for (size_t i = 0; i < P.items.size(); ++i)
        if (!item_seen[i]) fail("item not in a task");
    for (uint32_t j = 0; j < R; ++j) {
        const uint32_t v = P.trow[j];
        if (stored[v]++) fail("tiled row stored twice");
        double s = 0;
        for (uint32_t k = 0; k <= K; ++k) s += part[(size_t)k * R + j];
        for (uint32_t x = P.optr[j]; x < P.optr[j + 1]; ++x) {
            if (!ovf_set[P.oidx[x]]) fail("overflow partial never written");
            s += ovf[P.oidx[x]];
        }
        sums[v] = s;
    }
    for (uint32_t v = 0; v < n; ++v)
        if (!stored[v]) fail("row " + std::to_string(v) + " never stored");
}

}  // namespace ggpu
