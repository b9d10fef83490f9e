// BFS backend: device buffers, the host-driven level loop and the C-ABI.
//
// Host-driven loop = the reference protocol (src/metal/bfs_gpu.mm:326-429,
// runDirOpt; :384-412 runEdge; :177-236 runBinned): the host decides each
// level, launches that level's kernels, reads the level's stats back and
// synchronizes once per level. The decision rule is the same function the
// device-driven loops (bfs_loop.hip) evaluate on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "bfs_engine.hpp"
#include "graph.hpp"

namespace ggpu {

// ------------------------------------------------------------ kernels
__global__ void k_init(BfsDev d, uint32_t src, uint32_t src_bin) {
    ph_init(d, src, blockIdx.x, gridDim.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t dg = d.rp[src + 1] - d.rp[src];
        d.q[0][0] = src;
        d.deg[0][0] = dg;
        if (d.bins[0][src_bin]) d.bins[0][src_bin][0] = src;
    }
}
// gpu_vertex level (bfs.metal:45-79): a lane per vertex; vertices at depth
// `level` store level + 1 into their unreached neighbours. Concurrent stores
// to one vertex write the same value, so they are plain stores; a vertex
// stored twice is counted twice in S_FOUND, which only decides whether
// another level runs.
__global__ void k_vertex_level(BfsDev d, uint32_t level, unsigned long long* slot) {
    unsigned long long fr = 0, ed = 0, fo = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < d.n; v += gridDim.x * blockDim.x) {
        if (d.dist[v] != level) continue;
        const uint32_t e0 = d.rp[v], e1 = d.rp[v + 1];
        ++fr;
        ed += e1 - e0;
        for (uint32_t e = e0; e < e1; ++e) {
            const uint32_t u = d.ci[e];
            if (d.dist[u] == GGPU_INF) {
                d.dist[u] = level + 1;
                ++fo;
            }
        }
    }
    const int words[3] = {S_FRONTIER, S_EXAMINED, S_FOUND};
    const unsigned long long vals[3] = {fr, ed, fo};
    block_flush<3>(slot, words, vals);
}

// gpu_naive level (bfs.metal:95-126): a lane per queued vertex walks its whole
// list; a neighbour is claimed by compare-and-swap INF -> level + 1 (after a
// plain load filters visited ones) and every claim appends itself to the next
// queue with its own atomic on the tail, so the queue is duplicate-free and
// the tail is the exact count of discovered vertices.
__global__ void k_naive_level(BfsDev d, const uint32_t* q, uint32_t* nq, uint32_t size, uint32_t level,
                              unsigned long long* slot) {
    unsigned long long fr = 0, ed = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < size; i += gridDim.x * blockDim.x) {
        const uint32_t v = q[i];
        const uint32_t e0 = d.rp[v], e1 = d.rp[v + 1];
        ++fr;
        ed += e1 - e0;
        for (uint32_t e = e0; e < e1; ++e) {
            const uint32_t u = d.ci[e];
            if (d.dist[u] != GGPU_INF) continue;
            if (atomicCAS(&d.dist[u], GGPU_INF, level + 1) == GGPU_INF) {
                const unsigned long long pos = atomicAdd(&d.conv[kTailQueue], 1ull);
                nq[pos] = u;
            }
        }
    }
    const int words[2] = {S_FRONTIER, S_EXAMINED};
    const unsigned long long vals[2] = {fr, ed};
    block_flush<2>(slot, words, vals);
}

__global__ void k_scan_tiles(uint32_t* vals, uint32_t size, uint32_t* tile) {
    ph_scan_tiles(vals, size, tile, blockIdx.x, gridDim.x);
}
__global__ void k_scan_apply(uint32_t* vals, uint32_t size, const uint32_t* tile) {
    ph_scan_apply(vals, size, tile, blockIdx.x, gridDim.x);
}
__global__ void k_td_edge(BfsDev d, uint32_t cur, uint32_t level, uint32_t size, uint32_t total, uint32_t ept,
                          unsigned long long* stats) {
    ph_td_edge(d, cur, level, size, total, ept, stats, blockIdx.x, gridDim.x);
}
__global__ void k_td_binned(BfsDev d, uint32_t cur, uint32_t level, uint32_t small_max, uint32_t large_min,
                            uint32_t c0, uint32_t c1, uint32_t c2, unsigned long long* stats) {
    ph_td_binned(d, cur, level, small_max, large_min, c0, c1, c2, stats, blockIdx.x, gridDim.x);
}
__global__ void k_bu(BfsDev d, uint32_t cur, uint32_t level, unsigned long long* stats) {
    ph_bu(d, d.bm[cur], d.bm[cur ^ 1], level, stats, blockIdx.x, gridDim.x);
}
__global__ void k_bu_heavy(BfsDev d, uint32_t cur, uint32_t level, unsigned long long* stats) {
    ph_bu_heavy(d, d.bm[cur], d.bm[cur ^ 1], level, stats, blockIdx.x, gridDim.x);
}
__global__ void k_dist_to_bitmap(BfsDev d, uint32_t cur, uint32_t level) {
    ph_dist_to_bitmap(d, d.bm[cur], level, blockIdx.x, gridDim.x);
}
__global__ void k_bitmap_to_queue(BfsDev d, uint32_t bm_cur, uint32_t q_cur) {
    ph_bitmap_to_queue(d, d.bm[bm_cur], d.q[q_cur], d.deg[q_cur], &d.conv[0], blockIdx.x, gridDim.x);
}
__global__ void k_dist_collect(BfsDev d, uint32_t level, uint32_t to_bins, uint32_t dst, uint32_t small_max,
                               uint32_t large_min) {
    ph_dist_collect(d, level, to_bins != 0, dst, small_max, large_min, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------ helpers

uint32_t blocks_for(uint64_t items, uint32_t per_block, uint32_t cap) {
    const uint64_t b = (items + per_block - 1) / per_block;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, cap));
}

void check_tg(uint32_t tg) {
    if (tg < 64 || tg > 1024 || tg % 64 != 0)
        throw ArgError("workgroup size must be a multiple of 64 in [64, 1024]");
}

void BfsEngine::ensure_edge_buffers() {
    qa.alloc(n);
    qb.alloc(n);
    dega.alloc(n);
    degb.alloc(n);
    tile.alloc((n + 64 * kScanItems - 1) / (64 * kScanItems) + 1);
}

void BfsEngine::ensure_bins() {
    if (bins.count >= 6ull * n) return;
    bins.alloc(6ull * (n ? n : 1));
}

void BfsEngine::ensure_bitmaps() { bms.alloc(3ull * nw + 1); }

BfsDev BfsEngine::dev() {
    BfsDev d{};
    d.rp = rp.p;
    d.ci = ci.p;
    d.dist = dist.p;
    d.q[0] = qa.p;
    d.q[1] = qb.p;
    d.deg[0] = dega.p;
    d.deg[1] = degb.p;
    for (int s = 0; s < 2; ++s)
        for (int k = 0; k < 3; ++k) d.bins[s][k] = bins.p ? bins.p + (size_t)(s * 3 + k) * n : nullptr;
    d.bm[0] = bms.p;
    d.bm[1] = bms.p ? bms.p + nw : nullptr;
    d.visited = bms.p ? bms.p + 2ull * nw : nullptr;
    d.tile = tile.p;
    d.stats = stats.p;
    d.conv = stats.p + kSlotWords;
    d.heavy = heavy.p;
    d.n_heavy = n_heavy;
    d.heavy_min = kHeavyMin;
    d.n = n;
    d.nw = nw;
    d.m = m;
    return d;
}

BfsEngine::BfsEngine(ggpu_ctx* c, uint64_t n_, uint64_t m_, const uint32_t* hrp, const uint32_t* hci)
    : ctx(c), n((uint32_t)n_), m(m_) {
    if (n_ >= GGPU_INF) throw ArgError("graph too large for 32-bit vertex ids");
    if (m_ >= (1ull << 32)) throw ArgError("graph has 2^32 or more arcs");
    nw = (n + 63) / 64;
    host = from_arrays(n_, m_, hrp, hci, GGPU_FLAG_SORTED | GGPU_FLAG_DEDUP);
    GGPU_HIP_TRY(hipSetDevice(ctx->device));
    rp.alloc(n + 1);
    ci.alloc(m ? m : 1);
    dist.alloc(n ? n : 1);
    stats.alloc(kStatSlots * kSlotWords + kConvWords);
    GGPU_HIP_TRY(hipMemcpyAsync(rp.p, hrp, 4ull * (n + 1), hipMemcpyHostToDevice, ctx->stream));
    if (m) GGPU_HIP_TRY(hipMemcpyAsync(ci.p, hci, 4ull * m, hipMemcpyHostToDevice, ctx->stream));
    else GGPU_HIP_TRY(hipMemsetAsync(ci.p, 0, 4, ctx->stream));  // the placeholder arc: a valid vertex id
    ensure_edge_buffers();
    ensure_bitmaps();
    // Bottom-up heavy items: vertices of degree >= kHeavyMin, lists longer than
    // kHeavyChunk split into chunks; longest first so the tail stays short.
    {
        std::vector<uint4> items;
        for (uint32_t v = 0; v < n; ++v) {
            const uint32_t dg = host.deg(v);
            if (dg < kHeavyMin) continue;
            for (uint32_t x = host.rp[v]; x < host.rp[v + 1]; x += kHeavyChunk)
                items.push_back(make_uint4(v, x, std::min(x + kHeavyChunk, host.rp[v + 1]), dg));
        }
        std::stable_sort(items.begin(), items.end(), [](const uint4& a, const uint4& b) { return a.w > b.w; });
        n_heavy = (uint32_t)items.size();
        heavy.alloc(items.size() + 1);
        if (!items.empty())
            GGPU_HIP_TRY(hipMemcpyAsync(heavy.p, items.data(), items.size() * sizeof(uint4), hipMemcpyHostToDevice,
                                        ctx->stream));
    }
    GGPU_HIP_TRY(hipStreamSynchronize(ctx->stream));
    int dev_blocks = 0;
    GGPU_HIP_TRY(hipDeviceGetAttribute(&dev_blocks, hipDeviceAttributeMultiprocessorCount, ctx->device));
    grid_cap = (uint32_t)std::max(1, dev_blocks) * 16;
}

BfsEngine::~BfsEngine() {
    if (ms_state && ms_free) ms_free(ms_state);
    if (dl_state && dl_free_fn) dl_free_fn(dl_state);
    for (auto& e : events) (void)hipEventDestroy(e);
    if (h_mirror) (void)hipHostFree(h_mirror);
}

hipEvent_t BfsEngine::event(size_t i) {
    while (events.size() <= i) {
        hipEvent_t e;
        GGPU_HIP_TRY(hipEventCreate(&e));
        events.push_back(e);
    }
    return events[i];
}

// The reference's level loop with the decision of bfs_engine.hpp.
void BfsEngine::host_loop(int variant, uint32_t src, const ggpu_bfs_opts& o, std::vector<ggpu_level>* trace) {
    const uint32_t tg = o.tg;
    const uint64_t T = variant == GGPU_BFS_HYBRID ? o.hybrid_t : 0;
    const bool timing = o.time_levels != 0;
    hipStream_t s = ctx->stream;
    if (variant == GGPU_BFS_BINNED || o.policy != GGPU_POLICY_HEURISTIC) ensure_bins();
    BfsDev d = dev();
    LoopState st = loop_start(src, variant, o);
    const PolicyView pv =
        policy_view(variant, o, ctx->policy_set[GGPU_POLICY_SINGLE] ? ctx->policy_host[GGPU_POLICY_SINGLE].data() : nullptr);
    const uint32_t src_bin = bin_of(host.deg(src), o.small_max, o.large_min);
    last_levels = 0;
    last_gpu_ms = 0;
    size_t ev = 0;
    enum { NOWHERE, ON_DEV, ON_HOST } where = NOWHERE;  // who holds dist + frontier
    HybridMirror* hm = T > 0 ? &mirror() : nullptr;

    while (st.size > 0) {
        const auto t0 = std::chrono::steady_clock::now();
        decide_action(st, pv, host.n);
        const bool bu = st.action == GGPU_ACTION_BU;
        const uint64_t work = bu ? st.to_check : st.total;
        const bool cpu = work < T;
        LevelOut r;
        if (cpu) {
            if (where == NOWHERE) {
                steps::init_dist(hm->dist.data(), n, src, o.cpu_threads);
                hm->q[0][0] = src;
                hm->deg[0][0] = host.deg(src);
                st.fmt = FMT_QUEUE;
            } else if (where == ON_DEV) {
                download_state(st, *hm);
            }
            where = ON_HOST;
            r = cpu_level(st, *hm, o);
        } else {
            if (where == NOWHERE) {
                k_init<<<blocks_for(n, 256, grid_cap), 256, 0, s>>>(d, src, src_bin);
                GGPU_HIP_TRY(hipGetLastError());
                st.visited_fresh = 0;
            } else if (where == ON_HOST) {
                upload_state(st, *hm);
            }
            where = ON_DEV;
            if (timing) GGPU_HIP_TRY(hipEventRecord(event(ev), s));
            unsigned long long* slot = stats.p;
            const size_t words = kSlotWords + kConvWords;  // striped counters, then append tails
            GGPU_HIP_TRY(hipMemsetAsync(slot, 0, words * sizeof(unsigned long long), s));
            launch_level(d, st, o, slot, tg);
            if (timing) GGPU_HIP_TRY(hipEventRecord(event(ev + 1), s));
            ev += timing ? 2 : 0;
            auto* pin = (unsigned long long*)ctx_pinned(ctx, 4096);
            GGPU_HIP_TRY(hipMemcpyAsync(pin, slot, words * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
            GGPU_HIP_TRY(hipStreamSynchronize(s));
            unsigned long long sum[kStatWords] = {0};
            for (int t = 0; t < kStripes; ++t)
                for (int k = 0; k < kStatWords; ++k) sum[k] += pin[t * kStatWords + k];
            const unsigned long long* conv = pin + kSlotWords;
            if (st.action == GGPU_ACTION_TD_BINNED) {
                for (int k = 0; k < 3; ++k) r.bins[k] = conv[kTailBin0 + k];
                r.found = r.bins[0] + r.bins[1] + r.bins[2];
            } else {
                r.found = bu ? sum[S_FOUND] : conv[kTailQueue];
            }
            r.examined = bu ? sum[S_EXAMINED] : st.total;
            r.next_edges = sum[S_NEXT_EDGES];
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (trace) trace->push_back(trace_record(st, r, cpu ? 'C' : 'G', ms));
        fold_level(st, r, pv, host.n);
        ++last_levels;
    }
    if (where == ON_HOST) {
        // the last levels ran on the CPU: publish dist to the device
        GGPU_HIP_TRY(hipMemcpyAsync(dist.p, hm->dist.data(), 4ull * n, hipMemcpyHostToDevice, s));
    }
    GGPU_HIP_TRY(hipStreamSynchronize(s));
    if (timing) {
        for (size_t i = 0; i + 1 < ev; i += 2) {
            float ms = 0;
            GGPU_HIP_TRY(hipEventElapsedTime(&ms, events[i], events[i + 1]));
            last_gpu_ms += ms;
        }
    }
}

// gpu_vertex / gpu_naive (runVertex / runNaive, bfs_gpu.mm:115-175): one
// top-down launch per level; the level with nothing discovered is the last
// (it is recorded, as the reference records it).
void BfsEngine::simple_loop(int variant, uint32_t src, const ggpu_bfs_opts& o, std::vector<ggpu_level>* trace) {
    const uint32_t tg = o.tg;
    const bool timing = o.time_levels != 0;
    hipStream_t s = ctx->stream;
    BfsDev d = dev();
    k_init<<<blocks_for(n, 256, grid_cap), 256, 0, s>>>(d, src, 0);
    GGPU_HIP_TRY(hipGetLastError());
    last_levels = 0;
    last_gpu_ms = 0;
    size_t ev = 0;
    uint32_t size = 1, cur = 0;
    for (uint32_t level = 0;; ++level) {
        const auto t0 = std::chrono::steady_clock::now();
        unsigned long long* slot = stats.p;
        const size_t words = kSlotWords + kConvWords;
        GGPU_HIP_TRY(hipMemsetAsync(slot, 0, words * sizeof(unsigned long long), s));
        if (timing) GGPU_HIP_TRY(hipEventRecord(event(ev), s));
        if (variant == GGPU_BFS_VERTEX)
            k_vertex_level<<<blocks_for(n, tg, grid_cap), tg, 0, s>>>(d, level, slot);
        else
            k_naive_level<<<blocks_for(size, tg, grid_cap), tg, 0, s>>>(d, d.q[cur], d.q[cur ^ 1], size, level, slot);
        GGPU_HIP_TRY(hipGetLastError());
        if (timing) GGPU_HIP_TRY(hipEventRecord(event(ev + 1), s));
        ev += timing ? 2 : 0;
        auto* pin = (unsigned long long*)ctx_pinned(ctx, 4096);
        GGPU_HIP_TRY(hipMemcpyAsync(pin, slot, words * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
        unsigned long long sum[kStatWords] = {0};
        for (int t = 0; t < kStripes; ++t)
            for (int k = 0; k < kStatWords; ++k) sum[k] += pin[t * kStatWords + k];
        const unsigned long long found = variant == GGPU_BFS_VERTEX ? sum[S_FOUND] : pin[kSlotWords + kTailQueue];
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (trace) {
            ggpu_level L{};
            L.level = level;
            L.action = GGPU_ACTION_TD_EDGE;
            L.frontier = sum[S_FRONTIER];
            L.edges = sum[S_EXAMINED];
            L.frontier_edges = sum[S_EXAMINED];
            L.direction = 'T';
            L.device = 'G';
            L.time_ms = ms;
            trace->push_back(L);
        }
        ++last_levels;
        if (found == 0) break;
        size = (uint32_t)found;
        cur ^= 1;
    }
    if (timing) {
        for (size_t i = 0; i + 1 < ev; i += 2) {
            float ms = 0;
            GGPU_HIP_TRY(hipEventElapsedTime(&ms, events[i], events[i + 1]));
            last_gpu_ms += ms;
        }
    }
}

// Enqueues one GPU level (conversion if the frontier is in another format,
// then the action's kernels). Grids are sized exactly from the host state.
void BfsEngine::launch_level(BfsDev& d, LoopState& st, const ggpu_bfs_opts& o, unsigned long long* slot, uint32_t tg) {
    hipStream_t s = ctx->stream;
    const uint32_t level = st.level, cur = st.cur;
    const uint32_t act = st.action;
    if (act == GGPU_ACTION_BU) {
        if (st.fmt != FMT_BITMAP || !st.visited_fresh) {
            k_dist_to_bitmap<<<blocks_for(nw, tg / kWave, grid_cap), tg, 0, s>>>(d, cur, level);
            GGPU_HIP_TRY(hipGetLastError());
        }
        k_bu<<<blocks_for(nw, tg / kWave, grid_cap), tg, 0, s>>>(d, cur, level, slot);
        GGPU_HIP_TRY(hipGetLastError());
        if (n_heavy) {
            k_bu_heavy<<<blocks_for(n_heavy, tg / kWave, grid_cap), tg, 0, s>>>(d, cur, level, slot);
            GGPU_HIP_TRY(hipGetLastError());
        }
        return;
    }
    if (act == GGPU_ACTION_TD_EDGE) {
        if (st.fmt != FMT_QUEUE) {
            if (st.fmt == FMT_BITMAP)
                k_bitmap_to_queue<<<blocks_for(nw, tg, grid_cap), tg, 0, s>>>(d, cur, cur);
            else
                k_dist_collect<<<blocks_for(n, tg, grid_cap), tg, 0, s>>>(d, level, 0, cur, o.small_max, o.large_min);
            GGPU_HIP_TRY(hipGetLastError());
        }
        if (st.total == 0) return;
        const uint32_t size = (uint32_t)st.size;
        const uint32_t span = 256 * kScanItems;
        const uint32_t ntiles = (size + span - 1) / span;
        k_scan_tiles<<<blocks_for(ntiles, 1, grid_cap), 256, 0, s>>>(d.deg[cur], size, d.tile);
        k_scan_apply<<<blocks_for(ntiles, 1, grid_cap), 256, 0, s>>>(d.deg[cur], size, d.tile);
        k_td_edge<<<blocks_for(st.total, tg * o.ept, 1u << 20), tg, 0, s>>>(d, cur, level, size, (uint32_t)st.total,
                                                                             o.ept, slot);
        GGPU_HIP_TRY(hipGetLastError());
        return;
    }
    // binned top-down
    if (st.fmt != FMT_BINS) {
        k_dist_collect<<<blocks_for(n, tg, grid_cap), tg, 0, s>>>(d, level, 1, cur, o.small_max, o.large_min);
        GGPU_HIP_TRY(hipGetLastError());
        // bin sizes of a collected frontier: count on the host from the level's frontier size split
        unsigned long long* pin = (unsigned long long*)ctx_pinned(ctx, 4096);
        GGPU_HIP_TRY(hipMemcpyAsync(pin + 16, d.conv + 1, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
        for (int k = 0; k < 3; ++k) st.bins[k] = pin[16 + k];
    }
    const uint32_t wpb = tg / kWave;
    const uint64_t need =
        std::max<uint64_t>({st.bins[2], (st.bins[1] + wpb - 1) / wpb, (st.bins[0] + tg - 1) / tg, (uint64_t)1});
    k_td_binned<<<(uint32_t)std::min<uint64_t>(need, 1u << 20), tg, 0, s>>>(
        d, cur, level, o.small_max, o.large_min, (uint32_t)st.bins[0], (uint32_t)st.bins[1], (uint32_t)st.bins[2],
        slot);
    GGPU_HIP_TRY(hipGetLastError());
}

// ---------------------------------------------------------------- hybrid
HybridMirror& BfsEngine::mirror() {
    if (!mirror_ready) {
        hmirror.dist.resize(n);
        for (auto& v : hmirror.q) v.resize(std::max<uint32_t>(n, 1));
        for (auto& v : hmirror.deg) v.resize(std::max<uint32_t>(n, 1));
        for (auto& v : hmirror.bm) v.assign(nw + 1, 0);
        mirror_ready = true;
    }
    return hmirror;
}

// Device -> host: dist and the frontier in its current format.
void BfsEngine::download_state(LoopState& st, HybridMirror& hm) {
    hipStream_t s = ctx->stream;
    GGPU_HIP_TRY(hipMemcpyAsync(hm.dist.data(), dist.p, 4ull * n, hipMemcpyDeviceToHost, s));
    BfsDev d = dev();
    if (st.fmt == FMT_BITMAP) {
        GGPU_HIP_TRY(hipMemcpyAsync(hm.bm[st.cur].data(), d.bm[st.cur], 8ull * nw, hipMemcpyDeviceToHost, s));
    } else if (st.fmt == FMT_QUEUE) {
        GGPU_HIP_TRY(hipMemcpyAsync(hm.q[st.cur].data(), d.q[st.cur], 4ull * st.size, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipMemcpyAsync(hm.deg[st.cur].data(), d.deg[st.cur], 4ull * st.size, hipMemcpyDeviceToHost, s));
    } else {
        // bins: rebuild a queue from dist on the host
        uint64_t k = 0;
        GGPU_HIP_TRY(hipStreamSynchronize(s));
        for (uint32_t v = 0; v < n; ++v)
            if (hm.dist[v] == st.level) {
                hm.q[st.cur][k] = v;
                hm.deg[st.cur][k] = host.deg(v);
                ++k;
            }
        st.fmt = FMT_QUEUE;
    }
    GGPU_HIP_TRY(hipStreamSynchronize(s));
}

// Host -> device.
void BfsEngine::upload_state(LoopState& st, HybridMirror& hm) {
    hipStream_t s = ctx->stream;
    BfsDev d = dev();
    GGPU_HIP_TRY(hipMemcpyAsync(dist.p, hm.dist.data(), 4ull * n, hipMemcpyHostToDevice, s));
    if (st.fmt == FMT_BITMAP) {
        GGPU_HIP_TRY(hipMemcpyAsync(d.bm[st.cur], hm.bm[st.cur].data(), 8ull * nw, hipMemcpyHostToDevice, s));
    } else {
        GGPU_HIP_TRY(hipMemcpyAsync(d.q[st.cur], hm.q[st.cur].data(), 4ull * st.size, hipMemcpyHostToDevice, s));
        GGPU_HIP_TRY(hipMemcpyAsync(d.deg[st.cur], hm.deg[st.cur].data(), 4ull * st.size, hipMemcpyHostToDevice, s));
    }
    st.visited_fresh = 0;  // CPU levels do not maintain the device's visited bitmap
    GGPU_HIP_TRY(hipStreamSynchronize(s));
}

// One level on the CPU (bfs_steps.cpp) over the host mirror; the conversion
// into the action's format happens here when the direction changed.
LevelOut BfsEngine::cpu_level(LoopState& st, HybridMirror& hm, const ggpu_bfs_opts& o) {
    LevelOut r;
    const uint32_t cur = st.cur;
    if (st.action == GGPU_ACTION_BU) {
        if (st.fmt != FMT_BITMAP) {
            steps::queue_to_bitmap(hm.q[cur].data(), st.size, hm.bm[cur].data(), n, o.cpu_threads);
            st.fmt = FMT_BITMAP;
        }
        steps::Out so = steps::bottom_up(host, hm.dist.data(), hm.bm[cur].data(), hm.bm[cur ^ 1].data(), st.level,
                                         o.cpu_threads);
        r.found = so.found;
        r.examined = so.examined;
        r.next_edges = so.next_edges;
    } else {
        if (st.fmt != FMT_QUEUE) {
            steps::bitmap_to_queue(host, hm.bm[cur].data(), hm.q[cur].data(), hm.deg[cur].data(), o.cpu_threads);
            st.fmt = FMT_QUEUE;
        }
        steps::Out so = steps::top_down(host, hm.dist.data(), hm.q[cur].data(), st.size, st.total, hm.q[cur ^ 1].data(),
                                        hm.deg[cur ^ 1].data(), st.level, o.cpu_threads);
        r.found = so.found;
        r.examined = st.total;
        r.next_edges = so.next_edges;
    }
    return r;
}

}  // namespace ggpu

// ------------------------------------------------------------------ C-ABI
using namespace ggpu;

struct ggpu_bfs {
    ggpu_ctx* ctx;
    BfsEngine* eng;
};

extern "C" {

void ggpu_bfs_opts_default(ggpu_bfs_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->tg = 256;
    o->small_max = 16;
    o->large_min = 256;
    o->ept = 2;
    o->alpha = 15.0;
    o->beta = 18.0;
    o->hybrid_t = 100000;
    o->cpu_threads = 0;
    o->loop = GGPU_LOOP_HOST;
    o->policy = GGPU_POLICY_HEURISTIC;
    o->graph_levels = 8;
    o->time_levels = 1;
    o->batch_order = GGPU_BATCH_ORDER_AUTO;
}

int ggpu_bfs_parse_variant(const char* name, int* variant) {
    if (!name || !variant) return GGPU_EINVAL;
    const std::string s = name;
    if (s == "gpu_edge") *variant = GGPU_BFS_EDGE;
    else if (s == "gpu_binned") *variant = GGPU_BFS_BINNED;
    else if (s == "gpu_do") *variant = GGPU_BFS_DO;
    else if (s == "hybrid") *variant = GGPU_BFS_HYBRID;
    else if (s == "gpu_vertex") *variant = GGPU_BFS_VERTEX;
    else if (s == "gpu_naive") *variant = GGPU_BFS_NAIVE;
    else return GGPU_EINVAL;
    return GGPU_OK;
}

int ggpu_bfs_create(ggpu_ctx* ctx, uint64_t n, uint64_t m, const uint32_t* row_ptr, const uint32_t* col_idx,
                    ggpu_bfs** out) {
    if (!ctx || !out || !row_ptr || (m && !col_idx)) return set_error(ctx, GGPU_EINVAL, "null argument");
    *out = nullptr;
    BfsEngine* e = nullptr;
    int rc = guarded(ctx, [&] { e = new BfsEngine(ctx, n, m, row_ptr, col_idx); });
    if (rc != GGPU_OK) return rc;
    *out = new ggpu_bfs{ctx, e};
    return GGPU_OK;
}

void ggpu_bfs_destroy(ggpu_bfs* b) {
    if (!b) return;
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b->eng;
    delete b;
}

int ggpu_bfs_run(ggpu_bfs* b, int variant, uint32_t source, const ggpu_bfs_opts* opt, ggpu_level* trace,
                 size_t trace_cap, size_t* trace_len) {
    if (!b) return GGPU_EINVAL;
    ggpu_ctx* ctx = b->ctx;
    ggpu_bfs_opts o;
    ggpu_bfs_opts_default(&o);
    if (opt) o = *opt;
    return guarded(ctx, [&] {
        BfsEngine& e = *b->eng;
        if (variant < GGPU_BFS_EDGE || variant > GGPU_BFS_NAIVE) throw ArgError("unknown BFS variant");
        if (source >= e.n) throw ArgError("source out of range");
        check_tg(o.tg);
        if (o.ept == 0) throw ArgError("ept must be >= 1");
        if (o.tg * o.ept > (1u << 20)) throw ArgError("tg * ept too large");
        if (o.policy == GGPU_POLICY_SCRIPTED && (!o.actions || o.n_actions == 0))
            throw ArgError("scripted policy needs actions");
        if (o.policy > GGPU_POLICY_SCRIPTED) throw ArgError("unknown policy");
        if (o.loop > GGPU_LOOP_PERSISTENT) throw ArgError("unknown loop mode");
        if (variant == GGPU_BFS_HYBRID && (o.loop != GGPU_LOOP_HOST || o.policy != GGPU_POLICY_HEURISTIC))
            throw ArgError("hybrid runs with the host loop and the heuristic policy");
        const bool simple = variant == GGPU_BFS_VERTEX || variant == GGPU_BFS_NAIVE;
        if (simple && (o.loop != GGPU_LOOP_HOST || o.policy != GGPU_POLICY_HEURISTIC))
            throw ArgError("gpu_vertex and gpu_naive run top-down with the host loop and the heuristic policy");
        GGPU_HIP_TRY(hipSetDevice(ctx->device));
        std::vector<ggpu_level> tr;
        std::vector<ggpu_level>* trp = (trace || trace_len) ? &tr : nullptr;
        if (simple) e.simple_loop(variant, source, o, trp);
        else if (o.loop == GGPU_LOOP_HOST) e.host_loop(variant, source, o, trp);
        else e.device_loop(variant, source, o, trp);
        if (trace) std::memcpy(trace, tr.data(), std::min(tr.size(), trace_cap) * sizeof(ggpu_level));
        if (trace_len) *trace_len = tr.size();
    });
}

int ggpu_bfs_fetch_dist(ggpu_bfs* b, uint32_t* out) {
    if (!b || !out) return GGPU_EINVAL;
    return guarded(b->ctx, [&] {
        GGPU_HIP_TRY(hipMemcpyAsync(out, b->eng->dist.p, 4ull * b->eng->n, hipMemcpyDeviceToHost, b->ctx->stream));
        GGPU_HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    });
}

double ggpu_bfs_last_gpu_ms(const ggpu_bfs* b) { return b ? b->eng->last_gpu_ms : 0.0; }
uint32_t ggpu_bfs_last_levels(const ggpu_bfs* b) { return b ? b->eng->last_levels : 0; }

}  // extern "C"

// Internal accessor for the other translation units of this library.
ggpu::BfsEngine* ggpu_bfs_engine(ggpu_bfs* b) { return b ? b->eng : nullptr; }

// ------------------------------------------------------------- test hooks
// the level-loop primitives on their own, for the tests
#include "ggpu_hip_testing.h"

extern "C" {

int ggpu_test_scan(ggpu_ctx* ctx, uint32_t n, const uint32_t* in, uint32_t* out, uint32_t max_blocks) {
    if (!ctx || (n && (!in || !out))) return GGPU_EINVAL;
    return guarded(ctx, [&] {
        GGPU_HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        const uint32_t span = 256 * kScanItems;
        const uint32_t ntiles = (uint32_t)(((uint64_t)n + span - 1) / span);
        const uint32_t cap = max_blocks ? max_blocks : 4096;  // the engine's grid cap
        DBuf<uint32_t> v, tile;
        v.alloc(n ? n : 1);
        tile.alloc(ntiles ? ntiles : 1);
        if (n) GGPU_HIP_TRY(hipMemcpyAsync(v.p, in, 4ull * n, hipMemcpyHostToDevice, s));
        if (ntiles) {
            // as the edge-balanced top-down launches it: 256-thread blocks, one tile per block
            k_scan_tiles<<<blocks_for(ntiles, 1, cap), 256, 0, s>>>(v.p, n, tile.p);
            k_scan_apply<<<blocks_for(ntiles, 1, cap), 256, 0, s>>>(v.p, n, tile.p);
            GGPU_HIP_TRY(hipGetLastError());
        }
        if (n) GGPU_HIP_TRY(hipMemcpyAsync(out, v.p, 4ull * n, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
    });
}

int ggpu_test_frontier_convert(ggpu_bfs* b, const uint32_t* dist, uint32_t level, uint32_t tg, uint64_t* bitmap,
                               uint32_t* queue, uint32_t* degrees, uint32_t* count) {
    if (!b || !dist || !bitmap || !queue || !degrees || !count) return GGPU_EINVAL;
    return guarded(b->ctx, [&] {
        BfsEngine& e = *b->eng;
        check_tg(tg);
        GGPU_HIP_TRY(hipSetDevice(e.ctx->device));
        hipStream_t s = e.ctx->stream;
        e.ensure_edge_buffers();
        e.ensure_bitmaps();
        BfsDev d = e.dev();
        if (e.n) GGPU_HIP_TRY(hipMemcpyAsync(d.dist, dist, 4ull * e.n, hipMemcpyHostToDevice, s));
        // the bitmap starts all ones (the conversion must write every word), the counter at zero
        GGPU_HIP_TRY(hipMemsetAsync(d.bm[0], 0xFF, 8ull * std::max(e.nw, 1u), s));
        GGPU_HIP_TRY(hipMemsetAsync(d.conv, 0, sizeof(unsigned long long), s));
        k_dist_to_bitmap<<<blocks_for(e.nw, tg / kWave, e.grid_cap), tg, 0, s>>>(d, 0, level);
        k_bitmap_to_queue<<<blocks_for(e.nw, tg, e.grid_cap), tg, 0, s>>>(d, 0, 0);
        GGPU_HIP_TRY(hipGetLastError());
        unsigned long long c = 0;
        if (e.nw) GGPU_HIP_TRY(hipMemcpyAsync(bitmap, d.bm[0], 8ull * e.nw, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipMemcpyAsync(&c, d.conv, sizeof c, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
        if (c > e.n) throw ArgError("bitmap_to_queue counted more vertices than the graph has", GGPU_ESTATE);
        if (c) {
            GGPU_HIP_TRY(hipMemcpyAsync(queue, d.q[0], 4ull * c, hipMemcpyDeviceToHost, s));
            GGPU_HIP_TRY(hipMemcpyAsync(degrees, d.deg[0], 4ull * c, hipMemcpyDeviceToHost, s));
            GGPU_HIP_TRY(hipStreamSynchronize(s));
        }
        *count = (uint32_t)c;
    });
}

}  // extern "C"
