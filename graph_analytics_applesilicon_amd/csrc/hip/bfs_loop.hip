// Device-driven single-source BFS level loops (no reference counterpart; the
// reference syncs the host every level, src/metal/bfs_gpu.mm:83-104):
//   GGPU_LOOP_PERSISTENT  one persistent launch runs every level. Workgroup 0
//                         alone folds the counters and decides each level; it
//                         runs small top-down levels by itself and publishes
//                         the others as commands the grid polls
//                         (dl_protocol.hpp);
//   GGPU_LOOP_GRAPH       a captured hipGraph of K levels (a decide kernel and a
//                         level kernel per level) is replayed until the
//                         frontier is empty; the host checks once per K levels.
// Both make exactly the host loop's decisions: workgroup 0 (persistent) or the
// decide kernel (graph) runs decide_action / fold_level of bfs_engine.hpp on
// the same per-level counters, so the trace (frontier, direction) equals the
// host loop's and the oracle's.
//
// State: dist[n] (INF before the launch), visited bitmap vis, frontier
// bitmaps in three slots (level l reads slot l%3, writes (l+1)%3 and clears
// (l+2)%3, the slot of level l-1). One phase per level:
//   top-down: each frontier vertex pushes to its neighbours; atomicOr on vis
//     returns whether this push discovered the vertex (exactly once);
//   bottom-up: each unvisited vertex scans its neighbours for a frontier bit.
// Vertices of degree >= kHeavyMin are handled as kHeavyChunk-edge wave items
// in both directions (the engine's heavy list), the rest lane per vertex.
// In the persistent loop every datum another workgroup reads during the
// launch (counters, frontier and visited words, the command word) is written
// with device-scope atomics or write-through stores and read with coherent
// loads; the read-only CSR stays cached. dist has no reader during the
// launch (the host reads it afterwards); each entry is written once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>

#include "bfs_engine.hpp"
#include "dl_protocol.hpp"
#include "grid_sync.hpp"

namespace ggpu {

constexpr uint32_t kDlThreads = 1024;      // persistent: one block per CU
constexpr uint32_t kDlLevelThreads = 256;  // graph mode level kernel
constexpr uint32_t kDlParts = 4;           // counters per block: found, next_edges, examined, pad

// Per-run parameters and the live loop state, in device memory (the graph's
// kernels read them; their arguments are fixed at capture).
struct DlState {
    LoopState st;
    PolicyView pv;
    uint32_t started, prev_action, levels, done;
    uint64_t t_prev;
    double tick_ms;
};

struct DlDev {
    const uint32_t* rp;
    const uint32_t* ci;
    uint32_t* dist;
    uint64_t* vis;                  // nb
    uint64_t* fb;                   // 3 x nb
    unsigned long long* part;       // 2 x grid x kDlParts
    const uint4* heavy;             // {v, first edge, end edge, degree}, heaviest first
    DlState* ds;
    ggpu_level* trace;              // null: no trace
    uint32_t* bar;                  // persistent: arrival counters (grid_sync.hpp layout)
    uint32_t* abort;                // persistent: a workgroup gave up waiting
    unsigned long long* cmd;        // persistent: workgroup 0's command word (dl_protocol.hpp)
    uint32_t n_heavy, n, nb, trace_cap;
    uint32_t heavy_min;             // degree from which a vertex is left to the heavy items (none: ~0u)
    uint64_t m;
};

struct DlAcc {
    unsigned long long found = 0, next_edges = 0, examined = 0;
};

__device__ __forceinline__ uint32_t dl_gwave() { return (blockIdx.x * blockDim.x + threadIdx.x) / kWave; }
__device__ __forceinline__ uint32_t dl_nwaves() { return gridDim.x * blockDim.x / kWave; }

// u reached from the frontier at depth `depth`: the atomicOr on vis decides
// which pusher discovered it.
__device__ __forceinline__ void dl_discover(const DlDev& D, uint64_t* fbn, uint32_t u, uint32_t depth, DlAcc& a) {
    const uint64_t bit = 1ull << (u & 63);
    if (D.vis[u >> 6] & bit) return;  // visited before this level (a stale 0 only costs an atomic)
    const uint64_t old = atomicOr((unsigned long long*)&D.vis[u >> 6], (unsigned long long)bit);
    if (old & bit) return;
    store_through(&D.dist[u], depth);
    atomicOr((unsigned long long*)&fbn[u >> 6], (unsigned long long)bit);
    a.found += 1;
    a.next_edges += D.rp[u + 1] - D.rp[u];
}

__device__ void dl_top_down(const DlDev& D, uint32_t level, bool lane_per_vertex, DlAcc& a) {
    const uint64_t* fb = D.fb + (size_t)(level % 3) * D.nb;
    uint64_t* fbn = D.fb + (size_t)((level + 1) % 3) * D.nb;
    const uint32_t lane = lane_id(), depth = level + 1;
    const uint32_t stride = gridDim.x * blockDim.x;
    // light frontier vertices: each lane owns a frontier word, one vertex per round
    for (uint32_t wi = blockIdx.x * blockDim.x + threadIdx.x; wi - lane < D.nb; wi += stride) {
        uint64_t fw = wi < D.nb ? load_coherent(&fb[wi]) : 0ull;
        while (__ballot(fw != 0)) {
            uint32_t v = 0xFFFFFFFFu, b = 0, e = 0;
            if (fw) {
                v = wi * 64 + (uint32_t)(__ffsll((long long)fw) - 1);
                fw &= fw - 1;
                b = D.rp[v];
                e = D.rp[v + 1];
                if (e - b >= D.heavy_min) v = 0xFFFFFFFFu;  // heavy: pushed by its items below
            }
            if (lane_per_vertex) {
                if (v != 0xFFFFFFFFu)
                    for (uint32_t x = b; x < e; ++x) dl_discover(D, fbn, D.ci[x], depth, a);
            } else {
                uint64_t todo = __ballot(v != 0xFFFFFFFFu);
                while (todo) {
                    const int j = __ffsll((long long)todo) - 1;
                    todo &= todo - 1;
                    const uint32_t bj = __shfl(b, j, kWave), ej = __shfl(e, j, kWave);
                    for (uint32_t x = bj + lane; x < ej; x += kWave) dl_discover(D, fbn, D.ci[x], depth, a);
                }
            }
        }
    }
    // heavy frontier vertices: 64 items nwaves apart per wave, the frontier ones pushed in turn
    for (uint64_t r = 0; dl_gwave() + r * 64 * dl_nwaves() < D.n_heavy; ++r) {
        const uint64_t it = dl_gwave() + (r * 64 + lane) * dl_nwaves();
        uint4 item = make_uint4(0, 0, 0, 0);
        bool want = false;
        if (it < D.n_heavy) {
            item = D.heavy[it];
            want = (load_coherent(&fb[item.x >> 6]) >> (item.x & 63)) & 1ull;
        }
        uint64_t todo = __ballot(want);
        while (todo) {
            const int j = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            const uint32_t y = __shfl(item.y, j, kWave), z = __shfl(item.z, j, kWave);
            for (uint32_t x = y + lane; x < z; x += kWave) dl_discover(D, fbn, D.ci[x], depth, a);
        }
    }
}

__device__ void dl_bottom_up(const DlDev& D, uint32_t level, DlAcc& a) {
    const uint64_t* fb = D.fb + (size_t)(level % 3) * D.nb;
    uint64_t* fbn = D.fb + (size_t)((level + 1) % 3) * D.nb;
    const uint32_t lane = lane_id(), depth = level + 1;
    // light vertices: lane per vertex, first frontier neighbour ends the scan
    for (uint32_t wd = dl_gwave(); wd < D.nb; wd += dl_nwaves()) {
        const uint32_t v = wd * 64 + lane;
        const uint64_t seen = load_coherent(&D.vis[wd]);
        bool found = false;
        if (v < D.n && !((seen >> lane) & 1ull)) {
            const uint32_t b = D.rp[v], e = D.rp[v + 1];
            if (e > b && e - b < D.heavy_min) {
                for (uint32_t x = b; x < e && !found; x += 4) {
                    const uint32_t c = min(4u, e - x);
                    uint32_t u[4];
#pragma unroll
                    for (uint32_t i = 0; i < 4; ++i) u[i] = i < c ? D.ci[x + i] : 0u;
                    uint32_t hit = 4;
#pragma unroll
                    for (int i = 3; i >= 0; --i)
                        if ((uint32_t)i < c && ((load_coherent(&fb[u[i] >> 6]) >> (u[i] & 63)) & 1ull)) hit = (uint32_t)i;
                    if (hit < 4) {
                        found = true;
                        a.examined += hit + 1;
                    } else {
                        a.examined += c;
                    }
                }
                if (found) {
                    store_through(&D.dist[v], depth);
                    a.next_edges += e - b;
                }
            }
        }
        const uint64_t mask = __ballot(found);
        if (lane == 0 && mask) {
            atomicOr((unsigned long long*)&D.vis[wd], (unsigned long long)mask);
            atomicOr((unsigned long long*)&fbn[wd], (unsigned long long)mask);
            a.found += (uint64_t)__popcll(mask);
        }
    }
    // heavy vertices: items of unvisited vertices, scanned 256 edges at a time until a frontier hit
    for (uint64_t r = 0; dl_gwave() + r * 64 * dl_nwaves() < D.n_heavy; ++r) {
        const uint64_t it = dl_gwave() + (r * 64 + lane) * dl_nwaves();
        uint4 item = make_uint4(0, 0, 0, 0);
        bool want = false;
        if (it < D.n_heavy) {
            item = D.heavy[it];
            want = !((load_coherent(&D.vis[item.x >> 6]) >> (item.x & 63)) & 1ull);
        }
        uint64_t todo = __ballot(want);
        while (todo) {
            const int j = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            const uint32_t v = __shfl(item.x, j, kWave), y = __shfl(item.y, j, kWave), z = __shfl(item.z, j, kWave);
            bool any = false;
            for (uint32_t xb = y; xb < z && !any; xb += 4 * kWave) {
                bool hit = false;
#pragma unroll
                for (uint32_t i = 0; i < 4; ++i) {
                    const uint32_t x = xb + i * kWave + lane;
                    if (x < z) {
                        const uint32_t u = D.ci[x];
                        hit |= ((load_coherent(&fb[u >> 6]) >> (u & 63)) & 1ull) != 0;
                        a.examined += 1;
                    }
                }
                any = __ballot(hit) != 0;
            }
            if (any && lane == 0) {
                const uint64_t bit = 1ull << (v & 63);
                const uint64_t old = atomicOr((unsigned long long*)&D.vis[v >> 6], (unsigned long long)bit);
                if (!(old & bit)) {
                    store_through(&D.dist[v], depth);
                    atomicOr((unsigned long long*)&fbn[v >> 6], (unsigned long long)bit);
                    a.found += 1;
                    a.next_edges += D.rp[v + 1] - D.rp[v];
                }
            }
        }
    }
}

// One level: the action's phase, the clear of the slot of level - 1, the block's counters.
__device__ void dl_level(const DlDev& D, uint32_t level, uint32_t action, unsigned long long* part_slot) {
    DlAcc a;
    if (action == GGPU_ACTION_BU) dl_bottom_up(D, level, a);
    else dl_top_down(D, level, action == GGPU_ACTION_TD_BINNED, a);
    uint64_t* old = D.fb + (size_t)((level + 2) % 3) * D.nb;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < D.nb; i += gridDim.x * blockDim.x)
        store_through(&old[i], 0ull);
    __shared__ unsigned long long s_red[kDlThreads / kWave];
    const unsigned long long f = block_sum(a.found, s_red);
    const unsigned long long ne = block_sum(a.next_edges, s_red);
    const unsigned long long ex = block_sum(a.examined, s_red);
    if (threadIdx.x == 0) {
        unsigned long long* p = part_slot + (size_t)blockIdx.x * kDlParts;
        store_through(&p[0], f);
        store_through(&p[1], ne);
        store_through(&p[2], ex);
    }
}

// The level's counters summed over `blocks` entries, in every thread of the block.
__device__ LevelOut dl_totals(const unsigned long long* part_slot, uint32_t blocks) {
    __shared__ unsigned long long s_red[kDlThreads / kWave];
    unsigned long long v[3] = {0, 0, 0};
    for (uint32_t i = threadIdx.x; i < blocks; i += blockDim.x)
        for (int j = 0; j < 3; ++j) v[j] += load_coherent(&part_slot[(size_t)i * kDlParts + j]);
    LevelOut r;
    r.found = block_sum(v[0], s_red);
    r.next_edges = block_sum(v[1], s_red);
    r.examined = block_sum(v[2], s_red);
    return r;
}

// The loop state of one workgroup, in LDS: one copy per workgroup instead of
// one per lane (it is wave-uniform; in registers it was spilled every level).
struct DlBlock {
    LoopState st;
    PolicyView pv;
    uint64_t t_prev;   // writer: wall clock when the current level was decided
    uint32_t prev_action;
    float y[3];        // learned policy outputs
};

// The learned policy's outputs for B.st into B.y; wave 0. Out of line: the
// network's registers would otherwise be reserved across every level's code.
__device__ __noinline__ void dl_policy_eval(DlBlock& B, const float* mlp_lds) {
    float f[5], y[3];
    policy_features(B.st.size, B.st.total, B.st.unvisited_edges, B.st.level, B.pv.mean_degree, f);
    mlp_forward_wave(mlp_lds, f, y);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) B.y[q] = y[q];
}

// Thread 0's share of a decision.
__device__ void dl_fold_one(const DlDev& D, DlBlock& B, const LevelOut& r, bool writer, double tick_ms) {
    if (writer && D.trace && B.st.level < D.trace_cap) {
        const uint64_t now = wall_clock64();
        ggpu_level& L = D.trace[B.st.level];
        if (B.prev_action == GGPU_ACTION_BU) L.edges = r.examined;
        L.time_ms = (double)(now - B.t_prev) * tick_ms;
    }
    fold_level(B.st, r, B.pv, D.n);
}

__device__ void dl_decide_one(const DlDev& D, DlBlock& B, bool learned, bool writer) {
    decide_action(B.st, B.pv, D.n, learned ? B.y : nullptr);
    B.prev_action = B.st.action;
    if (writer) {
        B.t_prev = wall_clock64();
        if (D.trace && B.st.level < D.trace_cap) {
            LevelOut none;
            none.examined = B.st.total;
            D.trace[B.st.level] = trace_record(B.st, none, 'G', 0.0);
        }
    }
}

__device__ bool dl_decide_rest(const DlDev& D, DlBlock& B, bool writer, const float* mlp_lds);

// The decision step shared by both loops, called by every thread of the
// block: thread 0 folds the finished level's totals `r` (null before the
// first level) into B.st and decides the next action; wave 0 evaluates the
// learned policy. Returns false when the traversal is over. `writer`: this
// block writes the trace.
__device__ bool dl_decide_core(const DlDev& D, DlBlock& B, const LevelOut* r, bool writer, double tick_ms,
                               const float* mlp_lds) {
    if (r) {
        if (threadIdx.x == 0) dl_fold_one(D, B, *r, writer, tick_ms);
        __syncthreads();
    }
    return dl_decide_rest(D, B, writer, mlp_lds);
}

// The part of dl_decide_core after the fold (B.st folded and visible to the block).
__device__ bool dl_decide_rest(const DlDev& D, DlBlock& B, bool writer, const float* mlp_lds) {
    if (B.st.size == 0) return false;
    const bool learned = B.pv.mode == GGPU_POLICY_LEARNED && B.pv.mlp && B.pv.variant != GGPU_BFS_EDGE &&
                         B.pv.variant != GGPU_BFS_BINNED;
    if (learned && threadIdx.x < kWave) dl_policy_eval(B, mlp_lds);
    if (learned) __syncthreads();
    if (threadIdx.x == 0) dl_decide_one(D, B, learned, writer);
    __syncthreads();
    return true;
}

__device__ bool dl_decide(const DlDev& D, DlBlock& B, bool started, const unsigned long long* part_slot,
                          uint32_t blocks, bool writer, double tick_ms, const float* mlp_lds) {
    LevelOut r;
    if (started) r = dl_totals(part_slot, blocks);
    return dl_decide_core(D, B, started ? &r : nullptr, writer, tick_ms, mlp_lds);
}

// ------------------------------------------------- single-workgroup levels
// A level whose frontier is small costs the full-grid loop a chain of
// cross-XCD round trips (coherent counter reads, device-scope atomics, the
// arrival wait) that does not shrink with the frontier. While the decision is
// top-down and the frontier has at most kSbMax vertices and kSbEdges edges
// (sb_ok, dl_protocol.hpp), workgroup 0 runs levels alone: frontier queues in
// LDS, no arrivals; the other workgroups keep polling the command word. A
// level is edge-parallel: the frontier's row offsets are loaded once (their
// sum is also the previous level's next_edges) and scanned into an edge ->
// vertex map in LDS, then every edge's col_idx load and visited atomic is
// issued at once, so a level is about three memory round trips. The decisions
// come from the same fold / decide code with the same counters (found,
// next_edges), so the trace equals the other loops'. The frontier bitmaps:
// on entry the frontier is read out of its slot into LDS and the three slots
// are cleared (write-through); the levels then keep their frontiers in LDS
// and set no bitmap bits, except for vertices beyond the queue's capacity
// (which ends the run of single-workgroup levels); on leaving, the queued
// frontier's bits are added to its slot, so the grid finds the slots as it
// keeps them (frontier in slot L % 3, slot (L + 1) % 3 clear).
constexpr uint32_t kSbBatch = 8;  // edges per lane in flight (kSbEdges / blockDim rounds at most 2)

struct SbShared {
    uint32_t q[2][kSbMax];    // frontier queues: current, next
    uint32_t off[kSbMax];     // exclusive prefix of the current frontier's degrees
    uint32_t vb[kSbMax];      // row start of each current frontier vertex
    uint16_t map[kSbEdges];   // frontier index of each edge of the level
    uint32_t qn[2];           // queue counts (may exceed kSbMax: then the queue is cut)
    uint64_t wsum[2][kDlThreads / kWave];
};

// Exclusive prefix over the block (1024 threads) of a u64 per thread, and
// the block total of a second u64; returns the first's total.
__device__ __forceinline__ uint64_t sb_scan(uint64_t v, uint64_t& excl, uint64_t v2, uint64_t& total2,
                                            uint64_t (*wsum)[kDlThreads / kWave]) {
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
    uint64_t x = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint64_t y = __shfl_up(x, d, kWave);
        if (lane >= (uint32_t)d) x += y;
    }
    const uint64_t x2 = wave_sum(v2);
    __syncthreads();
    if (lane == kWave - 1) wsum[0][w] = x;
    if (lane == 0) wsum[1][w] = x2;
    __syncthreads();
    uint64_t before = 0, total = 0, t2 = 0;
    for (uint32_t i = 0; i < kDlThreads / kWave; ++i) {
        before += i < w ? wsum[0][i] : 0ull;
        total += wsum[0][i];
        t2 += wsum[1][i];
    }
    excl = before + x - v;
    total2 = t2;
    return total;
}

// Workgroup 0 only. On entry B.st is decided for level L = B.st.level with
// sb_ok(B.st); fb slot L % 3 holds the frontier, slot (L + 1) % 3 is clear.
// Runs levels until the traversal ends (returns false) or a level is decided
// that sb_ok rejects (returns true with B.st decided for that level L', its
// frontier in slot L' % 3 and the other two slots clear).
__device__ bool sb_levels(const DlDev& D, DlBlock& B, double tick_ms, const float* mlp_lds, SbShared& S) {
    const uint32_t tid = threadIdx.x, lane = lane_id();
    if (tid < 2) S.qn[tid] = 0;
    __syncthreads();
    {  // the frontier out of its slot; all three slots cleared
        uint64_t* fb = D.fb + (size_t)(B.st.level % 3) * D.nb;
        uint64_t* f2 = D.fb + (size_t)((B.st.level + 2) % 3) * D.nb;
        for (uint32_t i = tid; i < D.nb; i += blockDim.x) {
            uint64_t w = load_coherent(&fb[i]);
            if (w) store_through(&fb[i], 0ull);
            store_through(&f2[i], 0ull);
            while (w) {
                const uint32_t b = (uint32_t)__ffsll((long long)w) - 1;
                w &= w - 1;
                const uint32_t p = atomicAdd(&S.qn[0], 1u);
                if (p < kSbMax) S.q[0][p] = i * 64 + b;
            }
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    uint64_t cut_edges = 0;  // degrees of vertices found beyond the queue's capacity (this thread's)
    for (uint32_t cur = 0, first = 1;; first = 0) {
        const uint32_t nxt = cur ^ 1u;
        const uint32_t found = S.qn[cur], qn = min(found, kSbMax);
        // the frontier's degrees: four consecutive vertices per thread
        uint32_t b4[4], d4[4];
        uint64_t dsum = 0;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const uint32_t i = tid * 4 + k;
            b4[k] = d4[k] = 0;
            if (i < qn) {
                const uint32_t v = S.q[cur][i];
                b4[k] = D.rp[v];
                d4[k] = D.rp[v + 1] - b4[k];
            }
            dsum += d4[k];
        }
        uint64_t excl, cut;
        const uint64_t edges = sb_scan(dsum, excl, cut_edges, cut, S.wsum);
        if (!first && tid == 0) {  // the previous level's fold, beside the edge map
            LevelOut r;
            r.found = found;
            r.next_edges = edges + cut;
            dl_fold_one(D, B, r, true, tick_ms);
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {  // offsets and the edge map (used only while edges <= kSbEdges)
            const uint32_t i = tid * 4 + k;
            if (i < qn && edges <= kSbEdges) {
                S.off[i] = (uint32_t)excl;
                S.vb[i] = b4[k];
                for (uint32_t x = 0; x < d4[k]; ++x) S.map[excl + x] = (uint16_t)i;
            }
            excl += d4[k];
        }
        __syncthreads();
        if (!first) {
            if (!dl_decide_rest(D, B, true, mlp_lds)) return false;
            if (found > kSbMax || !sb_ok(B.st)) {
                // hand back: the queued part of the frontier into its slot (the
                // vertices beyond the queue's capacity set their bits when found)
                uint64_t* fbl = D.fb + (size_t)(B.st.level % 3) * D.nb;
                for (uint32_t i = tid; i < qn; i += blockDim.x) {
                    const uint32_t v = S.q[cur][i];
                    __hip_atomic_fetch_or(&fbl[v >> 6], 1ull << (v & 63), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                return true;
            }
        }
        cut_edges = 0;
        const uint32_t level = B.st.level, depth = level + 1;
        // heartbeat for the waiting workgroups' bounded wait
        if (tid == 0 && (level & 63) == 0) store_through(D.cmd, dl_cmd(kCmdNone, level, B.st.action));
        uint64_t* fbn = D.fb + (size_t)((level + 1) % 3) * D.nb;  // bits of vertices beyond the queue only
        if (tid == 0) S.qn[nxt] = 0;
        __syncthreads();
        // edge-parallel expansion: edge e0 + k * blockDim (consecutive edges in a wave), kSbBatch in flight
        const uint32_t E = (uint32_t)edges;
        for (uint32_t e0 = tid; e0 < E; e0 += blockDim.x * kSbBatch) {
            uint32_t u[kSbBatch];
#pragma unroll
            for (uint32_t k = 0; k < kSbBatch; ++k) {
                const uint32_t e = e0 + k * blockDim.x;
                u[k] = 0;
                if (e < E) {
                    const uint32_t i = S.map[e];
                    u[k] = D.ci[S.vb[i] + (e - S.off[i])];
                }
            }
            uint64_t old[kSbBatch];
#pragma unroll
            for (uint32_t k = 0; k < kSbBatch; ++k)
                old[k] = e0 + k * blockDim.x < E
                             ? __hip_atomic_fetch_or(&D.vis[u[k] >> 6], 1ull << (u[k] & 63), __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)
                             : ~0ull;
#pragma unroll
            for (uint32_t k = 0; k < kSbBatch; ++k) {
                const uint64_t bit = 1ull << (u[k] & 63);
                const bool fresh = !(old[k] & bit);
                if (fresh) {
                    // no workgroup reads dist during the launch: a plain store,
                    // written back by the end of the kernel
                    D.dist[u[k]] = depth;
                }
                // one queue reservation per wave
                const uint64_t m = __ballot(fresh);
                if (!m) continue;
                uint32_t base = 0;
                if (lane == (uint32_t)__ffsll((long long)m) - 1) base = atomicAdd(&S.qn[nxt], (uint32_t)__popcll(m));
                base = __shfl(base, __ffsll((long long)m) - 1, kWave);
                if (fresh) {
                    const uint32_t p = base + (uint32_t)__popcll(m & ((1ull << lane) - 1));
                    if (p < kSbMax) {
                        S.q[nxt][p] = u[k];
                    } else {  // beyond the queue: the bitmap holds it (this ends the single-workgroup levels)
                        cut_edges += D.rp[u[k] + 1] - D.rp[u[k]];
                        __hip_atomic_fetch_or(&fbn[u[k] >> 6], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            }
        }
        __syncthreads();  // the next queue is complete
        cur = nxt;
    }
}

// --------------------------------------------------------------- kernels
// Before either loop: dist = INF, vis and the frontier slots zero, the source seeded.
__global__ void k_dl_init(DlDev D, uint32_t src) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < D.n; i += stride) D.dist[i] = GGPU_INF;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < D.nb; i += stride) {
        const uint64_t w = (i == (src >> 6)) ? 1ull << (src & 63) : 0ull;
        D.vis[i] = w;
        D.fb[i] = w;
        D.fb[D.nb + i] = 0;
        D.fb[2 * (size_t)D.nb + i] = 0;
    }
}

__global__ void k_dl_seed(DlDev D, uint32_t src) { D.dist[src] = 0; }

// Persistent loop, workgroups other than 0 (thread 0): the next command for a
// workgroup whose last grid level was level1 - 1 (level1 = 0: none yet). The
// wait gives up after kBarSpinLimit polls without any change of the command
// word (workgroup 0 stores a heartbeat every 64 levels it runs alone), or when
// another workgroup has raised the abort flag.
__device__ uint64_t dl_wait_cmd(const DlDev& D, uint32_t level1) {
    uint64_t seen = ~0ull, spins = 0;
    for (uint32_t poll = 0;; ++poll) {
        const uint64_t c = load_coherent(D.cmd);
        const uint32_t kind = dl_cmd_kind(c);
        if (kind == kCmdDone || kind == kCmdAbort) return c;
        if (kind == kCmdGrid && dl_cmd_level1(c) > level1) return c;
        if (c != seen) {
            seen = c;
            spins = 0;
        } else if (++spins > kBarSpinLimit || ((poll & 15) == 0 && load_coherent(D.abort))) {
            __hip_atomic_store(D.abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return dl_cmd(kCmdAbort, 0, 0);
        }
        // while workgroup 0 runs levels alone, poll less often: 255 pollers of one
        // line load the memory channel workgroup 0's own accesses go through
        if (kind == kCmdNone) __builtin_amdgcn_s_sleep(32);
        else __builtin_amdgcn_s_sleep(4);
    }
}

// Persistent loop, workgroup 0 (every thread): wait until `target` group
// completions have reached the top counter. False when the wait gave up (then
// the abort flag is raised) or another workgroup raised it.
__device__ bool dl_wait_arrivals(const DlDev& D, uint32_t target) {
    __shared__ uint32_t s_ok;
    if (threadIdx.x == 0) {
        uint32_t ok = 1;
        uint64_t spins = 0;
        while (load_coherent(&D.bar[kBarTop]) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > kBarSpinLimit || ((spins & 15) == 0 && load_coherent(D.abort))) {
                __hip_atomic_store(D.abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
        }
        asm volatile("buffer_inv sc0" ::: "memory");  // this CU's L1
        s_ok = ok;
    }
    __syncthreads();
    const bool ok = s_ok != 0;
    __syncthreads();
    return ok;
}

// Workgroup 0 of the persistent loop: the only decider (dl_protocol.hpp).
__device__ __noinline__ void dl_decider(const DlDev& D, DlBlock& B, SbShared& S, float* mlp_lds) {
    const double tick_ms = D.ds->tick_ms;
    const uint32_t groups = dl_groups(gridDim.x, kBarGroups);
    if (threadIdx.x == 0) {
        B.st = D.ds->st;
        B.pv = D.ds->pv;
        B.t_prev = 0;
        B.prev_action = 0;
    }
    __syncthreads();
    if (B.pv.mlp)
        for (uint32_t i = threadIdx.x; i < GGPU_MLP_PARAMS; i += blockDim.x) mlp_lds[i] = B.pv.mlp[i];
    __syncthreads();
    bool more = dl_decide(D, B, false, nullptr, 0, true, tick_ms, mlp_lds);
    for (uint32_t g = 0; more; ++g) {
        if (sb_ok(B.st)) {
            more = sb_levels(D, B, tick_ms, mlp_lds, S);
            // every store and atomic of those levels has completed before the
            // command below lets the grid read the frontier and visited words
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (!more) break;
        }
        const uint32_t level = B.st.level, action = B.st.action;
        if (threadIdx.x == 0) store_through(D.cmd, dl_cmd(kCmdGrid, level, action));
        unsigned long long* slot = D.part + (size_t)dl_part_slot(g) * gridDim.x * kDlParts;
        dl_level(D, level, action, slot);
        grid_arrive(D.bar);
        if (!dl_wait_arrivals(D, (g + 1) * groups)) {
            if (threadIdx.x == 0) store_through(D.cmd, dl_cmd(kCmdAbort, level, action));
            return;
        }
        more = dl_decide(D, B, true, slot, gridDim.x, true, tick_ms, mlp_lds);
    }
    if (threadIdx.x == 0) {
        store_through(D.cmd, dl_cmd(kCmdDone, B.st.level, 0));
        D.ds->st = B.st;
        D.ds->levels = B.st.level;
        D.ds->done = 1;
    }
}

__global__ void __launch_bounds__(kDlThreads) k_dl_persistent(DlDev D) {
    __shared__ float s_mlp[GGPU_MLP_PARAMS];
    __shared__ DlBlock s_b;
    __shared__ SbShared s_sb;
    __shared__ uint64_t s_cmd;
    if (blockIdx.x == 0) {
        dl_decider(D, s_b, s_sb, s_mlp);
        return;
    }
    // the others run exactly the grid levels workgroup 0 publishes, in order
    uint32_t level1 = 0;
    for (uint32_t g = 0;; ++g) {
        if (threadIdx.x == 0) {
            s_cmd = dl_wait_cmd(D, level1);
            asm volatile("buffer_inv sc0" ::: "memory");  // this CU's L1
        }
        __syncthreads();
        const uint64_t c = s_cmd;
        __syncthreads();  // s_cmd is rewritten by the next wait
        if (dl_cmd_kind(c) != kCmdGrid) return;
        level1 = dl_cmd_level1(c);
        dl_level(D, level1 - 1, dl_cmd_action(c), D.part + (size_t)dl_part_slot(g) * gridDim.x * kDlParts);
        grid_arrive(D.bar);
    }
}

// Graph mode: one block of one wave advances the state...
__global__ void __launch_bounds__(kWave) k_dl_decide(DlDev D, uint32_t level_blocks) {
    __shared__ float s_mlp[GGPU_MLP_PARAMS];
    __shared__ DlBlock s_b;
    DlState* ds = D.ds;
    if (ds->done) return;
    if (threadIdx.x == 0) {
        s_b.st = ds->st;
        s_b.pv = ds->pv;
        s_b.prev_action = ds->prev_action;
        s_b.t_prev = ds->t_prev;
    }
    __syncthreads();
    if (s_b.pv.mlp)
        for (uint32_t i = threadIdx.x; i < GGPU_MLP_PARAMS; i += blockDim.x) s_mlp[i] = s_b.pv.mlp[i];
    __syncthreads();
    const bool started = ds->started != 0;
    const bool more = dl_decide(D, s_b, started, D.part + (size_t)((s_b.st.level + 1) & 1) * level_blocks * kDlParts,
                                level_blocks, true, ds->tick_ms, s_mlp);
    if (threadIdx.x == 0) {
        ds->st = s_b.st;
        ds->prev_action = s_b.prev_action;
        ds->t_prev = s_b.t_prev;
        ds->started = 1;
        if (!more) {
            ds->done = 1;
            ds->levels = s_b.st.level;
        }
    }
}

// ...and the level kernel runs the chosen action on the whole grid.
__global__ void __launch_bounds__(kDlLevelThreads) k_dl_level(DlDev D) {
    const DlState* ds = D.ds;
    if (ds->done) return;
    const uint32_t level = ds->st.level, action = ds->st.action;
    dl_level(D, level, action, D.part + (size_t)((level + 1) & 1) * gridDim.x * kDlParts);
}

// ------------------------------------------------------------------- host
struct DlResources {
    DBuf<uint64_t> vis, fb;
    DBuf<unsigned long long> part;
    DBuf<DlState> ds;
    DBuf<ggpu_level> trace;
    DBuf<uint8_t> actions;
    DBuf<uint32_t> ctl;
    uint32_t persist_grid = 0, level_grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t graph_k = 0;
    DlDev graph_dev{};  // the arguments the graph was captured with
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~DlResources() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

static void dl_free(void* p) { delete static_cast<DlResources*>(p); }

}  // namespace ggpu

namespace ggpu {

void BfsEngine::device_loop(int variant, uint32_t src, const ggpu_bfs_opts& o, std::vector<ggpu_level>* trace) {
    if (variant == GGPU_BFS_HYBRID) throw ArgError("hybrid runs with the host loop");
    if (o.loop == GGPU_LOOP_GRAPH && (o.graph_levels == 0 || o.graph_levels > 4096))
        throw ArgError("graph_levels must be in 1..4096");
    hipStream_t s = ctx->stream;
    if (!dl_state) {
        dl_state = new DlResources();
        dl_free_fn = dl_free;
    }
    DlResources& R = *static_cast<DlResources*>(dl_state);
    if (!R.ev[0]) {
        GGPU_HIP_TRY(hipEventCreate(&R.ev[0]));
        GGPU_HIP_TRY(hipEventCreate(&R.ev[1]));
    }
    if (!R.persist_grid) {
        int per_cu = 0;
        GGPU_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_dl_persistent, kDlThreads, 0));
        if (per_cu < 1) throw ArgError("persistent BFS kernel does not fit on a compute unit", GGPU_EDEVICE);
        R.persist_grid = (uint32_t)ctx->cus;
        R.level_grid = (uint32_t)ctx->cus * 4;
    }
    const uint32_t nb = (n + 63) / 64;
    R.vis.alloc(nb + 1);
    R.fb.alloc(3ull * nb + 1);
    R.part.alloc(2ull * std::max(R.persist_grid, R.level_grid) * kDlParts);
    R.ds.alloc(1);
    R.ctl.alloc(kBarWords + 16);
    const size_t tcap = trace ? (size_t)n + 2 : 0;
    if (trace) R.trace.alloc(tcap);
    const float* mlp = (o.policy == GGPU_POLICY_LEARNED && ctx->policy_set[GGPU_POLICY_SINGLE])
                           ? ctx->policy_dev[GGPU_POLICY_SINGLE] : nullptr;
    if (o.policy == GGPU_POLICY_SCRIPTED) {
        R.actions.alloc(o.n_actions);
        GGPU_HIP_TRY(hipMemcpyAsync(R.actions.p, o.actions, o.n_actions, hipMemcpyHostToDevice, s));
    }
    DlState h{};
    h.st = loop_start(src, variant, o);
    h.pv = policy_view(variant, o, mlp);
    h.pv.actions = o.policy == GGPU_POLICY_SCRIPTED ? R.actions.p : nullptr;
    int clock_khz = 0;
    GGPU_HIP_TRY(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, ctx->device));
    h.tick_ms = clock_khz > 0 ? 1.0 / clock_khz : 0.0;
    GGPU_HIP_TRY(hipMemcpyAsync(R.ds.p, &h, sizeof h, hipMemcpyHostToDevice, s));
    GGPU_HIP_TRY(hipMemsetAsync(R.ctl.p, 0, (kBarWords + 16) * sizeof(uint32_t), s));

    DlDev D{};
    D.rp = rp.p;
    D.ci = ci.p;
    D.dist = dist.p;
    D.vis = R.vis.p;
    D.fb = R.fb.p;
    D.part = R.part.p;
    D.heavy = heavy.p;
    D.ds = R.ds.p;
    D.trace = trace ? R.trace.p : nullptr;
    D.bar = R.ctl.p;
    D.abort = R.ctl.p + kBarWords;
    D.cmd = (unsigned long long*)(R.ctl.p + kBarWords + 2);  // 8-byte aligned: kBarWords is even
    D.n_heavy = n_heavy;
    D.heavy_min = kHeavyMin;
    D.n = n;
    D.nb = nb;
    D.trace_cap = (uint32_t)std::min<size_t>(tcap, 0xFFFFFFFFu);
    D.m = m;

    GGPU_HIP_TRY(hipEventRecord(R.ev[0], s));
    k_dl_init<<<blocks_for(std::max(n, nb), 256, grid_cap), 256, 0, s>>>(D, src);
    k_dl_seed<<<1, 1, 0, s>>>(D, src);
    GGPU_HIP_TRY(hipGetLastError());
    DlState out{};
    if (o.loop == GGPU_LOOP_PERSISTENT) {
        // A plain launch: the kernel synchronises itself (dl_protocol.hpp), so
        // it needs no grid-sync resource from the runtime (hipLaunchCooperativeKernel
        // has failed here with "cannot be performed in the present state" while
        // another process held the device). Residency: one workgroup per CU (its
        // LDS footprint allows no second), grid = CUs; a workgroup that cannot
        // start ends the launch through the bounded waits (GGPU_ETIMEOUT).
        void* args[] = {&D};
        GGPU_HIP_TRY(hipLaunchKernel((const void*)k_dl_persistent, dim3(R.persist_grid), dim3(kDlThreads), args, 0, s));
        GGPU_HIP_TRY(hipEventRecord(R.ev[1], s));
        uint32_t abort = 0;
        GGPU_HIP_TRY(hipMemcpyAsync(&abort, D.abort, 4, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipMemcpyAsync(&out, R.ds.p, sizeof out, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
        if (abort) {
            // where the grid stood: arrivals per group, group completions, the
            // last command word (kind, level)
            std::vector<uint32_t> c(kBarWords + 16);
            GGPU_HIP_TRY(hipMemcpy(c.data(), R.ctl.p, c.size() * 4, hipMemcpyDeviceToHost));
            uint64_t cmd = 0;
            std::memcpy(&cmd, &c[kBarWords + 2], sizeof cmd);
            std::string msg = "persistent BFS: a wait timed out (grid " + std::to_string(R.persist_grid) +
                              ", group arrivals";
            for (uint32_t g = 0; g < kBarGroups; ++g) msg += " " + std::to_string(c[g * kBarStride]);
            msg += ", group completions " + std::to_string(c[kBarTop]) + ", command kind " +
                   std::to_string(dl_cmd_kind(cmd)) + " level " + std::to_string((int64_t)dl_cmd_level1(cmd) - 1) + ")";
            throw ArgError(msg, GGPU_ETIMEOUT);
        }
    } else {
        // a graph of K levels; the captured kernels read everything else from device memory
        if (!R.exec || R.graph_k != o.graph_levels || std::memcmp(&R.graph_dev, &D, sizeof D) != 0) {
            if (R.exec) GGPU_HIP_TRY(hipGraphExecDestroy(R.exec));
            if (R.graph) GGPU_HIP_TRY(hipGraphDestroy(R.graph));
            R.exec = nullptr;
            R.graph = nullptr;
            GGPU_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (uint32_t l = 0; l < o.graph_levels; ++l) {
                k_dl_decide<<<1, kWave, 0, s>>>(D, R.level_grid);
                k_dl_level<<<R.level_grid, kDlLevelThreads, 0, s>>>(D);
            }
            GGPU_HIP_TRY(hipStreamEndCapture(s, &R.graph));
            GGPU_HIP_TRY(hipGraphInstantiate(&R.exec, R.graph, nullptr, nullptr, 0));
            R.graph_k = o.graph_levels;
            R.graph_dev = D;
        }
        uint32_t* done = (uint32_t*)ctx_pinned(ctx, 4096);
        for (uint64_t launches = 0;; ++launches) {
            if (launches > (uint64_t)n / o.graph_levels + 2) throw ArgError("graph loop did not finish", GGPU_ESTATE);
            GGPU_HIP_TRY(hipGraphLaunch(R.exec, s));
            GGPU_HIP_TRY(hipMemcpyAsync(done, &R.ds.p->done, 4, hipMemcpyDeviceToHost, s));
            GGPU_HIP_TRY(hipStreamSynchronize(s));
            if (*done) break;
        }
        GGPU_HIP_TRY(hipEventRecord(R.ev[1], s));
        GGPU_HIP_TRY(hipMemcpyAsync(&out, R.ds.p, sizeof out, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
    }
    float ms = 0;
    GGPU_HIP_TRY(hipEventElapsedTime(&ms, R.ev[0], R.ev[1]));
    last_gpu_ms = ms;
    last_levels = out.levels;
    if (trace) {
        trace->resize(out.levels);
        if (out.levels)
            GGPU_HIP_TRY(hipMemcpy(trace->data(), R.trace.p, out.levels * sizeof(ggpu_level), hipMemcpyDeviceToHost));
    }
}


// ------------------------------------------------- BFS without allocation
// ggpu_bfs_run_noalloc: the persistent gpu_do loop over a CSR the caller holds
// in device memory, with every working buffer in one fixed device array. The
// array's layout, in 8-byte words (each region starts on a 256-byte boundary):
//   vis   kNaNb + 32            visited bitmap (nb + 1 words used)
//   fb    3 kNaNb + 32          frontier bitmaps, three slots of nb words
//   part  2 kNaMaxGrid kDlParts per-workgroup level counters, two slots
//   ds    kNaDsWords            DlState
//   ctl   kNaCtlWords           arrival counters, abort flag, command word
// No heavy-vertex list: every vertex takes the light paths (heavy_min = ~0u).
// The counters that decide each level (found, next_edges) are exact on either
// path, so the decisions and the distances equal ggpu_bfs_run's.
constexpr uint32_t kNaMaxN = GGPU_NOALLOC_MAX_N;
constexpr uint32_t kNaMaxGrid = 256;
constexpr size_t kNaNb = kNaMaxN / 64;
constexpr size_t kNaAlign = 32;  // words
constexpr size_t na_round(size_t w) { return (w + kNaAlign - 1) / kNaAlign * kNaAlign; }
constexpr size_t kNaVis = 0;
constexpr size_t kNaFb = kNaVis + na_round(kNaNb + 1);
constexpr size_t kNaPart = kNaFb + na_round(3 * kNaNb + 1);
constexpr size_t kNaDs = kNaPart + na_round(2ull * kNaMaxGrid * kDlParts);
constexpr size_t kNaDsWords = na_round((sizeof(DlState) + 7) / 8);
constexpr size_t kNaCtl = kNaDs + kNaDsWords;
constexpr uint32_t kNaCtlU32 = kBarWords + 16;
constexpr size_t kNaCtlWords = na_round(kNaCtlU32 / 2);
constexpr size_t kNaWords = kNaCtl + kNaCtlWords;
static_assert(kNaWords * 8 == GGPU_NOALLOC_SCRATCH_BYTES, "GGPU_NOALLOC_SCRATCH_BYTES must match the layout");
static_assert(kBarWords % 2 == 0, "the command word must be 8-byte aligned");

// The array itself. Host code never uses this name as an address: its device
// address, per GPU, comes from hipGetSymbolAddress (noalloc_setup).
static __device__ __align__(256) uint64_t scratch[kNaWords];

constexpr int kNaMaxDevices = 64;
static uint64_t* g_na_ptr[kNaMaxDevices];   // scratch's device address on each GPU
static std::string g_na_err[kNaMaxDevices]; // why it is missing, if it is
static std::mutex g_na_mu;

// Called by ggpu_ctx_create with `device` current: looks up the array's address
// on that GPU once per process.
void noalloc_setup(int device) {
    if (device < 0 || device >= kNaMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (g_na_ptr[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        g_na_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
    }
    g_na_ptr[device] = static_cast<uint64_t*>(ptr);
    g_na_err[device].clear();
}

// Loop state and control words of one run: the start state from the source's
// degree (read here, from the caller's row offsets), the counters zeroed.
__global__ void k_na_setup(DlState* ds, PolicyView pv, const uint32_t* rp, uint32_t src, uint64_t m,
                           uint32_t small_max, uint32_t large_min, uint32_t* ctl) {
    for (uint32_t i = threadIdx.x; i < kNaCtlU32; i += blockDim.x) ctl[i] = 0;
    if (threadIdx.x == 0) {
        DlState h{};
        h.st = loop_start_state(rp[src + 1] - rp[src], m, GGPU_BFS_DO, small_max, large_min);
        h.pv = pv;
        *ds = h;
    }
}

int bfs_run_noalloc(ggpu_ctx* ctx, uint64_t n, uint64_t m, const uint32_t* rp, const uint32_t* ci,
                    uint32_t source, const ggpu_bfs_opts* opt, uint32_t* dist, uint32_t* levels) {
    ggpu_bfs_opts o;
    ggpu_bfs_opts_default(&o);
    if (opt) o = *opt;
    return guarded(ctx, [&] {
        if (n == 0 || n > kNaMaxN)
            throw ArgError("ggpu_bfs_run_noalloc: n = " + std::to_string(n) + " is outside 1.." +
                           std::to_string(kNaMaxN) + " (the fixed working array's capacity)");
        if (m >= (1ull << 32)) throw ArgError("ggpu_bfs_run_noalloc: the graph has 2^32 or more arcs");
        if (source >= n) throw ArgError("source out of range");
        if (!rp || !dist || (m && !ci)) throw ArgError("ggpu_bfs_run_noalloc: null device pointer");
        if (o.policy == GGPU_POLICY_SCRIPTED)
            throw ArgError("ggpu_bfs_run_noalloc: the scripted policy needs its actions in device memory");
        if (o.policy > GGPU_POLICY_SCRIPTED) throw ArgError("unknown policy");
        if (ctx->device < 0 || ctx->device >= kNaMaxDevices) throw ArgError("device index out of range");
        uint64_t* w = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_na_mu);
            w = g_na_ptr[ctx->device];
            if (!w) throw ArgError("ggpu_bfs_run_noalloc: no working array on this GPU: " + g_na_err[ctx->device],
                                   GGPU_EDEVICE);
        }
        if (!ctx->pinned || ctx->pinned_bytes < 16) throw ArgError("context has no staging memory", GGPU_ESTATE);
        GGPU_HIP_TRY(hipSetDevice(ctx->device));
        int per_cu = 0;
        GGPU_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_dl_persistent, kDlThreads, 0));
        if (per_cu < 1) throw ArgError("persistent BFS kernel does not fit on a compute unit", GGPU_EDEVICE);
        const uint32_t grid = std::min<uint32_t>((uint32_t)ctx->cus, kNaMaxGrid);
        const uint32_t nn = (uint32_t)n, nb = (nn + 63) / 64;
        hipStream_t s = ctx->stream;

        DlDev D{};
        D.rp = rp;
        D.ci = ci;
        D.dist = dist;
        D.vis = w + kNaVis;
        D.fb = w + kNaFb;
        D.part = reinterpret_cast<unsigned long long*>(w + kNaPart);
        D.heavy = nullptr;
        D.n_heavy = 0;
        D.heavy_min = ~0u;
        D.ds = reinterpret_cast<DlState*>(w + kNaDs);
        D.trace = nullptr;
        D.trace_cap = 0;
        uint32_t* ctl = reinterpret_cast<uint32_t*>(w + kNaCtl);
        D.bar = ctl;
        D.abort = ctl + kBarWords;
        D.cmd = reinterpret_cast<unsigned long long*>(ctl + kBarWords + 2);
        D.n = nn;
        D.nb = nb;
        D.m = m;

        const float* mlp = (o.policy == GGPU_POLICY_LEARNED && ctx->policy_set[GGPU_POLICY_SINGLE])
                               ? ctx->policy_dev[GGPU_POLICY_SINGLE] : nullptr;
        PolicyView pv = make_policy_view(GGPU_BFS_DO, o, mlp, n, m);
        pv.actions = nullptr;
        pv.n_actions = 0;
        k_na_setup<<<1, 256, 0, s>>>(D.ds, pv, rp, source, m, o.small_max, o.large_min, ctl);
        k_dl_init<<<blocks_for(std::max(nn, nb), 256, 4096), 256, 0, s>>>(D, source);
        k_dl_seed<<<1, 1, 0, s>>>(D, source);
        GGPU_HIP_TRY(hipGetLastError());
        void* args[] = {&D};
        GGPU_HIP_TRY(hipLaunchKernel((const void*)k_dl_persistent, dim3(grid), dim3(kDlThreads), args, 0, s));
        uint32_t* out = static_cast<uint32_t*>(ctx->pinned);
        GGPU_HIP_TRY(hipMemcpyAsync(&out[0], D.abort, 4, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipMemcpyAsync(&out[1], &D.ds->levels, 4, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
        if (out[0]) throw ArgError("ggpu_bfs_run_noalloc: a wait of the persistent loop timed out", GGPU_ETIMEOUT);
        if (levels) *levels = out[1];
    });
}

}  // namespace ggpu

extern "C" int ggpu_bfs_run_noalloc(ggpu_ctx* ctx, uint64_t n, uint64_t m, const uint32_t* d_row_ptr,
                                    const uint32_t* d_col_idx, uint32_t source, const ggpu_bfs_opts* opt,
                                    uint32_t* d_dist, uint32_t* levels) {
    if (!ctx) return GGPU_EINVAL;
    return ggpu::bfs_run_noalloc(ctx, n, m, d_row_ptr, d_col_idx, source, opt, d_dist, levels);
}
