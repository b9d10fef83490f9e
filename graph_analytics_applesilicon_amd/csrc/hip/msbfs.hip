// Batched multi-source BFS (no reference counterpart): k <= 256 sources
// traverse together in ONE persistent launch (a workgroup per CU), one grid barrier
// per level (a renumbered run below adds an un-permute kernel, and a relaunch
// per window of levels on deep graphs).
//
// State, W = ceil(k/64) u64 words per vertex, bit s%64 of word s/64 = source s:
//   visit[v][w]       sources that have reached v
//   F_l[v][w]         sources for which v is at depth l       (C level slots)
//   fb_l[v/64]        bit v%64: F_l[v] != 0 (the level-l frontier)
// Level l reads F_l / fb_l and writes F_{l+1} / fb_{l+1} / visit:
//   bottom-up: an unfinished vertex ORs the F_l words of its frontier
//     neighbours until every bit it still needs is present (testing fb_l
//     first, unless most arcs lead into the frontier);
//   top-down: a frontier vertex ORs F_l[v] (& ~visit[u] once most pair-edges
//     are reached) into next[u] of its neighbours; after a barrier each touched
//     vertex keeps next & ~visit as its new bits, so each (source, vertex)
//     pair is discovered once.
// Vertices of degree >= heavy_min are scanned / pushed as chunk-edge items
// by whole waves (both directions), the rest lane per vertex. A top-down
// level with a small frontier and many arcs is routed instead of pushed
// (batches of up to 128 sources): its records are binned by target range and
// each bin's pushes are applied with LDS atomics (ms_route_*, below).
// Depths are not written during the traversal. A launch of k_msbfs ends when
// its C level slots are used up (saving the loop state; the next launch
// resumes there) or when the traversal is over; then k_ms_finish turns the
// window's level slots into depths through bit planes (plane p = OR of F over
// levels whose offset has bit p), written source-major with coalesced stores.
// That pass also clears what it read, so visit / next / F / fb are all zero
// between runs. As a kernel of its own it issues every load of a vertex at
// once (inside the persistent kernel the pass cost its level loop registers,
// and its per-level loads ran one after another).
// A relabeled run (batch_order DEGREE, or AUTO on a graph whose maximum degree is
// at least 8x the mean) traverses the graph renumbered by descending degree (the
// frontier words most arcs gather then share a few MB of lines, which stay in
// each XCD's L2): the depth pass writes a 32-byte record per vertex (four
// planes, a 4-bit code per source) in that numbering, and a separate kernel, k_ms_unpermute,
// reads the record of perm[u] for every caller id u and writes u's depths,
// after each k_ms_finish.
//
// Cross-block data written inside a level goes through device-scope atomics
// or write-through stores only (grid_sync.hpp), and every decision input
// (per-block counters) is reduced identically by every block, so all blocks
// take the same branch at every level without a second barrier.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "bfs_engine.hpp"
#include "ggpu_hip_testing.h"
#include "grid_sync.hpp"

namespace ggpu {

constexpr uint32_t kMsThreads = 1024;
// Degree from which a vertex is handled as wave items, and edges per item: in the
// caller's numbering a wave's 64 consecutive vertices mix degrees, so vertices of
// degree >= 32 go to wave items; in the degree numbering they are alike, and
// lane-per-vertex scans pay up to degree 63 (64-source rmat22 batch: 1.72 ms at
// 32 / 1024, 1.54 ms at 64 / 2048; caller numbering 2.55 ms at 32 / 1024, 2.84 ms
// at 64 / 2048).
constexpr uint32_t kMsHeavyGiven = 32, kMsChunkGiven = 1024;
constexpr uint32_t kMsHeavyDegree = 64, kMsChunkDegree = 2048;
constexpr uint32_t kMsSlots = 15;       // level slots: a window retires <= 14 levels (four depth planes)
constexpr uint32_t kMsPlanes = 4;
constexpr uint32_t kMsRec = 4;          // u64 words per (vertex, source word) depth record: the planes
// Codes of a depth record (the four plane bits of a source): 0..13 depth - base,
// kCodeSkip: not reached in this window (an earlier window wrote it), kCodeInf: never reached.
constexpr uint32_t kCodeSkip = 14, kCodeInf = 15;
// Loop state of a relabeled run between launches (u64 words; zero before the first
// launch of a run: zeroed at allocation and by the run's last launch):
enum : uint32_t {
    kStResume = 0,   // 1: resume at kStLevel, whose action kStAct was decided by the previous launch
    kStLevel,
    kStAct,
    kStBottomUp,     // heuristic policy state
    kStPrevSize,
    kStCounts,       // 3 words: levels run per action
    kStCumPairs = kStCounts + 3,
    kStBase,         // first level of the window whose records the un-permute reads
    kStLoopTicks,    // device clock spent in the level loop by earlier launches
    kStSetupTicks,
    kStCnt,          // the window k_ms_finish retires: levels [kStBase, kStBase + kStCnt)
    kStFinal,        // 1: the traversal is over (k_ms_finish also writes INF and clears visit)
    kMsState
};
static_assert(kMsSlots - 1 <= kCodeSkip, "depth offsets within a window stay below the two special codes");
constexpr uint32_t kMsTraceCap = 1u << 16;
constexpr uint32_t kPartWords = 4;
// Bottom-up LDS cache: the frontier words of the highest-degree vertices
// (kMsHotWords / W of them; 16384 at W = 1 take 41% of rmat22's arcs). In
// the bottom-up copy of col_idx an arc to such a vertex holds kMsHotBit | slot.
constexpr uint32_t kMsHotWords = 16384;  // 128 KiB of LDS
constexpr uint32_t kMsHotBit = 1u << 31;
constexpr uint32_t kMsZeroWords = 64;
// The kernel's big LDS buffer: the bottom-up hot words, or the routed
// top-down's queue prefix and bin counters, then its next words of one bin.
__shared__ uint64_t g_ms_lds[kMsHotWords];

// Routed top-down (ms_route_*): a top-down level whose frontier has at most
// kRouteQ vertices and at least kRouteMinEdges arcs. Records are (vertex - bin
// start) << kRouteQBits | queue index; a bin is a word-aligned vertex range of
// at most kRouteBinV / W vertices (its next words fill the LDS buffer).
constexpr uint32_t kRouteQ = 4096;
constexpr uint32_t kRouteQBits = 12;
constexpr uint32_t kRouteBinV = kMsHotWords;
constexpr uint32_t kRouteBins = 2048;
constexpr uint32_t kRouteArcs = 8;                        // consecutive arcs a thread bins per round
constexpr uint32_t kRouteChunk = kRouteArcs * kMsThreads;  // arcs a workgroup bins per round
constexpr uint64_t kRouteMinEdges = 65536;
static_assert((kRouteBinV << kRouteQBits) <= (1ull << 32) && kRouteQ <= (1u << kRouteQBits), "record fields");
// Abort code of a routed level whose queue, arc total or bin reservations do not
// match the level's counters (the phases then skip their writes; every
// workgroup leaves at the next barrier and the host reports it).
constexpr uint32_t kAbortRoute = 2;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

struct MsDev {
    const uint32_t* rp;
    const uint32_t* ci;
    const uint32_t* cib;           // bottom-up col_idx: kMsHotBit | slot for arcs to hot vertices
    const uint32_t* hot;           // n_hot hot vertex ids, by slot
    uint64_t* visit;               // n * W
    uint64_t* next;                // n * W, top-down push target (zero between levels)
    uint64_t* F;                   // C * n * W
    uint64_t* fb;                  // C * nb
    const uint64_t* zero;          // kMsZeroWords zero words: the load of a lane that takes nothing
    const uint64_t* iso;           // nb words, bit v % 64: v has no arcs or v >= n (never scanned)
    uint64_t* done;                // nb words, bit v % 64: every source has reached v (zero between runs)
    unsigned long long* part;      // 2 * G * kPartWords
    uint32_t* dist;                // k * n, source-major
    const uint32_t* sources;       // k
    const uint4* heavy;            // {v, first edge, end edge, degree}
    const uint32_t* perm;          // relabeled runs: perm[caller id] = traversal id (else null)
    uint64_t* rec;                 // relabeled runs: n * W depth records (four planes of codes)
    unsigned long long* state;     // kMsState words: the loop state between the launches of one run
    uint32_t* bar;                 // kBarWords
    uint32_t* abort;
    uint32_t* out;                 // {levels, done, actions[3], 0, setup, loop, finish clock ticks}
    ggpu_level* trace;             // trace_cap entries
    float* feat;                   // trace_cap x 5 policy features (as the policy sees them)
    const uint8_t* actions;        // scripted policy
    const float* mlp;              // learned policy parameters or null
    double alpha, beta;
    double tick_ms;                // wall_clock64() period in ms (per-level times in the trace)
    uint64_t m;
    uint64_t valid_last;           // valid bits of the last source word
    uint32_t n, k, nb, n_heavy, C, trace_cap, n_actions, policy;
    uint32_t n_hot;                // <= kMsHotWords / W
    uint32_t heavy_min;            // degree from which a vertex is in the heavy item list
    // routed top-down (n_bins == 0: not available)
    uint32_t* rq;                  // kRouteQ frontier vertices
    uint32_t* rqd;                 // their degrees
    uint64_t* rqf;                 // their frontier words, kRouteQ * W
    uint32_t* rctl;                // [0] queue reservations, [1 + b] records of bin b (zero between levels)
    const uint32_t* blo;           // n_bins + 1 bin starts (multiples of 64; the last is nb * 64)
    const uint32_t* bbase;         // n_bins record-area starts (area b: the in-degree sum of bin b)
    const uint16_t* wbin;          // nb: the bin of each 64-vertex word
    uint32_t* recs;                // records
    uint32_t n_bins;
    uint64_t route_min;            // frontier arcs from which a small-frontier top-down level is routed
};

struct MsAcc {
    unsigned long long size = 0, edges = 0, pairs = 0, examined = 0;
};

template <int W>
__device__ __forceinline__ uint64_t valid_word(const MsDev& D, int w) {
    return w == W - 1 ? D.valid_last : ~0ull;
}

__device__ __forceinline__ uint32_t gwave() { return (blockIdx.x * blockDim.x + threadIdx.x) / kWave; }
__device__ __forceinline__ uint32_t nwaves() { return gridDim.x * blockDim.x / kWave; }

// The frontier / mark word a lane owns in the round starting at r0: consecutive
// waves take consecutive words and the lanes of a wave words nw apart, so a
// run of busy words (the degree numbering packs the hubs, and the low-degree
// tail, into a few) spreads over many waves. Words past nb read as empty.
__device__ __forceinline__ uint32_t ms_word_of(uint64_t r0, uint32_t lane, uint32_t nw) {
    return (uint32_t)(r0 + (uint64_t)lane * nw + gwave());
}

// Sums the kPartWords counters over the workgroup together (one LDS exchange
// and one pair of barriers for all four), into every thread.
__device__ __forceinline__ void ms_block_sum4(unsigned long long (&v)[kPartWords]) {
    __shared__ unsigned long long s_red4[kPartWords][kMsThreads / kWave];
    const uint32_t w = threadIdx.x / kWave, nwv = blockDim.x / kWave;
#pragma unroll
    for (uint32_t j = 0; j < kPartWords; ++j) v[j] = wave_sum(v[j]);
    __syncthreads();  // the previous call's readers are done with s_red4
    if (lane_id() == 0)
#pragma unroll
        for (uint32_t j = 0; j < kPartWords; ++j) s_red4[j][w] = v[j];
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kPartWords; ++j) {
        unsigned long long t = 0;
        for (uint32_t i = 0; i < nwv; ++i) t += s_red4[j][i];
        v[j] = t;
    }
}

// Per-block counters of the level being produced, one write-through store each.
__device__ void ms_flush(const MsDev& D, uint32_t slot, const MsAcc& a) {
    unsigned long long t[kPartWords] = {a.size, a.edges, a.pairs, a.examined};
    ms_block_sum4(t);
    if (threadIdx.x == 0) {
        unsigned long long* p = D.part + ((size_t)slot * gridDim.x + blockIdx.x) * kPartWords;
#pragma unroll
        for (uint32_t j = 0; j < kPartWords; ++j) store_through(&p[j], t[j]);
    }
}

// Totals of one counter slot, in every thread.
__device__ MsAcc ms_totals(const MsDev& D, uint32_t slot) {
    unsigned long long v[kPartWords] = {0, 0, 0, 0};
    for (uint32_t i = threadIdx.x; i < gridDim.x; i += blockDim.x) {
        const unsigned long long* p = D.part + ((size_t)slot * gridDim.x + i) * kPartWords;
#pragma unroll
        for (uint32_t j = 0; j < kPartWords; ++j) v[j] += p[j];
    }
    ms_block_sum4(v);
    MsAcc a;
    a.size = v[0];
    a.edges = v[1];
    a.pairs = v[2];
    a.examined = v[3];
    return a;
}

// Marks bits x of vertex u as reached at the level being produced; only the
// bits that atomicOr on visit reports as new count, so concurrent
// discoverers of the same pair count it once.
template <int W>
__device__ __forceinline__ void ms_discover(const MsDev& D, uint64_t* Fn, uint64_t* fbn, uint32_t u,
                                            const uint64_t (&x)[W], MsAcc& a) {
    uint32_t pc = 0;
    bool full = true;  // every word of u is now complete (a word this call does not touch: unknown)
#pragma unroll
    for (int w = 0; w < W; ++w) {
        if (!x[w]) {
            full = false;
            continue;
        }
        const uint64_t old = atomicOr((unsigned long long*)&D.visit[(size_t)u * W + w], (unsigned long long)x[w]);
        const uint64_t fresh = x[w] & ~old;
        full &= (old | x[w]) == valid_word<W>(D, w);
        if (fresh) {
            atomicOr((unsigned long long*)&Fn[(size_t)u * W + w], (unsigned long long)fresh);
            pc += (uint32_t)__popcll(fresh);
        }
    }
    if (pc) {
        const uint64_t dg = D.rp[u + 1] - D.rp[u];
        a.pairs += dg * pc;
        const uint64_t bit = 1ull << (u & 63);
        if (full) atomicOr((unsigned long long*)&D.done[u >> 6], (unsigned long long)bit);
        const uint64_t old = atomicOr((unsigned long long*)&fbn[u >> 6], (unsigned long long)bit);
        if (!(old & bit)) {
            a.size += 1;
            a.edges += dg;
        }
    }
}

template <int W>
__device__ void ms_seed(const MsDev& D, MsAcc& a) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < D.k) {
        uint64_t x[W];
#pragma unroll
        for (int w = 0; w < W; ++w) x[w] = (uint32_t)w == t / 64 ? 1ull << (t % 64) : 0ull;
        ms_discover<W>(D, D.F, D.fb, D.sources[t], x, a);
    }
}

// ---------------------------------------------------------------- bottom-up
template <int W>
// direct: gather the frontier words of every neighbour without testing the
// frontier bitmap first (a word is zero outside the frontier); pays when most
// arcs lead into the frontier, as the test then filters little and adds a
// dependent gather.
__device__ void ms_bottom_up(const MsDev& D, uint32_t level, bool use_hot, bool direct, MsAcc& a) {
    const uint32_t cs = level % D.C, ns = (level + 1) % D.C;
    const uint64_t* F = D.F + (size_t)cs * D.n * W;
    const uint64_t* fb = D.fb + (size_t)cs * D.nb;
    uint64_t* Fn = D.F + (size_t)ns * D.n * W;
    uint64_t* fbn = D.fb + (size_t)ns * D.nb;
    const uint32_t lane = lane_id();
    // this level's frontier words of the hot vertices (zero: not in the frontier);
    // a level with little left to scan does not pay for the copy and reads col_idx
    auto& s_hot = g_ms_lds;
    const uint32_t* cx = use_hot ? D.cib : D.ci;
    if (use_hot) {
        for (uint32_t i = threadIdx.x; i < D.n_hot * W; i += blockDim.x)
            s_hot[i] = F[(size_t)D.hot[i / W] * W + i % W];
        __syncthreads();
    }

    // light vertices: lane per vertex, 64 consecutive vertices per wave. The
    // wave's words are gwave + i * nwaves; lane i loads the skip word of the
    // i-th at once, and a word whose vertices are all isolated or complete is
    // not visited (in the degree numbering the isolated vertices fill the
    // last words)
    const uint64_t gw = gwave(), nwv = nwaves();
    for (uint64_t i0 = 0; gw + i0 * nwv < D.nb; i0 += kWave) {
        const uint64_t my_wd = gw + (i0 + lane) * nwv;
        const uint64_t my_skip = my_wd < D.nb ? (D.iso[my_wd] | D.done[my_wd]) : ~0ull;
        uint64_t words = __ballot(my_skip != ~0ull);
        while (words) {
        const int jw = __ffsll((long long)words) - 1;
        words &= words - 1;
        const uint32_t wd = (uint32_t)(gw + (i0 + (uint64_t)jw) * nwv);
        const uint64_t skip = __shfl(my_skip, jw, kWave);
        const uint32_t v = wd * 64 + lane;
        bool found = false, complete = false;
        if (!((skip >> lane) & 1ull)) {
            // v < n (past-the-end vertices are in iso); visit with the offsets
            const uint32_t b = D.rp[v], e = D.rp[v + 1];
            uint64_t vis[W];
#pragma unroll
            for (int w = 0; w < W; ++w) vis[w] = D.visit[(size_t)v * W + w];
            const uint32_t dg = e - b;
            if (dg > 0 && dg < D.heavy_min) {
                uint64_t need[W], acc[W];
                bool any = false;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    need[w] = ~vis[w] & valid_word<W>(D, w);
                    acc[w] = 0;
                    any |= need[w] != 0;
                }
                if (any) {
                    for (uint32_t x = b; x < e; x += 4) {
                        // branch-free, so that the four arcs' loads are in flight
                        // together: clamped arc loads, every test a select, and a
                        // lane that takes nothing gathers from the zero words
                        const uint32_t c = min(4u, e - x);
                        uint32_t u[4];
#pragma unroll
                        for (uint32_t i = 0; i < 4; ++i) u[i] = cx[min(x + i, e - 1)];
                        uint64_t fw[4];
#pragma unroll
                        for (uint32_t i = 0; i < 4; ++i) fw[i] = direct ? ~0ull : fb[(u[i] & ~kMsHotBit) >> 6];
                        uint64_t g[4][W], h[4][W];
#pragma unroll
                        for (uint32_t i = 0; i < 4; ++i) {
                            const uint32_t id = u[i] & ~kMsHotBit;
                            const bool hot = (u[i] & kMsHotBit) != 0;
                            const bool take = i < c && !hot && ((fw[i] >> (id & 63)) & 1ull);
                            const uint64_t* src = take ? F + (size_t)id * W : D.zero;
                            const uint32_t slot = hot ? id : 0u;
#pragma unroll
                            for (int w = 0; w < W; ++w) {
                                g[i][w] = src[w];
                                h[i][w] = s_hot[slot * W + w];
                            }
                        }
#pragma unroll
                        for (uint32_t i = 0; i < 4; ++i) {
                            const bool hot = i < c && (u[i] & kMsHotBit);
#pragma unroll
                            for (int w = 0; w < W; ++w) acc[w] |= g[i][w] | (hot ? h[i][w] : 0ull);
                        }
                        a.examined += c;
                        bool all = true;
#pragma unroll
                        for (int w = 0; w < W; ++w) all &= (acc[w] & need[w]) == need[w];
                        if (all) break;
                    }
                    uint32_t pc = 0;
                    uint64_t nw[W];
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        nw[w] = acc[w] & need[w];
                        pc += (uint32_t)__popcll(nw[w]);
                    }
                    if (pc) {
                        // a light vertex is scanned by this lane only and nothing
                        // else writes its words during the level: its bits are new,
                        // its visit word is ~need, its next-level word is zero
                        // (write-through stores: no atomic fetches the line)
                        found = true;
                        complete = true;
#pragma unroll
                        for (int w = 0; w < W; ++w) complete &= nw[w] == need[w];
#pragma unroll
                        for (int w = 0; w < W; ++w)
                            if (nw[w]) {
                                store_through(&D.visit[(size_t)v * W + w], (~need[w] & valid_word<W>(D, w)) | nw[w]);
                                store_through(&Fn[(size_t)v * W + w], nw[w]);
                            }
                        a.pairs += (uint64_t)dg * pc;
                        a.edges += dg;
                    }
                }
            }
        }
        const uint64_t mask = __ballot(found);
        const uint64_t cmask = __ballot(complete);
        if (lane == 0 && mask) {
            atomicOr((unsigned long long*)&fbn[wd], (unsigned long long)mask);
            a.size += (uint64_t)__popcll(mask);
        }
        if (lane == 0 && cmask) atomicOr((unsigned long long*)&D.done[wd], (unsigned long long)cmask);
        }
    }

    // heavy items: the lanes of a wave look at 64 items nwaves apart (the
    // chunks of one vertex land on different waves), the wave then scans each
    // item that still needs bits, 256 edges at a time
    for (uint64_t r = 0; gwave() + r * 64 * nwaves() < D.n_heavy; ++r) {
        const uint64_t it = gwave() + (r * 64 + lane) * nwaves();
        uint4 item = make_uint4(0, 0, 0, 0);
        uint64_t need[W];
        bool want = false;
#pragma unroll
        for (int w = 0; w < W; ++w) need[w] = 0;
        if (it < D.n_heavy) {
            item = D.heavy[it];
#pragma unroll
            for (int w = 0; w < W; ++w) {
                need[w] = ~D.visit[(size_t)item.x * W + w] & valid_word<W>(D, w);
                want |= need[w] != 0;
            }
        }
        uint64_t todo = __ballot(want);
        uint64_t mine[W];  // bits the wave found for this lane's item
#pragma unroll
        for (int w = 0; w < W; ++w) mine[w] = 0;
        while (todo) {
            const int j = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            const uint32_t y = __shfl(item.y, j, kWave), z = __shfl(item.z, j, kWave);
            uint64_t nd[W], acc[W];
#pragma unroll
            for (int w = 0; w < W; ++w) {
                nd[w] = __shfl(need[w], j, kWave);
                acc[w] = 0;
            }
            for (uint32_t xb = y; xb < z; xb += 4 * kWave) {
                // branch-free as in the light scan (z > y: the item is not empty)
                uint32_t u[4];
                bool in[4];
#pragma unroll
                for (uint32_t i = 0; i < 4; ++i) {
                    const uint32_t x = xb + i * kWave + lane;
                    in[i] = x < z;
                    u[i] = cx[min(x, z - 1)];
                }
                uint64_t fw[4];
#pragma unroll
                for (uint32_t i = 0; i < 4; ++i) fw[i] = direct ? ~0ull : fb[(u[i] & ~kMsHotBit) >> 6];
                uint64_t g[4][W], h[4][W];
#pragma unroll
                for (uint32_t i = 0; i < 4; ++i) {
                    const uint32_t id = u[i] & ~kMsHotBit;
                    const bool hot = (u[i] & kMsHotBit) != 0;
                    const bool take = in[i] && !hot && ((fw[i] >> (id & 63)) & 1ull);
                    const uint64_t* src = take ? F + (size_t)id * W : D.zero;
                    const uint32_t slot = hot ? id : 0u;
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        g[i][w] = src[w];
                        h[i][w] = s_hot[slot * W + w];
                    }
                }
#pragma unroll
                for (uint32_t i = 0; i < 4; ++i) {
                    const bool hot = in[i] && (u[i] & kMsHotBit);
#pragma unroll
                    for (int w = 0; w < W; ++w) acc[w] |= g[i][w] | (hot ? h[i][w] : 0ull);
                    a.examined += in[i];
                }
                bool all = true;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    uint64_t r = acc[w];
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) r |= __shfl_xor(r, off, kWave);
                    acc[w] = r;
                    all &= (r & nd[w]) == nd[w];
                }
                if (all) break;
            }
#pragma unroll
            for (int w = 0; w < W; ++w)
                if (lane == (uint32_t)j) mine[w] = acc[w] & nd[w];
        }
        // the round's discoveries together, each by the lane that holds the item
        // (one set of atomic round trips per round instead of one per item)
        bool any = false;
#pragma unroll
        for (int w = 0; w < W; ++w) any |= mine[w] != 0;
        if (any) ms_discover<W>(D, Fn, fbn, item.x, mine, a);
    }
}

// ----------------------------------------------------------------- top-down
// Two phases. Push: every frontier edge v -> u ORs F_l[v] (less the bits
// already in visit[u], when filtering) into next[u] and, when marking, sets
// u's bit in fb_{l+1} (fire-and-forget atomics; visit is not written, so the
// filter is exact for earlier levels and duplicates of this level merge in next). Collect (after a barrier): each marked vertex
// takes its new bits, clears next, and fb_{l+1} is rewritten exactly.
// kMark: also mark u in the next frontier bitmap fbn (the collect then visits
// marked words only); without it the collect scans every vertex. kFilter: bits
// already in visit or next (plain reads; a stale one only costs the atomic) are
// not pushed again, which pays once most pair-edges lead to reached vertices (on
// a skewed graph most pushes of a late big level land on a few hubs); while
// most are unreached the two reads are two random gathers per edge that filter
// nothing, and every edge is pushed (the collect keeps next & ~visit either way).
template <int W, bool kMark, bool kFilter>
__device__ __forceinline__ void ms_push4(const MsDev& D, uint64_t* fbn, const uint32_t (&u)[4], uint32_t c,
                                         const uint64_t (&f)[W]) {
    static_assert(kFilter || !kMark, "an unfiltered push marks nothing");
    if (!kFilter) {
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i)
#pragma unroll
            for (int w = 0; w < W; ++w)
                if (i < c && f[w]) atomicOr((unsigned long long*)&D.next[(size_t)u[i] * W + w], (unsigned long long)f[w]);
        return;
    }
    uint64_t vis[4][W], nxt[4][W];
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i)
#pragma unroll
        for (int w = 0; w < W; ++w) {
            // both loads unconditional (a lane that takes nothing reads the zero
            // words), so the eight are in flight together
            const bool ok = i < c && f[w];
            const uint64_t pv = *(ok ? D.visit + (size_t)u[i] * W + w : D.zero);
            const uint64_t pn = *(ok ? D.next + (size_t)u[i] * W + w : D.zero);
            vis[i][w] = ok ? pv : ~0ull;
            nxt[i][w] = ok ? pn : ~0ull;
        }
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        bool any = false;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint64_t x = f[w] & ~vis[i][w] & ~nxt[i][w];
            if (x) {
                atomicOr((unsigned long long*)&D.next[(size_t)u[i] * W + w], (unsigned long long)x);
                any = true;
            }
        }
        if (kMark && any) atomicOr((unsigned long long*)&fbn[u[i] >> 6], (unsigned long long)(1ull << (u[i] & 63)));
    }
}

// Pushes edges [b, e) of a vertex with frontier bits f, `step` lanes apart.
template <int W, bool kMark, bool kFilter>
__device__ __forceinline__ void ms_push_range(const MsDev& D, uint64_t* fbn, uint32_t b, uint32_t e, uint32_t first,
                                              uint32_t step, const uint64_t (&f)[W]) {
    for (uint32_t x = b + first; x < e; x += 4 * step) {
        uint32_t u[4];
        uint32_t c = 0;
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            const uint32_t y = x + i * step;
            u[i] = D.ci[min(y, e - 1)];  // x < e: a valid arc
            c += y < e;
        }
        ms_push4<W, kMark, kFilter>(D, fbn, u, c, f);
    }
}

template <int W, bool kMark, bool kFilter>
__device__ void ms_td_push(const MsDev& D, uint32_t level, bool lane_per_vertex) {
    const uint32_t cs = level % D.C, ns = (level + 1) % D.C;
    const uint64_t* F = D.F + (size_t)cs * D.n * W;
    const uint64_t* fb = D.fb + (size_t)cs * D.nb;
    uint64_t* fbn = D.fb + (size_t)ns * D.nb;
    const uint32_t lane = lane_id();
    const uint32_t nw = nwaves();

    // light frontier vertices: each lane owns a frontier word and, round by
    // round, takes its next set bit, so 64 vertices of 64 words run side by side
    for (uint64_t r0 = 0; r0 < D.nb; r0 += (uint64_t)nw * kWave) {
        const uint32_t wi = ms_word_of(r0, lane, nw);
        uint64_t fw = wi < D.nb ? fb[wi] : 0ull;
        while (__ballot(fw != 0)) {
            uint32_t v = 0xFFFFFFFFu;
            if (fw) {
                v = wi * 64 + (uint32_t)(__ffsll((long long)fw) - 1);
                fw &= fw - 1;
            }
            if (lane_per_vertex) {
                // binned: the lane pushes its vertex
                if (v != 0xFFFFFFFFu) {
                    const uint32_t b = D.rp[v], e = D.rp[v + 1];
                    if (e - b < D.heavy_min) {
                        uint64_t f[W];
#pragma unroll
                        for (int w = 0; w < W; ++w) f[w] = F[(size_t)v * W + w];
                        ms_push_range<W, kMark, kFilter>(D, fbn, b, e, 0, 1, f);
                    }
                }
            } else {
                // edge: the wave pushes the round's vertices one after another
                uint32_t b = 0, e = 0;
                if (v != 0xFFFFFFFFu) {
                    b = D.rp[v];
                    e = D.rp[v + 1];
                }
                uint64_t todo = __ballot(v != 0xFFFFFFFFu && e - b < D.heavy_min);
                while (todo) {
                    const int j = __ffsll((long long)todo) - 1;
                    todo &= todo - 1;
                    const uint32_t vj = __shfl(v, j, kWave), bj = __shfl(b, j, kWave), ej = __shfl(e, j, kWave);
                    uint64_t f[W];
#pragma unroll
                    for (int w = 0; w < W; ++w) f[w] = F[(size_t)vj * W + w];
                    ms_push_range<W, kMark, kFilter>(D, fbn, bj, ej, lane, kWave, f);
                }
            }
        }
    }
    // heavy frontier vertices: lanes select, among 64 items nwaves apart, those
    // whose vertex is in the frontier; the wave pushes each in turn
    for (uint64_t r = 0; gwave() + r * 64 * nwaves() < D.n_heavy; ++r) {
        const uint64_t it = gwave() + (r * 64 + lane) * nwaves();
        uint4 item = make_uint4(0, 0, 0, 0);
        bool want = false;
        if (it < D.n_heavy) {
            item = D.heavy[it];
            want = (fb[item.x >> 6] >> (item.x & 63)) & 1ull;
        }
        uint64_t todo = __ballot(want);
        while (todo) {
            const int j = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            const uint32_t v = __shfl(item.x, j, kWave), y = __shfl(item.y, j, kWave), z = __shfl(item.z, j, kWave);
            uint64_t f[W];
#pragma unroll
            for (int w = 0; w < W; ++w) f[w] = F[(size_t)v * W + w];
            ms_push_range<W, kMark, kFilter>(D, fbn, y, z, lane, kWave, f);
        }
    }
}

// ------------------------------------------------------- routed top-down
// A top-down level with a small frontier of hubs (the level after the sources:
// 1454 vertices, 2.9M arcs on rmat22_c2) spends its time in device-scope
// atomicOr pushes, which execute beyond the XCD's L2. Routed, the same level
// keeps every atomic in LDS:
//  1. queue (ms_route_queue): the frontier's vertices, degrees and words, one
//     reservation per wave; grid barrier;
//  2. bin (ms_route_bin): every workgroup takes an equal slice of the
//     frontier's arcs in queue order and, per round of kRouteChunk arcs,
//     counts its records per bin in LDS, reserves each bin's run with one
//     device atomic and writes the records there (write-through); grid barrier;
//  3. apply (ms_route_apply): a workgroup takes whole bins, ORs each record's
//     frontier words into an LDS copy of the bin's next words, then settles the
//     bin's vertices (new bits = next & ~visit) as the collect does.
// A bin's record area holds the in-degree sum of its vertices, the most
// records one level can put there. Discoveries equal the pushed level's.
template <int W>
__device__ void ms_route_queue(const MsDev& D, uint32_t level) {
    const uint32_t cs = level % D.C;
    const uint64_t* F = D.F + (size_t)cs * D.n * W;
    const uint64_t* fb = D.fb + (size_t)cs * D.nb;
    const uint32_t lane = lane_id(), nw = nwaves();
    for (uint64_t r0 = 0; r0 < D.nb; r0 += (uint64_t)nw * kWave) {
        const uint32_t wi = ms_word_of(r0, lane, nw);
        const uint64_t fw = wi < D.nb ? fb[wi] : 0ull;
        // the wave takes the nonzero words one at a time, a lane per bit (the
        // hubs of the degree numbering fill whole words)
        uint64_t todo = __ballot(fw != 0);
        while (todo) {
            const int j = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            const uint64_t word = __shfl(fw, j, kWave);
            const uint32_t wj = __shfl(wi, j, kWave);
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&D.rctl[0], (uint32_t)__popcll(word));
            base = __shfl(base, 0, kWave);
            if (base + (uint32_t)__popcll(word) > kRouteQ) {  // more frontier vertices than counted
                if (lane == 0) atomicExch(D.abort, kAbortRoute);
                continue;
            }
            if ((word >> lane) & 1ull) {
                const uint32_t q = base + (uint32_t)__popcll(word & ((1ull << lane) - 1ull));
                const uint32_t v = wj * 64 + lane;
                store_through(&D.rq[q], v);
                store_through(&D.rqd[q], D.rp[v + 1] - D.rp[v]);
#pragma unroll
                for (int w = 0; w < W; ++w) store_through(&D.rqf[(size_t)q * W + w], F[(size_t)v * W + w]);
            }
        }
    }
}

// Exclusive prefix of x[0, count) in place (count <= 8 * blockDim.x); x[count] = total.
__device__ __forceinline__ void block_exclusive_scan(uint32_t* x, uint32_t count) {
    __shared__ uint32_t s_wsum[kMsThreads / kWave];
    const uint32_t t = threadIdx.x, per = 8, lane = lane_id(), w = t / kWave;
    uint32_t v[8], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < per; ++j) {
        const uint32_t i = t * per + j;
        v[j] = i < count ? x[i] : 0u;
        sum += v[j];
    }
    const uint32_t incl = wave_inclusive_scan(sum);
    if (lane == kWave - 1) s_wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t j = 0; j < w; ++j) before += s_wsum[j];
    uint32_t run = before + incl - sum;
#pragma unroll
    for (uint32_t j = 0; j < per; ++j) {
        const uint32_t i = t * per + j;
        if (i < count) x[i] = run;
        run += v[j];
    }
    if (t == blockDim.x - 1) x[count] = run;  // the last thread's run is the total
    __syncthreads();
}

template <int W>
__device__ void ms_route_bin(const MsDev& D, uint32_t S, uint64_t E) {
    uint32_t* sD = (uint32_t*)g_ms_lds;   // S + 1 prefix of the queue's degrees
    uint32_t* sR = sD + kRouteQ + 1;      // S row starts of the queue's vertices
    uint32_t* sCnt = sR + kRouteQ;        // per bin: records of this round, then their LDS offsets
    uint32_t* sGb = sCnt + kRouteBins + 1;  // per bin: where this round's run starts in the bin's area
    uint32_t* sBlo = sGb + kRouteBins;    // n_bins + 1 bin starts
    uint32_t* sRec = sBlo + kRouteBins + 1;  // the round's records, by bin
    uint32_t* sDst = sRec + kRouteChunk;     // and where each goes
    const uint32_t t = threadIdx.x, T = blockDim.x, NB = D.n_bins;
    // the queue was written with write-through stores and the barrier's acquire
    // has invalidated stale lines: plain loads from here on
    for (uint32_t i = t; i < S; i += T) {
        const uint32_t v = D.rq[i];
        sD[i] = D.rqd[i];
        sR[i] = D.rp[v];
    }
    for (uint32_t b = t; b <= NB; b += T) sBlo[b] = D.blo[b];
    for (uint32_t b = t; b <= NB; b += T) sCnt[b] = 0;
    __syncthreads();
    block_exclusive_scan(sD, S);
    if ((uint64_t)sD[S] != E) {  // the queue's arcs are not the counted frontier arcs
        if (t == 0) atomicExch(D.abort, kAbortRoute);
        return;  // workgroup-uniform: sD[S] and E are the same in every thread
    }
    const uint64_t s0 = E * blockIdx.x / gridDim.x, s1 = E * (blockIdx.x + 1) / gridDim.x;
    for (uint64_t c0 = s0; c0 < s1; c0 += kRouteChunk) {  // rounds are uniform per workgroup
        // thread t takes arcs [i0, i0 + kRouteArcs): one search for the queue entry
        // holding i0 (the last k with sD[k] <= i0), then a walk across entries
        const uint64_t i0 = c0 + (uint64_t)t * kRouteArcs;
        uint32_t k = 0;
        if (i0 < s1) {
            uint32_t hi = S;
            while (hi - k > 1) {
                const uint32_t mid = (k + hi) >> 1;
                if (sD[mid] <= (uint32_t)i0) k = mid;
                else hi = mid;
            }
        }
        uint32_t u[kRouteArcs], kk[kRouteArcs];
        bool ok[kRouteArcs];
#pragma unroll
        for (uint32_t j = 0; j < kRouteArcs; ++j) {
            const uint32_t i = (uint32_t)(i0 + j);
            ok[j] = i0 + j < s1;
            if (ok[j])
                while (sD[k + 1] <= i) ++k;  // sD[S] = E > i: stops inside the queue
            kk[j] = k;
            u[j] = ok[j] ? D.ci[sR[k] + (i - sD[k])] : 0u;
        }
        uint32_t rec[kRouteArcs], bin[kRouteArcs], slot[kRouteArcs];
#pragma unroll
        for (uint32_t j = 0; j < kRouteArcs; ++j) bin[j] = ok[j] ? D.wbin[u[j] >> 6] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < kRouteArcs; ++j) {
            rec[j] = ((u[j] - sBlo[bin[j]]) << kRouteQBits) | kk[j];
            slot[j] = ok[j] ? atomicAdd(&sCnt[bin[j]], 1u) : 0u;
        }
        __syncthreads();
        for (uint32_t b = t; b < NB; b += T) {
            const uint32_t c = sCnt[b];
            uint32_t g = 0;
            if (c) {
                const uint32_t off = atomicAdd(&D.rctl[1 + b], c);
                g = D.bbase[b] + off;
                if ((uint64_t)off + c > (uint64_t)(D.bbase[b + 1] - D.bbase[b])) {  // past the bin's area
                    atomicExch(D.abort, kAbortRoute);
                    g = kNoSlot;
                }
            }
            sGb[b] = g;
        }
        block_exclusive_scan(sCnt, NB);  // counts -> LDS offsets of the bins' runs
        const uint32_t total = sCnt[NB];  // read before any thread clears sCnt at the round's end
#pragma unroll
        for (uint32_t j = 0; j < kRouteArcs; ++j)
            if (ok[j]) {
                const uint32_t p = sCnt[bin[j]] + slot[j];
                const uint32_t g = sGb[bin[j]];
                sRec[p] = rec[j];
                sDst[p] = g == kNoSlot ? kNoSlot : g + slot[j];
            }
        __syncthreads();
        // a bin's run is consecutive in LDS and in its area: consecutive lanes write consecutive words
        for (uint32_t p = t; p < total; p += T)
            if (sDst[p] != kNoSlot) store_through(&D.recs[sDst[p]], sRec[p]);
        for (uint32_t b = t; b <= NB; b += T) sCnt[b] = 0;
        __syncthreads();
    }
}

template <int W>
__device__ void ms_route_apply(const MsDev& D, uint32_t level, uint32_t S, MsAcc& a) {
    __shared__ uint32_t s_rc;
    uint64_t* sNext = g_ms_lds;
    const uint32_t ns = (level + 1) % D.C;
    uint64_t* Fn = D.F + (size_t)ns * D.n * W;
    uint64_t* fbn = D.fb + (size_t)ns * D.nb;
    const uint32_t t = threadIdx.x, T = blockDim.x, lane = lane_id(), wv = t / kWave, nwv = T / kWave;
    for (uint32_t b = blockIdx.x; b < D.n_bins; b += gridDim.x) {
        if (t == 0) s_rc = load_coherent(&D.rctl[1 + b]);  // device atomics made it
        __syncthreads();
        const uint32_t cnt = s_rc;
        if (cnt) {  // workgroup-uniform
            const uint32_t lo = D.blo[b], hi = D.blo[b + 1];
            for (uint32_t x = t; x < (hi - lo) * W; x += T) sNext[x] = 0;
            __syncthreads();
            const uint32_t* rec = D.recs + D.bbase[b];
            // kApply records per thread per round: their loads, then their frontier
            // words, then the LDS atomics (two memory round trips per round)
            constexpr uint32_t kApply = W == 1 ? 8 : 4;
            for (uint32_t base = 0; base < cnt; base += kApply * T) {
                uint32_t r[kApply];
#pragma unroll
                for (uint32_t j = 0; j < kApply; ++j) {
                    // written through before the barrier, whose acquire dropped stale lines: plain loads
                    const uint32_t i = base + j * T + t;
                    r[j] = i < cnt ? rec[i] : 0xFFFFFFFFu;
                }
                uint64_t f[kApply][W];
#pragma unroll
                for (uint32_t j = 0; j < kApply; ++j) {
                    const uint32_t off = r[j] >> kRouteQBits, q = r[j] & (kRouteQ - 1);
                    // past the count, or not a record of this level (never with matching counters): nothing to OR
                    const bool ok = off < hi - lo && q < S;
#pragma unroll
                    for (int w = 0; w < W; ++w) f[j][w] = ok ? D.rqf[(size_t)q * W + w] : 0ull;
                }
#pragma unroll
                for (uint32_t j = 0; j < kApply; ++j) {
                    const uint32_t off = r[j] >> kRouteQBits;
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        if (f[j][w]) atomicOr((unsigned long long*)&sNext[(size_t)off * W + w], (unsigned long long)f[j][w]);
                }
            }
            __syncthreads();
            // settle [lo, hi) as ms_td_collect_all does, a wave taking kSettle 64-vertex
            // words per round: the round's visit words and row offsets are loaded for
            // every vertex of the bin (unconditional loads: made conditional on the
            // vertex's next words, they took half again as long on a 16384-vertex bin)
            // before any is used
            constexpr uint32_t kSettle = W == 1 ? 4 : 2;
            for (uint32_t u00 = lo + wv * kWave * kSettle; u00 < hi; u00 += nwv * kWave * kSettle) {
                uint64_t nx[kSettle][W], vis[kSettle][W];
                uint32_t r0[kSettle], r1[kSettle];
#pragma unroll
                for (uint32_t j = 0; j < kSettle; ++j) {
                    const uint32_t u = u00 + j * kWave + lane;
                    const bool in = u < hi && u < D.n;
                    const uint32_t uc = in ? u : 0u;
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        nx[j][w] = in ? sNext[(size_t)(u - lo) * W + w] : 0ull;
                        vis[j][w] = D.visit[(size_t)uc * W + w];
                    }
                    r0[j] = D.rp[uc];
                    r1[j] = D.rp[uc + 1];
                }
#pragma unroll
                for (uint32_t j = 0; j < kSettle; ++j) {
                    const uint32_t u0 = u00 + j * kWave, u = u0 + lane;
                    if (u0 >= hi) break;  // wave-uniform
                    uint32_t pc = 0;
                    bool full = u < D.n;
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        full &= nx[j][w] != 0;
                        if (nx[j][w]) {
                            const size_t i = (size_t)u * W + w;
                            const uint64_t nwb = nx[j][w] & ~vis[j][w];
                            full &= (vis[j][w] | nwb) == valid_word<W>(D, w);
                            if (nwb) {
                                store_through(&Fn[i], nwb);
                                store_through(&D.visit[i], vis[j][w] | nwb);
                                pc += (uint32_t)__popcll(nwb);
                            }
                        }
                    }
                    const uint64_t found = __ballot(pc != 0);
                    const uint64_t cmask = __ballot(pc != 0 && full);
                    if (lane == 0 && cmask) atomicOr((unsigned long long*)&D.done[u0 >> 6], (unsigned long long)cmask);
                    if (pc) {
                        const uint64_t dg = r1[j] - r0[j];
                        a.pairs += dg * pc;
                        a.edges += dg;
                    }
                    if (lane == 0 && found) {
                        store_through(&fbn[u0 >> 6], found);
                        a.size += (uint64_t)__popcll(found);
                    }
                }
            }
            if (t == 0) store_through(&D.rctl[1 + b], 0u);  // the bin is empty for the next routed level
        }
        __syncthreads();  // s_rc and the LDS words are reused by the next bin
    }
    if (blockIdx.x == 0 && t == 0) store_through(&D.rctl[0], 0u);  // the queue, likewise
}

// Settles marked vertex u: keeps next & ~visit as its new bits, clears next.
// Returns the number of new (source, vertex) pairs.
template <int W>
__device__ __forceinline__ uint32_t ms_settle(const MsDev& D, uint64_t* Fn, uint32_t u, MsAcc& a) {
    uint32_t pc = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const size_t i = (size_t)u * W + w;
        const uint64_t vis = D.visit[i];
        const uint64_t nw = D.next[i] & ~vis;
        store_through(&D.next[i], 0ull);
        if (nw) {
            store_through(&Fn[i], nw);
            store_through(&D.visit[i], vis | nw);
            pc += (uint32_t)__popcll(nw);
        }
    }
    if (pc) {
        const uint64_t dg = D.rp[u + 1] - D.rp[u];
        a.pairs += dg * pc;
        a.edges += dg;
    }
    return pc;
}

// Marked words with more set bits than this are settled by the whole wave.
constexpr uint32_t kMsDenseBits = 8;

template <int W>
__device__ void ms_td_collect(const MsDev& D, uint32_t level, MsAcc& a) {
    const uint32_t ns = (level + 1) % D.C;
    uint64_t* Fn = D.F + (size_t)ns * D.n * W;
    uint64_t* fbn = D.fb + (size_t)ns * D.nb;
    const uint32_t lane = lane_id();
    const uint32_t nw = nwaves();
    for (uint64_t r0 = 0; r0 < D.nb; r0 += (uint64_t)nw * kWave) {
        const uint32_t wi = ms_word_of(r0, lane, nw);
        const uint64_t marked = wi < D.nb ? fbn[wi] : 0ull;
        // sparse words: each lane settles its word's vertices round by round
        const bool dense = (uint32_t)__popcll(marked) > kMsDenseBits;
        uint64_t mk = dense ? 0ull : marked, found = 0;
        while (__ballot(mk != 0)) {
            if (mk) {
                const uint32_t j = (uint32_t)(__ffsll((long long)mk) - 1);
                mk &= mk - 1;
                if (ms_settle<W>(D, Fn, wi * 64 + j, a)) found |= 1ull << j;
            }
        }
        if (marked && !dense) {
            store_through(&fbn[wi], found);
            a.size += (uint64_t)__popcll(found);
        }
        // dense words (in the degree numbering the hubs fill whole words): the
        // wave takes them one at a time, a lane per vertex
        uint64_t dw = __ballot(dense);
        while (dw) {
            const int j = __ffsll((long long)dw) - 1;
            dw &= dw - 1;
            const uint32_t wj = __shfl(wi, j, kWave);
            const uint64_t word = __shfl(marked, j, kWave);
            const bool f = ((word >> lane) & 1ull) && ms_settle<W>(D, Fn, wj * 64 + lane, a) != 0;
            const uint64_t fm = __ballot(f);
            if (lane == 0) {
                store_through(&fbn[wj], fm);
                a.size += (uint64_t)__popcll(fm);
            }
        }
    }
}

// The collect of an unmarked push: every vertex's next words are read
// (coalesced, 8 B per word and vertex); a wave owns 64 consecutive vertices,
// so it writes their frontier word once instead of one atomic per push.
template <int W>
__device__ void ms_td_collect_all(const MsDev& D, uint32_t level, MsAcc& a) {
    const uint32_t ns = (level + 1) % D.C;
    uint64_t* Fn = D.F + (size_t)ns * D.n * W;
    uint64_t* fbn = D.fb + (size_t)ns * D.nb;
    const uint32_t lane = lane_id();
    for (uint64_t b = (uint64_t)gwave() * kWave; b < D.n; b += (uint64_t)nwaves() * kWave) {
        const uint32_t u = (uint32_t)b + lane;
        uint32_t pc = 0;
        bool full = u < D.n;
        if (u < D.n) {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const size_t i = (size_t)u * W + w;
                const uint64_t nx = D.next[i];
                full &= nx != 0;
                if (nx) {
                    const uint64_t vis = D.visit[i];
                    const uint64_t nw = nx & ~vis;
                    full &= (vis | nw) == valid_word<W>(D, w);
                    store_through(&D.next[i], 0ull);
                    if (nw) {
                        store_through(&Fn[i], nw);
                        store_through(&D.visit[i], vis | nw);
                        pc += (uint32_t)__popcll(nw);
                    }
                }
            }
        }
        const uint64_t found = __ballot(pc != 0);
        const uint64_t cmask = __ballot(pc != 0 && full);
        if (lane == 0 && cmask) atomicOr((unsigned long long*)&D.done[b >> 6], (unsigned long long)cmask);
        if (pc) {
            const uint64_t dg = D.rp[u + 1] - D.rp[u];
            a.pairs += dg * pc;
            a.edges += dg;
        }
        if (lane == 0 && found) {
            store_through(&fbn[b >> 6], found);
            a.size += (uint64_t)__popcll(found);
        }
    }
}

// ---------------------------------------------------------- depths, cleanup
// Depth rows of sources s0 + b (b < ns) at one vertex: out[b * stride] =
// base + (bit b of the four planes) where `any` has bit b, INF where
// `unreached` has it too (a pair never reached has no plane bits). The loop
// is unrolled over the 64 bits, so every shift count is a constant and a
// source costs a few 32-bit bit-field operations; the store is streamed out,
// as the traversal does not read the depth rows again.
__device__ __forceinline__ void ms_write_depths(uint32_t* out, size_t stride, uint32_t ns, uint32_t base,
                                                const uint64_t (&plane)[kMsPlanes], uint64_t unreached, uint64_t any) {
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t sh = 32 * h;
        const uint32_t p0 = (uint32_t)(plane[0] >> sh), p1 = (uint32_t)(plane[1] >> sh);
        const uint32_t p2 = (uint32_t)(plane[2] >> sh), p3 = (uint32_t)(plane[3] >> sh);
        const uint32_t un = (uint32_t)(unreached >> sh), an = (uint32_t)(any >> sh);
#pragma unroll
        for (uint32_t j = 0; j < 32; ++j) {
            if (sh + j >= ns) return;
            const uint32_t d = ((p0 >> j) & 1u) | (((p1 >> j) & 1u) << 1) | (((p2 >> j) & 1u) << 2) |
                               (((p3 >> j) & 1u) << 3);
            const uint32_t val = (base + d) | (0u - ((un >> j) & 1u));
            if ((an >> j) & 1u) __builtin_nontemporal_store(val, out + (size_t)(sh + j) * stride);
        }
    }
}

// The same rows from a record's codes: base + code, INF for kCodeInf, no store
// for kCodeSkip.
__device__ __forceinline__ void ms_write_codes(uint32_t* out, size_t stride, uint32_t ns, uint32_t base, uint64_t q0,
                                               uint64_t q1, uint64_t q2, uint64_t q3) {
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t sh = 32 * h;
        const uint32_t p0 = (uint32_t)(q0 >> sh), p1 = (uint32_t)(q1 >> sh);
        const uint32_t p2 = (uint32_t)(q2 >> sh), p3 = (uint32_t)(q3 >> sh);
#pragma unroll
        for (uint32_t j = 0; j < 32; ++j) {
            if (sh + j >= ns) return;
            const uint32_t c = ((p0 >> j) & 1u) | (((p1 >> j) & 1u) << 1) | (((p2 >> j) & 1u) << 2) |
                               (((p3 >> j) & 1u) << 3);
            const uint32_t val = c == kCodeInf ? GGPU_INF : base + c;
            if (c != kCodeSkip) __builtin_nontemporal_store(val, out + (size_t)(sh + j) * stride);
        }
    }
}

typedef unsigned long long ms_u64x2 __attribute__((ext_vector_type(2)));

constexpr uint32_t kMsWin = kMsSlots - 1;  // levels one window retires at most

// After each launch of k_msbfs: the window's levels [base, base + cnt) are in
// slots (base + i) % C. Writes the depth of every pair first reached in them
// (relabeled runs: the depth records); final: also INF for pairs never
// reached, and visit cleared. Clears the slots it read (F where the frontier
// bit is set: elsewhere F is already zero). A wave owns 64 consecutive
// vertices, so their frontier words are one word per level; all of them, and
// then every F word of the vertex's levels, are loaded before any is used.
template <int W>
__global__ void __launch_bounds__(256) k_ms_finish(MsDev D) {
    const uint32_t base = (uint32_t)D.state[kStBase], cnt = (uint32_t)D.state[kStCnt];
    const bool final = D.state[kStFinal] != 0;
    const uint32_t lane = lane_id();
    uint32_t sl[kMsWin];
#pragma unroll
    for (uint32_t i = 0; i < kMsWin; ++i) sl[i] = (base + i) % D.C;
    const uint32_t stride = gridDim.x * blockDim.x;  // a multiple of 64: lane 0 owns v % 64 == 0
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < D.n; v += stride) {
        const uint32_t wd = v >> 6;
        uint64_t fw[kMsWin];
#pragma unroll
        for (uint32_t i = 0; i < kMsWin; ++i) fw[i] = i < cnt ? D.fb[(size_t)sl[i] * D.nb + wd] : 0ull;
        uint32_t levels = 0;  // bit i: v is in the frontier of level base + i
#pragma unroll
        for (uint32_t i = 0; i < kMsWin; ++i) levels |= (uint32_t)((fw[i] >> (v & 63)) & 1ull) << i;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const size_t vw = (size_t)v * W + w;
            uint64_t f[kMsWin];
#pragma unroll
            for (uint32_t i = 0; i < kMsWin; ++i)
                f[i] = ((levels >> i) & 1u) ? D.F[(size_t)sl[i] * D.n * W + vw] : 0ull;
            const uint64_t unreached = final ? ~D.visit[vw] : 0ull;
            uint64_t plane[kMsPlanes] = {0, 0, 0, 0}, reached = 0;
#pragma unroll
            for (uint32_t i = 0; i < kMsWin; ++i) {
                reached |= f[i];
#pragma unroll
                for (uint32_t p = 0; p < kMsPlanes; ++p)
                    if ((i >> p) & 1u) plane[p] |= f[i];
            }
            if (final) D.visit[vw] = 0ull;
            const uint64_t any = reached | unreached;
            if (D.perm) {
                // the record of (v, w), read by k_ms_unpermute: a pair with no
                // plane bits gets kCodeInf if never reached, else kCodeSkip
                const uint64_t hi = unreached | ~any;
                ms_u64x2* r = (ms_u64x2*)(D.rec + vw * kMsRec);
                r[0] = ms_u64x2{plane[0] | unreached, plane[1] | hi};
                r[1] = ms_u64x2{plane[2] | hi, plane[3] | hi};
                continue;
            }
            if (!__any(any != 0)) continue;
            const uint32_t s0 = (uint32_t)w * 64;
            ms_write_depths(D.dist + (size_t)s0 * D.n + v, D.n, min(64u, D.k - s0), base, plane, unreached, any);
        }
#pragma unroll
        for (uint32_t i = 0; i < kMsWin; ++i) {
            if (i >= cnt) continue;
            if ((levels >> i) & 1u) {
#pragma unroll
                for (int w = 0; w < W; ++w) D.F[(size_t)sl[i] * D.n * W + (size_t)v * W + w] = 0ull;
            }
            // the whole wave has read the frontier word of its 64-vertex group
            if (lane == 0) D.fb[(size_t)sl[i] * D.nb + wd] = 0ull;
        }
        if (final && lane == 0) D.done[wd] = 0ull;
    }
}

// Relabeled runs, after each k_ms_finish: caller id u takes the record of
// perm[u] (one random 32-byte read) and writes its depth rows with the wave's
// 64 consecutive u, as the direct depth pass does (four ids per thread with
// 16-byte stores measured slower: 256 against 232 us on the scale-22 batch).
template <int W>
__global__ void __launch_bounds__(256) k_ms_unpermute(MsDev D) {
    const uint32_t base = (uint32_t)D.state[kStBase];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < D.n; u += stride) {
        const uint32_t v = D.perm[u];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const ms_u64x2* r = (const ms_u64x2*)(D.rec + ((size_t)v * W + w) * kMsRec);
            const ms_u64x2 a = r[0], b = r[1];
            const uint32_t s0 = (uint32_t)w * 64;
            const uint32_t ns = min(64u, D.k - s0);
            const uint64_t valid = ns == 64 ? ~0ull : (1ull << ns) - 1;
            const uint64_t skip = ~a.x & a.y & b.x & b.y;  // kCodeSkip
            if (!__any((skip & valid) != valid)) continue;
            ms_write_codes(D.dist + (size_t)s0 * D.n + u, D.n, ns, base, a.x, a.y, b.x, b.y);
        }
    }
}

// ---------------------------------------------------------------- decisions
struct MsPolicyState {
    uint32_t bottom_up = 0;
    uint64_t prev_size = 0;
};

__device__ __forceinline__ void ms_features(const MsDev& D, uint32_t level, const MsAcc& t, uint64_t pairs_u,
                                            float* f) {
    // frontier vertices, frontier and unreached edge counts per source (pair-edges / k)
    const double kk = (double)D.k;
    policy_features(t.size, (uint64_t)((double)t.pairs / kk), (uint64_t)((double)pairs_u / kk), level,
                    D.n ? (double)D.m / (double)D.n : 0.0, f);
}

// y: the MLP's predictions for this level (learned policy with parameters).
__device__ uint32_t ms_choose(const MsDev& D, const float (&y)[3], uint32_t level, const MsAcc& t, uint64_t pairs_u,
                              MsPolicyState& ps) {
    uint32_t act;
    if (D.policy == GGPU_POLICY_SCRIPTED) {
        act = D.actions[level % D.n_actions];
        if (act > GGPU_ACTION_BU) act = GGPU_ACTION_BU;
        ps.bottom_up = act == GGPU_ACTION_BU;
    } else if (D.policy == GGPU_POLICY_LEARNED && D.mlp) {
        act = 0;
        for (uint32_t q = 1; q < 3; ++q)
            if (y[q] < y[act]) act = q;
        if (!(y[act] == y[act])) act = GGPU_ACTION_BU;
        ps.bottom_up = act == GGPU_ACTION_BU;
    } else {
        // the GAP rule on pair counts: frontier pair-edges against unreached pair-edges
        if (!ps.bottom_up && (double)t.pairs > (double)pairs_u / D.alpha) ps.bottom_up = 1;
        else if (ps.bottom_up && t.size < ps.prev_size && (double)t.size <= (double)D.n / D.beta) ps.bottom_up = 0;
        act = ps.bottom_up ? GGPU_ACTION_BU : GGPU_ACTION_TD_BINNED;
    }
    ps.prev_size = t.size;
    return act;
}

template <int W>
__global__ void __launch_bounds__(kMsThreads) k_msbfs(MsDev D) {
    __shared__ uint32_t s_act;
    __shared__ float s_mlp[GGPU_MLP_PARAMS];  // the learned policy, read at every level boundary
    if (D.mlp)
        for (uint32_t i = threadIdx.x; i < GGPU_MLP_PARAMS; i += blockDim.x) s_mlp[i] = D.mlp[i];
    uint64_t clk[4] = {0, 0, 0, 0};
    if (blockIdx.x == 0 && threadIdx.x == 0) clk[0] = wall_clock64();
    const bool resume = D.state[kStResume] != 0;
    if (!resume) {
        MsAcc acc;
        ms_seed<W>(D, acc);
        ms_flush(D, 0, acc);
        if (!grid_barrier(D.bar, D.abort)) return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) clk[1] = wall_clock64();

    MsPolicyState ps;
    uint32_t counts[3] = {0, 0, 0};
    uint32_t prev_action = GGPU_ACTION_TD_BINNED;
    uint32_t base = 0, level = 0;
    const unsigned long long total_pairs = (unsigned long long)D.k * D.m;
    uint64_t t_prev = 0;
    unsigned long long cum_pairs = 0;  // identical in every thread: t is the same everywhere
    uint32_t pending = 0xFFFFFFFFu;    // a resumed level's action, decided by the previous launch
    if (resume) {
        level = base = (uint32_t)D.state[kStLevel];
        pending = (uint32_t)D.state[kStAct];
        ps.bottom_up = (uint32_t)D.state[kStBottomUp];
        ps.prev_size = D.state[kStPrevSize];
        for (int i = 0; i < 3; ++i) counts[i] = (uint32_t)D.state[kStCounts + i];
        cum_pairs = D.state[kStCumPairs];
        t_prev = clk[1];
    }
    for (;; ++level) {
        const MsAcc t = ms_totals(D, level & 1);
        if (pending == 0xFFFFFFFFu) cum_pairs += t.pairs;
        const uint64_t pairs_u = total_pairs > cum_pairs ? total_pairs - cum_pairs : 0;
        uint32_t act = pending;
        if (pending == 0xFFFFFFFFu) {
            float y[3] = {0.f, 0.f, 0.f};
            if (D.policy == GGPU_POLICY_LEARNED && D.mlp && threadIdx.x < kWave && t.size) {
                // the level boundary's inference: wave 0 evaluates the MLP on this level's features
                float f[5];
                ms_features(D, level, t, pairs_u, f);
                mlp_forward_wave(s_mlp, f, y);
            }
            if (threadIdx.x == 0) {
                if (blockIdx.x == 0) {
                    const uint64_t now = wall_clock64();
                    if (level > 0 && level - 1 < D.trace_cap) {
                        if (prev_action == GGPU_ACTION_BU) D.trace[level - 1].edges = t.examined;
                        D.trace[level - 1].time_ms = (double)(now - t_prev) * D.tick_ms;
                    }
                    t_prev = now;
                }
                uint32_t a = 0xFFFFFFFFu;
                if (t.size) {
                    a = ms_choose(D, y, level, t, pairs_u, ps);
                    counts[a]++;
                    if (blockIdx.x == 0 && level < D.trace_cap) {
                        ggpu_level L{};
                        L.level = level;
                        L.action = a;
                        L.frontier = t.size;
                        L.edges = t.edges;
                        L.frontier_edges = t.edges;
                        L.unvisited_edges = (uint64_t)((double)pairs_u / (double)D.k);
                        L.direction = a == GGPU_ACTION_BU ? 'B' : 'T';
                        L.device = 'G';
                        D.trace[level] = L;
                        if (D.feat) ms_features(D, level, t, pairs_u, D.feat + (size_t)level * 5);
                    }
                }
                s_act = a;
            }
            __syncthreads();
            act = s_act;
            __syncthreads();
            if (act == 0xFFFFFFFFu) break;
            if (level + 1 - base == D.C) {
                // slot (level + 1) % C still holds level `base`: end the launch;
                // k_ms_finish retires levels [base, level), the next launch resumes here
                if (blockIdx.x == 0 && threadIdx.x == 0) {
                    unsigned long long* st = D.state;
                    st[kStResume] = 1;
                    st[kStLevel] = level;
                    st[kStAct] = act;
                    st[kStBottomUp] = ps.bottom_up;
                    st[kStPrevSize] = ps.prev_size;
                    for (int i = 0; i < 3; ++i) st[kStCounts + i] = counts[i];
                    st[kStCumPairs] = cum_pairs;
                    st[kStBase] = base;
                    st[kStCnt] = level - base;
                    st[kStFinal] = 0;
                    st[kStLoopTicks] += wall_clock64() - clk[1];
                    if (!resume) st[kStSetupTicks] = clk[1] - clk[0];
                }
                return;
            }
        }
        pending = 0xFFFFFFFFu;
        prev_action = act;
        MsAcc a;
        if (act == GGPU_ACTION_BU) {
            // the hot copy (~25 us per level on 256 workgroups) pays while at least
            // 1/32 of the pair-edges are unreached
            // frontier edges: arcs into the frontier (the graph is symmetric)
            ms_bottom_up<W>(D, level, pairs_u * 32 > total_pairs, (uint64_t)t.edges * 2 > D.m, a);
        } else {
            // a big push marks nothing and the collect scans every vertex (fewer
            // memory operations than one contended bitmap atomic per push)
            const bool scan_all = (uint64_t)t.edges * 8 > (uint64_t)D.n;
            const bool binned = act == GGPU_ACTION_TD_BINNED;
            // (routed only with one or two words per vertex: wider batches have bigger
            // first frontiers than the queue holds, and the branch would cost their
            // level loop registers)
            if (W <= 2 && D.n_bins && t.size <= kRouteQ && t.edges >= D.route_min) {
                // a small frontier with many arcs: routed through LDS bins
                if (blockIdx.x == 0 && threadIdx.x == 0) D.rctl[1 + kRouteBins]++;  // routed levels of this run
                ms_route_queue<W>(D, level);
                if (!grid_barrier(D.bar, D.abort)) return;
                ms_route_bin<W>(D, (uint32_t)t.size, t.edges);
                if (!grid_barrier(D.bar, D.abort)) return;
                ms_route_apply<W>(D, level, (uint32_t)t.size, a);
            } else {
                if (!scan_all) ms_td_push<W, true, true>(D, level, binned);
                else if (pairs_u * 2 > total_pairs) ms_td_push<W, false, false>(D, level, binned);
                else ms_td_push<W, false, true>(D, level, binned);
                if (!grid_barrier(D.bar, D.abort)) return;
                if (scan_all) ms_td_collect_all<W>(D, level, a);
                else ms_td_collect<W>(D, level, a);
            }
        }
        ms_flush(D, (level + 1) & 1, a);
        if (!grid_barrier(D.bar, D.abort)) return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        clk[2] = wall_clock64();
        D.out[0] = level;
        D.out[1] = 1;  // the traversal has ended (a launch that fills its window leaves 0)
        D.out[2] = counts[0];
        D.out[3] = counts[1];
        D.out[4] = counts[2];
        // setup of the first launch, level loops of all of them
        unsigned long long* st = D.state;
        D.out[6] = (uint32_t)(resume ? st[kStSetupTicks] : clk[1] - clk[0]);
        D.out[7] = (uint32_t)(clk[2] - clk[1] + (resume ? st[kStLoopTicks] : 0ull));
        st[kStBase] = base;
        st[kStCnt] = level - base;
        st[kStFinal] = 1;
        st[kStResume] = 0;  // the next run starts afresh
        st[kStLoopTicks] = 0;
    }
}

// ------------------------------------------------------------------ host
struct MsState {
    DBuf<uint64_t> visit, next, F, fb;
    DBuf<unsigned long long> part;
    DBuf<uint32_t> dist, sources, ctl;
    DBuf<uint4> heavy;
    DBuf<uint32_t> cib, hot;  // bottom-up col_idx with hot slots, hot ids (for hot_w words per vertex)
    DBuf<uint64_t> zero;      // kMsZeroWords zero words, never written
    DBuf<uint64_t> iso, done; // skip bitmaps: isolated / past-the-end vertices (per numbering), complete ones
    uint32_t hot_w = 0, n_hot = 0;
    // the numbering heavy / cib / hot were built in; for DEGREE the renumbered
    // graph (host copy for rebuilding them, device copy to traverse), perm[caller
    // id] = traversal id, and the depth records
    uint32_t order = 0xFFFFFFFFu;
    uint32_t auto_order = 0xFFFFFFFFu;  // what AUTO resolves to on this graph (computed once)
    Csr relabeled;
    std::vector<uint32_t> perm_host;
    DBuf<uint32_t> rp2, ci2, perm;
    DBuf<uint64_t> rec;
    DBuf<unsigned long long> st;  // kMsState loop-state words
    DBuf<ggpu_level> trace;
    DBuf<float> feat;
    DBuf<uint8_t> actions;
    uint32_t n_heavy = 0;
    bool heavy_ready = false;
    // routed top-down: bins of the traversal numbering for route_w words per vertex
    uint32_t route_w = 0, n_bins = 0;
    uint64_t route_min = kRouteMinEdges;  // test hook override
    DBuf<uint32_t> rq, rqd, rctl, blo, bbase, recs;
    DBuf<uint64_t> rqf;
    DBuf<uint16_t> wbin;
    uint32_t W = 0, C = 0, k = 0;
    uint32_t grid[5] = {0, 0, 0, 0, 0};
    bool clean = false;  // visit / next / F / fb all zero
    bool have = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // batch start / end, depth pass start / end
    ~MsState() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

static void ms_state_free(void* p) { delete static_cast<MsState*>(p); }

static MsState& ms_state(BfsEngine& e) {
    if (!e.ms_state) {
        e.ms_state = new MsState();
        e.ms_free = ms_state_free;
    }
    return *static_cast<MsState*>(e.ms_state);
}

static const void* ms_kernel(uint32_t W) {
    switch (W) {
        case 1: return (const void*)k_msbfs<1>;
        case 2: return (const void*)k_msbfs<2>;
        case 3: return (const void*)k_msbfs<3>;
        default: return (const void*)k_msbfs<4>;
    }
}

static const void* ms_finish_kernel(uint32_t W) {
    switch (W) {
        case 1: return (const void*)k_ms_finish<1>;
        case 2: return (const void*)k_ms_finish<2>;
        case 3: return (const void*)k_ms_finish<3>;
        default: return (const void*)k_ms_finish<4>;
    }
}

static const void* ms_unpermute_kernel(uint32_t W) {
    switch (W) {
        case 1: return (const void*)k_ms_unpermute<1>;
        case 2: return (const void*)k_ms_unpermute<2>;
        case 3: return (const void*)k_ms_unpermute<3>;
        default: return (const void*)k_ms_unpermute<4>;
    }
}
// enough 256-lane blocks to cover n, at most 16 per CU (grid-stride beyond)
static uint32_t ms_unpermute_grid(ggpu_ctx* ctx, uint32_t n) {
    const uint64_t want = ((uint64_t)n + 255) / 256;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)ctx->cus * 16));
}
static uint32_t ms_grid(ggpu_ctx* ctx, MsState& S, uint32_t W) {
    if (!S.grid[W]) {
        int per_cu = 0;
        GGPU_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ms_kernel(W), kMsThreads, 0));
        if (per_cu < 1) throw ArgError("batched BFS kernel does not fit on a compute unit", GGPU_EDEVICE);
        S.grid[W] = (uint32_t)ctx->cus;  // one block per CU: the barrier cost grows with the block count
    }
    return S.grid[W];
}

}  // namespace ggpu

using namespace ggpu;

ggpu::BfsEngine* ggpu_bfs_engine(ggpu_bfs* b);

extern "C" {

int ggpu_msbfs_run(ggpu_bfs* b, const uint32_t* sources, uint32_t k, const ggpu_bfs_opts* opt,
                   ggpu_msbfs_result* res, ggpu_level* trace, size_t trace_cap, size_t* trace_len) {
    BfsEngine* ep = ggpu_bfs_engine(b);
    if (!ep) return GGPU_EINVAL;
    BfsEngine& e = *ep;
    ggpu_ctx* ctx = e.ctx;
    ggpu_bfs_opts o;
    ggpu_bfs_opts_default(&o);
    if (opt) o = *opt;
    return guarded(ctx, [&] {
        if (!sources) throw ArgError("sources is null");
        if (k == 0 || k > 256) throw ArgError("batched BFS takes 1 to 256 sources");
        for (uint32_t i = 0; i < k; ++i)
            if (sources[i] >= e.n) throw ArgError("source out of range");
        if (o.policy > GGPU_POLICY_SCRIPTED) throw ArgError("unknown policy");
        if (o.policy == GGPU_POLICY_SCRIPTED && (!o.actions || !o.n_actions))
            throw ArgError("scripted policy needs actions");
        if (!(o.alpha > 0) || !(o.beta > 0)) throw ArgError("alpha and beta must be positive");
        if (o.batch_order > GGPU_BATCH_ORDER_AUTO) throw ArgError("unknown batch_order");
        GGPU_HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        MsState& S = ms_state(e);
        uint32_t order = o.batch_order;
        if (order == GGPU_BATCH_ORDER_AUTO) {
            // skewed degrees: the renumbering packs the most-gathered frontier words
            // into few lines; on near-regular graphs (grids, paths) it only adds the
            // un-permute and, on deep traversals, a relaunch per window of levels
            // (64-source batch on the 2932-level grid_c3: 178 ms given, 188 ms degree)
            if (S.auto_order == 0xFFFFFFFFu) {
                uint32_t dmax = 0;
                for (uint32_t v = 0; v < e.n; ++v) dmax = std::max(dmax, e.host.deg(v));
                S.auto_order = e.n && (double)dmax >= 8.0 * (double)e.m / (double)e.n ? GGPU_BATCH_ORDER_DEGREE
                                                                                      : GGPU_BATCH_ORDER_GIVEN;
            }
            order = S.auto_order;
        }
        const bool relabel = order == GGPU_BATCH_ORDER_DEGREE && e.n > 0;
        if (S.order != order) {
            S.heavy_ready = false;
            S.hot_w = 0;
            S.route_w = 0;
            S.relabeled = Csr();
            S.perm_host.clear();
            S.rp2.release();
            S.ci2.release();
            S.perm.release();
            S.rec.release();
            if (relabel) {
                S.perm_host = order_perm(e.host, "degree", 1);
                S.relabeled = permute(e.host, S.perm_host);
                S.rp2.alloc(S.relabeled.rp.size());
                S.ci2.alloc(S.relabeled.ci.size() + 1);
                S.perm.alloc(e.n);
                GGPU_HIP_TRY(hipMemcpy(S.rp2.p, S.relabeled.rp.data(), S.relabeled.rp.size() * 4, hipMemcpyHostToDevice));
                if (!S.relabeled.ci.empty())
                    GGPU_HIP_TRY(hipMemcpy(S.ci2.p, S.relabeled.ci.data(), S.relabeled.ci.size() * 4, hipMemcpyHostToDevice));
                GGPU_HIP_TRY(hipMemcpy(S.perm.p, S.perm_host.data(), (size_t)e.n * 4, hipMemcpyHostToDevice));
            }
            S.order = order;
        }
        const Csr& hg = relabel ? S.relabeled : e.host;
        const uint32_t heavy_min = relabel ? kMsHeavyDegree : kMsHeavyGiven;
        const uint32_t chunk = relabel ? kMsChunkDegree : kMsChunkGiven;
        const uint32_t W = (k + 63) / 64;
        const uint32_t n = e.n;
        const uint32_t nb = (n + 63) / 64;
        const size_t words = (size_t)n * W;
        if (!S.ev[0])
            for (auto& ev : S.ev) GGPU_HIP_TRY(hipEventCreate(&ev));
        if (!S.heavy_ready) {
            std::vector<uint4> items;
            for (uint32_t v = 0; v < n; ++v) {
                const uint32_t dg = hg.deg(v);
                if (dg < heavy_min) continue;
                for (uint32_t x = hg.rp[v]; x < hg.rp[v + 1]; x += chunk)
                    items.push_back(make_uint4(v, x, std::min(x + chunk, hg.rp[v + 1]), dg));
            }
            S.n_heavy = (uint32_t)items.size();
            S.heavy.alloc(items.size() + 1);
            if (!items.empty())
                GGPU_HIP_TRY(hipMemcpy(S.heavy.p, items.data(), items.size() * sizeof(uint4), hipMemcpyHostToDevice));
            // vertices without arcs (and the tail of the last word) are never scanned
            std::vector<uint64_t> iso(nb, 0);
            for (uint32_t v = 0; v < nb * 64; ++v)
                if (v >= n || hg.deg(v) == 0) iso[v >> 6] |= 1ull << (v & 63);
            S.iso.alloc(nb);
            GGPU_HIP_TRY(hipMemcpy(S.iso.p, iso.data(), (size_t)nb * 8, hipMemcpyHostToDevice));
            S.heavy_ready = true;
        }
        if (S.hot_w != W) {
            // the highest-degree vertices (ties by id), kMsHotWords / W of them
            const uint32_t H = n < (1u << 31) ? std::min<uint32_t>(kMsHotWords / W, n) : 0u;
            std::vector<uint32_t> order(n);
            for (uint32_t v = 0; v < n; ++v) order[v] = v;
            std::partial_sort(order.begin(), order.begin() + H, order.end(), [&](uint32_t a, uint32_t b) {
                const uint32_t da = hg.deg(a), db = hg.deg(b);
                return da != db ? da > db : a < b;
            });
            std::vector<uint32_t> slot(n, 0xFFFFFFFFu);
            for (uint32_t i = 0; i < H; ++i) slot[order[i]] = i;
            std::vector<uint32_t> cib(hg.ci.size());
            for (size_t x = 0; x < cib.size(); ++x) {
                const uint32_t u = hg.ci[x];
                cib[x] = slot[u] != 0xFFFFFFFFu ? (kMsHotBit | slot[u]) : u;
            }
            S.cib.alloc(cib.size() + 1);
            if (!cib.empty())
                GGPU_HIP_TRY(hipMemcpy(S.cib.p, cib.data(), cib.size() * 4, hipMemcpyHostToDevice));
            S.hot.alloc(H + 1);
            if (H) GGPU_HIP_TRY(hipMemcpy(S.hot.p, order.data(), (size_t)H * 4, hipMemcpyHostToDevice));
            S.n_hot = H;
            S.hot_w = W;
        }
        if (S.route_w != W) {
            // bins of the routed top-down: word-aligned vertex ranges of at most
            // kRouteBinV / W vertices; a bin also closes before a word that would take
            // its in-degree sum past cap_d (m / 384 to start with, so that bins carry
            // similar record counts and the hub words get bins of their own; doubled
            // until there are at most kRouteBins). Bin b's record area is its in-degree sum.
            std::vector<uint64_t> wdeg(nb, 0);
            for (const uint32_t u : hg.ci) wdeg[u >> 6]++;
            const uint64_t cap_v = kRouteBinV / W;
            std::vector<uint32_t> blo;
            std::vector<uint64_t> bdeg;
            std::vector<uint16_t> wbin(std::max<uint32_t>(nb, 1), 0);
            for (uint64_t cap_d = std::max<uint64_t>(e.m / 384, 1);; cap_d *= 2) {
                blo.assign(1, 0);
                bdeg.clear();
                uint64_t cur_v = 0, cur_d = 0;
                for (uint32_t wd = 0; wd < nb; ++wd) {
                    if (cur_v && (cur_v + 64 > cap_v || cur_d + wdeg[wd] > cap_d)) {
                        bdeg.push_back(cur_d);
                        blo.push_back(wd * 64);
                        cur_v = cur_d = 0;
                    }
                    cur_v += 64;
                    cur_d += wdeg[wd];
                    wbin[wd] = (uint16_t)std::min<size_t>(blo.size() - 1, 0xFFFF);
                }
                bdeg.push_back(cur_d);
                blo.push_back(nb * 64);
                if (blo.size() - 1 <= kRouteBins || cap_d > e.m) break;
            }
            const uint32_t NB = blo.size() - 1 <= kRouteBins ? (uint32_t)(blo.size() - 1) : 0u;
            std::vector<uint32_t> bbase(bdeg.size() + 1);  // + the end of the last area
            uint64_t acc = 0;
            for (size_t b = 0; b < bdeg.size(); ++b) {
                bbase[b] = (uint32_t)acc;
                acc += bdeg[b];
            }
            bbase[bdeg.size()] = (uint32_t)acc;
            S.blo.alloc(blo.size());
            S.bbase.alloc(bbase.size());
            S.wbin.alloc(wbin.size());
            GGPU_HIP_TRY(hipMemcpy(S.blo.p, blo.data(), blo.size() * 4, hipMemcpyHostToDevice));
            GGPU_HIP_TRY(hipMemcpy(S.bbase.p, bbase.data(), bbase.size() * 4, hipMemcpyHostToDevice));
            GGPU_HIP_TRY(hipMemcpy(S.wbin.p, wbin.data(), wbin.size() * 2, hipMemcpyHostToDevice));
            S.recs.alloc(std::max<uint64_t>(e.m, 1));
            S.rq.alloc(kRouteQ);
            S.rqd.alloc(kRouteQ);
            S.rqf.alloc((size_t)kRouteQ * 4);
            S.rctl.alloc(2 + kRouteBins);  // + the routed-level count of the last run
            GGPU_HIP_TRY(hipMemset(S.rctl.p, 0, (2 + kRouteBins) * sizeof(uint32_t)));
            S.n_bins = NB;
            S.route_w = W;
        }
        // level slots: kMsSlots, fewer if device memory is short (at least 2)
        if (S.W != W) {
            S.visit.release();
            S.next.release();
            S.F.release();
            S.fb.release();
            size_t free_b = 0, total_b = 0;
            GGPU_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
            const size_t slot_bytes = words * 8 + (size_t)nb * 8;
            const size_t dist_bytes = (size_t)256 * n * 4;
            const size_t avail = free_b > dist_bytes ? (free_b - dist_bytes) / 2 : 0;
            const uint32_t C = (uint32_t)std::min<size_t>(kMsSlots, std::max<size_t>(2, avail / std::max<size_t>(slot_bytes, 1)));
            S.visit.alloc(words);
            S.done.alloc(nb);
            S.next.alloc(words);
            S.F.alloc((size_t)C * words);
            S.fb.alloc((size_t)C * nb);
            S.W = W;
            S.C = C;
            S.clean = false;
        }
        if (!S.clean) {
            GGPU_HIP_TRY(hipMemsetAsync(S.visit.p, 0, words * 8, s));
            GGPU_HIP_TRY(hipMemsetAsync(S.done.p, 0, (size_t)nb * 8, s));
            GGPU_HIP_TRY(hipMemsetAsync(S.next.p, 0, words * 8, s));
            GGPU_HIP_TRY(hipMemsetAsync(S.F.p, 0, (size_t)S.C * words * 8, s));
            GGPU_HIP_TRY(hipMemsetAsync(S.fb.p, 0, (size_t)S.C * nb * 8, s));
            if (S.st.p) GGPU_HIP_TRY(hipMemsetAsync(S.st.p, 0, kMsState * 8, s));
            if (S.rctl.p) GGPU_HIP_TRY(hipMemsetAsync(S.rctl.p, 0, (2 + kRouteBins) * sizeof(uint32_t), s));
            S.clean = true;
        }
        if (!S.zero.p) {
            S.zero.alloc(kMsZeroWords);
            GGPU_HIP_TRY(hipMemsetAsync(S.zero.p, 0, kMsZeroWords * 8, s));
        }
        const uint32_t G = ms_grid(ctx, S, W);
        S.part.alloc((size_t)2 * G * kPartWords);
        S.dist.alloc((size_t)k * n);
        if (relabel) S.rec.alloc(words * kMsRec);
        if (!S.st.p) {
            S.st.alloc(kMsState);
            GGPU_HIP_TRY(hipMemsetAsync(S.st.p, 0, kMsState * 8, s));
        }
        S.k = k;
        S.sources.alloc(k);
        S.ctl.alloc(kBarWords + 16);
        const uint32_t tcap = (uint32_t)std::min<uint64_t>((uint64_t)n + 1, kMsTraceCap);
        S.trace.alloc(tcap);
        S.feat.alloc((size_t)tcap * 5);
        if (o.policy == GGPU_POLICY_SCRIPTED) {
            S.actions.alloc(o.n_actions);
            GGPU_HIP_TRY(hipMemcpyAsync(S.actions.p, o.actions, o.n_actions, hipMemcpyHostToDevice, s));
        }
        const bool learned = o.policy == GGPU_POLICY_LEARNED && ctx->policy_set[GGPU_POLICY_BATCHED];
        // pinned staging (the context's): sources (mapped to the traversal numbering)
        // in the first KiB, the control words read back after each launch behind them
        uint32_t* pin = (uint32_t*)ctx_pinned(ctx, 4096);
        for (uint32_t i = 0; i < k; ++i) pin[i] = relabel ? S.perm_host[sources[i]] : sources[i];
        GGPU_HIP_TRY(hipMemcpyAsync(S.sources.p, pin, 4ull * k, hipMemcpyHostToDevice, s));
        GGPU_HIP_TRY(hipMemsetAsync(S.ctl.p, 0, (kBarWords + 16) * sizeof(uint32_t), s));

        MsDev D{};
        D.rp = relabel ? S.rp2.p : e.rp.p;
        D.ci = relabel ? S.ci2.p : e.ci.p;
        D.perm = relabel ? S.perm.p : nullptr;
        D.rec = relabel ? S.rec.p : nullptr;
        D.state = S.st.p;
        D.cib = S.cib.p;
        D.hot = S.hot.p;
        D.n_hot = S.n_hot;
        D.heavy_min = heavy_min;
        D.rq = S.rq.p;
        D.rqd = S.rqd.p;
        D.rqf = S.rqf.p;
        D.rctl = S.rctl.p;
        D.blo = S.blo.p;
        D.bbase = S.bbase.p;
        D.wbin = S.wbin.p;
        D.recs = S.recs.p;
        D.n_bins = S.n_bins;
        D.route_min = S.route_min;
        GGPU_HIP_TRY(hipMemsetAsync(S.rctl.p + 1 + kRouteBins, 0, sizeof(uint32_t), s));
        D.visit = S.visit.p;
        D.next = S.next.p;
        D.F = S.F.p;
        D.fb = S.fb.p;
        D.zero = S.zero.p;
        D.iso = S.iso.p;
        D.done = S.done.p;
        D.part = S.part.p;
        D.dist = S.dist.p;
        D.sources = S.sources.p;
        D.heavy = S.heavy.p;
        D.bar = S.ctl.p;
        D.abort = S.ctl.p + kBarWords;
        D.out = S.ctl.p + kBarWords + 4;
        D.trace = S.trace.p;
        D.feat = S.feat.p;
        D.actions = o.policy == GGPU_POLICY_SCRIPTED ? S.actions.p : nullptr;
        D.mlp = learned ? ctx->policy_dev[GGPU_POLICY_BATCHED] : nullptr;
        D.alpha = o.alpha;
        D.beta = o.beta;
        int clock_khz = 0;
        GGPU_HIP_TRY(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, ctx->device));
        D.tick_ms = clock_khz > 0 ? 1.0 / clock_khz : 0.0;
        D.m = e.m;
        D.valid_last = (k % 64) ? ((1ull << (k % 64)) - 1) : ~0ull;
        D.n = n;
        D.k = k;
        D.nb = nb;
        D.n_heavy = S.n_heavy;
        D.C = S.C;
        D.trace_cap = tcap;
        D.n_actions = o.n_actions;
        D.policy = o.policy;

        void* args[] = {&D};
        S.have = false;
        S.clean = false;  // until the launch is known to have finished
        GGPU_HIP_TRY(hipEventRecord(S.ev[0], s));
        // plain launches, one workgroup per CU (the LDS footprint allows no second):
        // the barrier is the kernel's own, bounded (see the single-source loop).
        // One launch per window of C levels, each followed by the depth pass of
        // that window and, for a relabeled run, the un-permute of its records.
        uint32_t* ctl = pin + 256;
        double finish_ms = 0;
        for (uint32_t launch = 0;; ++launch) {
            if (launch) GGPU_HIP_TRY(hipMemsetAsync(S.ctl.p, 0, kBarWords * sizeof(uint32_t), s));
            GGPU_HIP_TRY(hipLaunchKernel(ms_kernel(W), dim3(G), dim3(kMsThreads), args, 0, s));
            GGPU_HIP_TRY(hipEventRecord(S.ev[2], s));
            GGPU_HIP_TRY(hipLaunchKernel(ms_finish_kernel(W), dim3(ms_unpermute_grid(ctx, n)), dim3(256), args, 0, s));
            if (relabel)
                GGPU_HIP_TRY(hipLaunchKernel(ms_unpermute_kernel(W), dim3(ms_unpermute_grid(ctx, n)), dim3(256),
                                             args, 0, s));
            GGPU_HIP_TRY(hipEventRecord(S.ev[3], s));
            GGPU_HIP_TRY(hipEventRecord(S.ev[1], s));  // the last launch's record is the one timed
            GGPU_HIP_TRY(hipMemcpyAsync(ctl, S.ctl.p + kBarWords, 16 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            GGPU_HIP_TRY(hipStreamSynchronize(s));
            if (ctl[0] == kAbortRoute)
                throw ArgError("batched BFS: a routed top-down level's queue or bins disagree with its counters",
                               GGPU_EDEVICE);
            if (ctl[0]) throw ArgError("batched BFS: grid barrier timed out", GGPU_ETIMEOUT);
            float fms = 0;
            GGPU_HIP_TRY(hipEventElapsedTime(&fms, S.ev[2], S.ev[3]));
            finish_ms += fms;
            if (ctl[5]) break;  // out[1]: the traversal has ended
        }

        S.clean = true;
        float ms = 0;
        GGPU_HIP_TRY(hipEventElapsedTime(&ms, S.ev[0], S.ev[1]));
        const uint32_t levels = ctl[4];
        e.last_gpu_ms = ms;
        e.last_levels = levels;
        if (res) {
            res->levels = levels;
            res->pad = 0;
            res->gpu_ms = ms;
            res->level_actions[0] = ctl[6];
            res->level_actions[1] = ctl[7];
            res->level_actions[2] = ctl[8];
            res->setup_ms = ctl[10] * D.tick_ms;
            res->levels_ms = ctl[11] * D.tick_ms;
            res->finish_ms = finish_ms;
        }
        if (trace_len) *trace_len = levels;
        if (trace && trace_cap) {
            const size_t cnt = std::min<size_t>({(size_t)levels, trace_cap, (size_t)tcap});
            if (cnt) GGPU_HIP_TRY(hipMemcpy(trace, S.trace.p, cnt * sizeof(ggpu_level), hipMemcpyDeviceToHost));
        }
        S.have = true;
    });
}

int ggpu_msbfs_record_levels(ggpu_bfs* b, const uint32_t* sources, uint32_t k, const ggpu_bfs_opts* opt,
                             float* features, float* times_ms, ggpu_level* trace, size_t cap, size_t* len) {
    BfsEngine* ep = ggpu_bfs_engine(b);
    if (!ep) return GGPU_EINVAL;
    ggpu_ctx* ctx = ep->ctx;
    return guarded(ctx, [&] {
        ggpu_bfs_opts o;
        ggpu_bfs_opts_default(&o);
        if (opt) o = *opt;
        o.policy = GGPU_POLICY_SCRIPTED;
        o.n_actions = 1;
        std::vector<ggpu_level> runs[3];
        std::vector<float> feat;
        size_t levels = 0;
        for (uint8_t a = 0; a < 3; ++a) {
            o.actions = &a;
            runs[a].resize(kMsTraceCap);
            size_t ln = 0;
            const int rc = ggpu_msbfs_run(b, sources, k, &o, nullptr, runs[a].data(), runs[a].size(), &ln);
            if (rc != GGPU_OK) throw ArgError(ctx->err, rc);
            if (a == 0) {
                levels = ln;
                MsState& S = *static_cast<MsState*>(ep->ms_state);
                const size_t kept = std::min<size_t>(levels, S.trace.count);
                feat.resize(kept * 5);
                if (kept) GGPU_HIP_TRY(hipMemcpy(feat.data(), S.feat.p, kept * 5 * sizeof(float), hipMemcpyDeviceToHost));
            } else if (ln != levels) {
                throw ArgError("level counts differ between actions", GGPU_ESTATE);
            }
        }
        const size_t cnt = std::min({levels, cap, feat.size() / 5});
        for (size_t l = 0; l < cnt; ++l) {
            if (features) std::memcpy(features + l * 5, feat.data() + l * 5, 5 * sizeof(float));
            if (times_ms)
                for (int a = 0; a < 3; ++a) times_ms[l * 3 + a] = (float)runs[a][l].time_ms;
            if (trace) trace[l] = runs[0][l];
        }
        if (len) *len = levels;
    });
}

int ggpu_test_msbfs_route(ggpu_bfs* b, uint64_t min_edges, uint32_t* routed_last) {
    BfsEngine* ep = ggpu_bfs_engine(b);
    if (!ep) return GGPU_EINVAL;
    return guarded(ep->ctx, [&] {
        MsState& S = ms_state(*ep);
        S.route_min = min_edges ? min_edges : kRouteMinEdges;
        if (routed_last) {
            *routed_last = 0;
            if (S.rctl.p) {
                GGPU_HIP_TRY(hipSetDevice(ep->ctx->device));
                GGPU_HIP_TRY(hipStreamSynchronize(ep->ctx->stream));
                GGPU_HIP_TRY(hipMemcpy(routed_last, S.rctl.p + 1 + kRouteBins, sizeof(uint32_t), hipMemcpyDeviceToHost));
            }
        }
    });
}

int ggpu_msbfs_fetch_dist(ggpu_bfs* b, uint32_t* out) {
    BfsEngine* ep = ggpu_bfs_engine(b);
    if (!ep) return GGPU_EINVAL;
    return guarded(ep->ctx, [&] {
        if (!out) throw ArgError("out is null");
        if (!ep->ms_state || !static_cast<MsState*>(ep->ms_state)->have)
            throw ArgError("no batched BFS result yet", GGPU_ESTATE);
        MsState& S = *static_cast<MsState*>(ep->ms_state);
        GGPU_HIP_TRY(hipSetDevice(ep->ctx->device));
        GGPU_HIP_TRY(hipMemcpyAsync(out, S.dist.p, 4ull * S.k * ep->n, hipMemcpyDeviceToHost, ep->ctx->stream));
        GGPU_HIP_TRY(hipStreamSynchronize(ep->ctx->stream));
    });
}

}  // extern "C"
