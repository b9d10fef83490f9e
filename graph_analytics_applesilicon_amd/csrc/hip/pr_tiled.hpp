// gpu_pull_adaptive's TILED layout (Jacobi sweeps): source-tiled partial sums
// for the rows with long in-lists, a form of two-dimensional SpMV blocking.
//
// On a skewed graph in degree order the rows of in-degree >= D ("tiled rows",
// 4% of rmat23_c5's vertices) hold ~80% of the arcs, and their in-lists draw
// on every part of the contribution vector, which is 8x one XCD's L2: one
// gather per arc through L2, a third of them misses, bounds the plain pull.
// Here the sources are cut into tiles of kTileSrc consecutive ids; a tile's
// contributions fit in LDS. For every tiled row j and tile t, the arcs of j
// from t form one contiguous segment of j's sorted in-list, and each
// non-empty segment yields one compensated partial sum P[t][j]:
//   * a workgroup loads tile t into LDS once per task and its waves take the
//     task's items: runs of <= kTileRows segments with <= kTileNnz arcs
//     (lanes per segment by length, Kahan per lane, TwoSum across lanes), or
//     chunks of kTileNnz arcs of a longer segment, whose partials are
//     combined in chunk order by the task at its end;
//   * tiles with few arcs into the tiled rows are merged into one "cold"
//     pseudo-tile whose arcs keep their source ids and gather from global
//     memory (no LDS);
//   * the merge kernel adds P[0..K][j] for each tiled row in tile order
//     (TwoSum, deterministic), then the partials of pieces of segments too
//     long for one task, and stores the row;
//   * the other rows are pulled in the same launch as runs of consecutive
//     rows, with sources below kTileSrc read from the LDS copy of tile 0.
// Arcs of dense tiles are stored as 16-bit offsets within the tile. P is
// dense per tile (zero where a row has no arcs from the tile; written once
// at build time and never again there).
#pragma once

#include "pagerank_kernels.hpp"

namespace ggpu {

constexpr uint32_t kTileSrc = 32768;  // sources per tile: 128 KiB of LDS
constexpr uint32_t kTileThreads = 1024;
constexpr uint32_t kTileWaves = kTileThreads / kWave;
constexpr uint32_t kTileNnz = 1024;   // arcs per item (longer segments are chunked)
constexpr uint32_t kTileRows = 64;    // segments (or untiled rows) per run
constexpr uint32_t kTileSlots = 512;  // chunk partials per task
constexpr uint32_t kNoTile = 0xFFFFFFFFu;

enum TileItemKind : uint32_t {
    TI_RUN = 0,      // {kind | k << 2, first entry, end entry, first arc}
    TI_CHUNK = 1,    // {kind | k << 2, task-local slot, first arc, end arc}
    TI_UNTILED = 2,  // {kind, first row, end row, 0}: rows outside the tiled set
};

struct TileTask {
    uint32_t item_begin, item_end;
    uint32_t lds_tile;  // source tile copied to LDS, or kNoTile
    uint32_t k;         // partial array written by the task's segments
    uint32_t seg_begin, seg_end;  // long segments: {tiled row, first slot, chunks, dest}
    uint32_t pad[2];
};
static_assert(sizeof(TileTask) == 32, "TileTask is 8 words (the host plan's layout)");

struct PrTiledDev {
    const uint4* items;
    const TileTask* tasks;
    const uint2* wg_tasks;    // per workgroup: [first task, end task)
    const uint4* segs;        // long segments of all tasks
    const uint16_t* tci;      // arcs of dense tiles: source - tile * kTileSrc
    const uint32_t* cci;      // arcs of the cold pseudo-tile: source ids
    const uint32_t* ent_row;  // segment entries: tiled row index
    const uint16_t* ent_len;  // segment entries: arcs (<= kTileNnz)
    float2* part;             // (K + 1) x R compensated partial sums
    float2* ovf;              // pieces of segments longer than a task's slots
    const uint32_t* optr;     // R + 1: each tiled row's overflow partials ...
    const uint32_t* oidx;     // ... oidx[optr[j] .. optr[j + 1])
    const uint32_t* trow;     // tiled row index -> vertex
    uint32_t R;               // tiled rows
    uint32_t K;               // dense tiles; partial array K is the cold pseudo-tile's
};

// Exclusive prefix over the wave.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v) { return wave_inclusive_scan(v) - v; }

// Lanes per segment for `rows` <= 64 segments: the largest power of two <= 64 / rows.
__device__ __forceinline__ uint32_t lanes_for(uint32_t rows) {
    const uint32_t g = rows ? kWave / rows : kWave;
    return g >= 64 ? 64 : (g >= 32 ? 32 : (g >= 16 ? 16 : (g >= 8 ? 8 : (g >= 4 ? 4 : (g >= 2 ? 2 : 1)))));
}

// Arcs one lane takes in an item: the planner keeps every run's and chunk's
// lanes at or below this, so a lane issues all its index loads at once.
constexpr uint32_t kLaneArcs = kTileNnz / kWave;

// An item's descriptor and, per lane, the values its processing starts from:
// runs (len, row) of the lane's segment entry, untiled runs (row start, len).
struct TileItem {
    uint4 I;
    uint32_t a, b;
};

__device__ __forceinline__ TileItem item_desc(const PrTiledDev& Z, uint32_t it, uint32_t end) {
    TileItem t;
    t.I = it < end ? Z.items[it] : make_uint4(3u, 0u, 0u, 0u);  // kind 3: no item
    t.a = t.b = 0;
    return t;
}

__device__ __forceinline__ void item_lanes(const PrSweep& P, const PrTiledDev& Z, TileItem& t) {
    const uint32_t kind = t.I.x & 3u, lane = lane_id(), nr = t.I.z - t.I.y;
    if (kind == TI_RUN) {
        if (lane < nr) {
            t.a = Z.ent_len[t.I.y + lane];
            t.b = Z.ent_row[t.I.y + lane];
        }
    } else if (kind == TI_UNTILED) {
        if (lane < nr) {
            const uint32_t b = P.rp[t.I.y + lane];
            t.a = P.rp[t.I.y + lane + 1] - b;
            t.b = b;
        }
    }
}

// Kahan sum of the arcs gl, gl + G, ... < len at idx_of(i): every index load of
// a round of kLaneArcs is issued before any is used, then value_of(index) for
// each, all unconditional (offsets clamped into [0, len), and idx_of(0) must be
// a valid address even when len == 0), so no load waits for another; only the
// additions are masked. `between` runs once, after the first index loads.
// Rounds of A arcs per lane (A = kLaneArcs: one round per item).
template <uint32_t A, class IdxOf, class ValOf, class Between>
__device__ __forceinline__ Kahan lane_sum(uint32_t len, uint32_t gl, uint32_t G, IdxOf idx_of, ValOf value_of,
                                          Between between) {
    Kahan k;
    const uint32_t last = len ? len - 1 : 0u;
    uint32_t r0 = 0;
    do {
        uint32_t idx[A];
#pragma unroll
        for (uint32_t j = 0; j < A; ++j) idx[j] = idx_of(min(gl + (r0 + j) * G, last));
        if (r0 == 0) between();
        float v[A];
#pragma unroll
        for (uint32_t j = 0; j < A; ++j) v[j] = value_of(idx[j]);
#pragma unroll
        for (uint32_t j = 0; j < A; ++j)
            if (gl + (r0 + j) * G < len) k.add(v[j]);
        r0 += A;
    } while (r0 * G < len);
    return k;
}

// lane_sum for untiled rows (arcs ci[b + i]): both loads for every arc, LDS
// for sources below `hot` and global for the others (the LDS-served lanes
// read element 0, one shared line), each kind issued for the whole round
// before any value is used.
template <uint32_t A, class Between>
__device__ __forceinline__ Kahan untiled_sum(const PrSweep& P, const float* s_tile, uint32_t hot, uint32_t b,
                                             uint32_t len, uint32_t gl, uint32_t G, Between between) {
    Kahan k;
    const uint32_t last = len ? len - 1 : 0u;
    uint32_t r0 = 0;
    do {
        uint32_t u[A];
#pragma unroll
        for (uint32_t j = 0; j < A; ++j) u[j] = __builtin_nontemporal_load(&P.ci[b + min(gl + (r0 + j) * G, last)]);
        if (r0 == 0) between();
        float g[A], l[A];
#pragma unroll
        for (uint32_t j = 0; j < A; ++j) g[j] = P.contrib_in[u[j] < hot ? 0u : u[j]];
#pragma unroll
        for (uint32_t j = 0; j < A; ++j) l[j] = s_tile[u[j] < hot ? u[j] : 0u];
#pragma unroll
        for (uint32_t j = 0; j < A; ++j)
            if (gl + (r0 + j) * G < len) k.add(u[j] < hot ? l[j] : g[j]);
        r0 += A;
    } while (r0 * G < len);
    return k;
}

// One item; `nx` (the wave's next item, descriptor loaded) gets its lane
// values loaded while this item's index loads are in flight.
__device__ __forceinline__ void tile_item(const PrSweep& P, const PrTiledDev& Z, const float* s_tile, uint32_t hot,
                                          const TileItem& t, TileItem& nx, float2* s_slots, float add, float d,
                                          float inv_n, PrPart& part) {
    const uint32_t lane = lane_id(), kind = t.I.x & 3u, k = t.I.x >> 2;
    auto prefetch = [&]() { item_lanes(P, Z, nx); };
    if (kind == TI_CHUNK) {
        const uint32_t a = t.I.z, len = t.I.w - t.I.z;
        Kahan acc;
        if (k < Z.K)
            acc = lane_sum<kLaneArcs>(len, lane, kWave, [&](uint32_t i) { return (uint32_t)Z.tci[a + i]; },
                           [&](uint32_t u) { return s_tile[u]; }, prefetch);
        else
            acc = lane_sum<kLaneArcs>(len, lane, kWave, [&](uint32_t i) { return Z.cci[a + i]; },
                           [&](uint32_t u) { return P.contrib_in[u]; }, prefetch);
        float hi = acc.s, lo = -acc.c;
        comp_reduce(hi, lo, kWave);
        if (lane == 0) s_slots[t.I.y] = make_float2(hi, lo);
        return;
    }
    if (kind != TI_RUN && kind != TI_UNTILED) {
        prefetch();
        return;
    }
    const uint32_t nr = t.I.z - t.I.y;
    const uint32_t G = lanes_for(nr), r = lane / G, gl = lane % G;
    const bool ok = r < nr;
    if (kind == TI_RUN) {
        const uint32_t off = wave_excl_scan(t.a);
        const uint32_t len_r = __shfl(t.a, (int)r, kWave);  // shuffles in every lane
        const uint32_t my_len = ok ? len_r : 0u;
        const uint32_t my_off = __shfl(off, (int)r, kWave), my_row = __shfl(t.b, (int)r, kWave);
        const uint32_t base = my_len ? t.I.w + my_off : 0u;  // empty lanes read arc 0
        Kahan acc;
        if (k < Z.K)
            acc = lane_sum<kLaneArcs>(my_len, gl, G, [&](uint32_t i) { return (uint32_t)Z.tci[base + i]; },
                           [&](uint32_t u) { return s_tile[u]; }, prefetch);
        else
            acc = lane_sum<kLaneArcs>(my_len, gl, G, [&](uint32_t i) { return Z.cci[base + i]; },
                           [&](uint32_t u) { return P.contrib_in[u]; }, prefetch);
        float hi = acc.s, lo = -acc.c;
        comp_reduce(hi, lo, G);
        if (ok && gl == 0) Z.part[(size_t)k * Z.R + my_row] = make_float2(hi, lo);
        return;
    }
    // untiled rows t.I.y .. t.I.z: the plain pull, sources below `hot` from the LDS copy of tile 0
    // shuffles in every lane (a lane may read a lane whose own row is empty)
    const uint32_t len_r = __shfl(t.a, (int)r, kWave), b_r = __shfl(t.b, (int)r, kWave);
    const uint32_t my_len = ok ? len_r : 0u;
    const uint32_t my_b = my_len ? b_r : 0u;  // empty lanes read arc 0
    Kahan acc = untiled_sum<kLaneArcs / 2>(P, s_tile, hot, my_b, my_len, gl, G, prefetch);
    float hi = acc.s, lo = -acc.c;
    comp_reduce(hi, lo, G);
    if (ok && gl == 0) pr_store(P, t.I.y + r, add + d * (hi + lo), part, inv_n);
}

// Copies src[0, cnt) (cnt <= kTileSrc) into LDS: every thread issues its loads
// in two rounds of 16 independent ones (addresses clamped, stores masked), so
// a fill costs two memory round trips instead of one per element.
__device__ __forceinline__ void tile_fill(const float* src, uint32_t cnt, float* s_tile) {
    constexpr uint32_t kPer = kTileSrc / kTileThreads, kRound = 16;
    static_assert(kPer % kRound == 0, "fill rounds");
    const uint32_t last = cnt ? cnt - 1 : 0u;
#pragma unroll
    for (uint32_t r = 0; r < kPer; r += kRound) {
        float v[kRound];
#pragma unroll
        for (uint32_t j = 0; j < kRound; ++j) v[j] = src[min(threadIdx.x + (r + j) * kTileThreads, last)];
#pragma unroll
        for (uint32_t j = 0; j < kRound; ++j) {
            const uint32_t i = threadIdx.x + (r + j) * kTileThreads;
            if (i < cnt) s_tile[i] = v[j];
        }
    }
}

// The tile sweep: one kTileThreads workgroup per CU runs its task list; each
// wave takes the task's items w, w + kTileWaves, ..., one item ahead.
__device__ __forceinline__ void pr_tiled_body(const PrSweep& P, const PrTiledDev& Z) {
    __shared__ float s_tile[kTileSrc];
    __shared__ float2 s_slots[kTileSlots];
    const double addd =
        (1.0 - (double)P.damping) / (double)P.n + (double)P.damping * P.acc_in->dangling / (double)P.n;
    const float add = (float)addd, d = P.damping, inv_n = 1.0f / (float)P.n;
    // wave-uniform: item descriptors come through scalar loads, item kinds are scalar branches
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    PrPart part;
    uint32_t loaded = kNoTile, hot = 0;
    const uint2 mine = Z.wg_tasks[blockIdx.x];
    for (uint32_t t = mine.x; t < mine.y; ++t) {
        const TileTask T = Z.tasks[t];
        if (T.lds_tile != kNoTile && T.lds_tile != loaded) {
            __syncthreads();  // every wave is done with the previous tile
            const uint32_t s0 = T.lds_tile * kTileSrc;
            hot = min(kTileSrc, P.n - s0);
            tile_fill(P.contrib_in + s0, hot, s_tile);
            loaded = T.lds_tile;
            __syncthreads();
        }
        TileItem cur = item_desc(Z, T.item_begin + w, T.item_end);
        item_lanes(P, Z, cur);
        for (uint32_t it = T.item_begin + w; it < T.item_end; it += kTileWaves) {
            TileItem nx = item_desc(Z, it + kTileWaves, T.item_end);
            tile_item(P, Z, s_tile, hot, cur, nx, s_slots, add, d, inv_n, part);
            cur = nx;
        }
        if (T.seg_end > T.seg_begin) {
            __syncthreads();  // every chunk partial of the task is in s_slots
            for (uint32_t s = T.seg_begin + threadIdx.x; s < T.seg_end; s += blockDim.x) {
                const uint4 S = Z.segs[s];
                float H = 0.f, L = 0.f;
                for (uint32_t c = 0; c < S.z; ++c) {
                    const float2 q = s_slots[S.y + c];
                    float a, e;
                    two_sum(H, q.x, a, e);
                    H = a;
                    L = L + q.y + e;
                }
                if (S.w == 0) Z.part[(size_t)T.k * Z.R + S.x] = make_float2(H, L);
                else Z.ovf[S.w - 1] = make_float2(H, L);
            }
            __syncthreads();  // s_slots is rewritten by the next task
        }
    }
    write_block_partials(part, P.partial, P.partial_cap, P.partial_base + blockIdx.x);
}

// The merge: lane per tiled row, the K + 1 partials in tile order.
__device__ __forceinline__ void pr_tiled_merge_body(const PrSweep& P, const PrTiledDev& Z) {
    const double addd =
        (1.0 - (double)P.damping) / (double)P.n + (double)P.damping * P.acc_in->dangling / (double)P.n;
    const float add = (float)addd, d = P.damping, inv_n = 1.0f / (float)P.n;
    PrPart part;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < Z.R; j += gridDim.x * blockDim.x) {
        float H = 0.f, L = 0.f;
        const float2* q = Z.part + j;
        uint32_t k = 0;
        for (; k + 4 <= Z.K + 1; k += 4) {
            float2 v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // read once: a streaming (non-temporal) load of the pair as one 64-bit word
                const unsigned long long w =
                    __builtin_nontemporal_load((const unsigned long long*)(q + (size_t)(k + i) * Z.R));
                v[i] = make_float2(__uint_as_float((uint32_t)w), __uint_as_float((uint32_t)(w >> 32)));
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float a, e;
                two_sum(H, v[i].x, a, e);
                H = a;
                L = L + v[i].y + e;
            }
        }
        for (; k < Z.K + 1; ++k) {
            const float2 v = q[(size_t)k * Z.R];
            float a, e;
            two_sum(H, v.x, a, e);
            H = a;
            L = L + v.y + e;
        }
        for (uint32_t x = Z.optr[j]; x < Z.optr[j + 1]; ++x) {
            const float2 v = Z.ovf[Z.oidx[x]];
            float a, e;
            two_sum(H, v.x, a, e);
            H = a;
            L = L + v.y + e;
        }
        pr_store(P, Z.trow[j], add + d * (H + L), part, inv_n);
    }
    write_block_partials(part, P.partial, P.partial_cap, P.partial_base + blockIdx.x);
}

}  // namespace ggpu
