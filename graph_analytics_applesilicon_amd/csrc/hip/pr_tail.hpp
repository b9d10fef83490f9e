// The HOT_TAIL PageRank layout's device side (plan: csrc/host/pr_tail_plan.hpp).
// No reference counterpart: the reference pulls every row from the whole
// contribution vector (src/shaders/pagerank.metal:306-381).
//
// In the degree-ordered numbering the HOT sweep serves sources [0, 30720) from
// LDS and gathers the rest through L2. The split rows (the heavy ones) give
// most of those gathers; their arcs from sources >= tail_begin are copied
// tile-major and summed per (tile, row) piece (<= 512 arcs of one split row in
// one tile) with Kahan + TwoSum into a (hi, lo) partial, by two kernels:
//  * k_pr_band, for the band tiles (band_tile <= kBandTile sources each): a
//    workgroup per task of <= kBandTaskItems items of one tile copies the
//    tile's contributions to LDS, and G lanes per piece sum the piece's arcs
//    from there. 64 KiB of LDS and few registers leave room for two 1024-lane
//    workgroups per CU: the waves are chains of dependent loads (item, piece
//    bounds, arc ids), and twice the waves in flight is what made the band pay;
//  * k_pr_tail, for the tail tiles beyond the band (tiles small enough for
//    one XCD's L2): one wave per item in item order (tile-major), so the waves
//    in flight read one or two tiles, which stay in L2; each wave stages its
//    item's gathers in LDS, then sums each piece.
// The HOT sweep leaves the split rows' heads (sources below tail_begin) as
// chunk partials; k_pr_tail_merge adds, per split row, its head partials in
// chunk order, then its tile partials in tile order (band tiles first), then its
// overflow partials, with TwoSum, and stores the row (pr_store: ranks,
// contributions and the FP64 dangling / sum / change partials).
#pragma once

#include "pagerank_kernels.hpp"

namespace ggpu {

constexpr uint32_t kTailWaves = 4;             // waves per k_pr_tail workgroup (one item each)
constexpr uint32_t kBandTile = 16384;          // most sources per band tile: 64 KiB of LDS
constexpr uint32_t kBandWaves = 16;            // waves per k_pr_band workgroup
constexpr uint32_t kBandTaskItems = 64;        // items per band task
constexpr uint32_t kNoHead = 0xFFFFFFFFu;      // a split row whose in-list lies entirely in the tail
constexpr uint32_t kMergeLanePartials = 8;     // split rows with more head + overflow partials: a wave each

struct PrTailDev {
    const uint32_t* tci;    // tail arcs (source ids), tile-major
    const uint32_t* px;     // piece p: tci[px[p], px[p + 1])
    const uint32_t* pdst;   // piece p: partial index
    const uint2* items;     // {first piece, end piece}
    uint32_t n_items;
    float2* part;           // [T][R] tile partials (zero where a row has no arc in a tile), then overflow slots
    uint64_t n_part;        // entries of part
    const uint32_t* srow;   // split row -> vertex
    const uint32_t* slr;    // split row -> long-row index of its head chunks in the stream layout, or kNoHead
    const uint32_t* ovptr;  // split row -> overflow slots [ovptr[j], ovptr[j + 1])
    const uint32_t* big;    // split rows merged by a whole wave
    uint32_t n_big;
    uint32_t R, T;
};

// One wave per item: the item's arcs are gathered (kGatherBatch loads in flight
// per lane, tci streamed with non-temporal loads) into the wave's LDS staging,
// then G lanes per piece sum it (G: the largest power of two with G * pieces
// <= 64).
// Items [item_base, n_items): the tail tiles' items.
__device__ __forceinline__ void pr_tail_body(const PrSweep& P, const PrTailDev& Z, uint32_t item_base) {
    __shared__ float s_stage[kTailWaves][kWaveNnz];
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    const uint32_t it = item_base + blockIdx.x * kTailWaves + w;
    if (it >= Z.n_items) return;  // wave-uniform; no workgroup barrier follows
    const uint2 item = Z.items[it];
    const uint32_t x0 = Z.px[item.x], cnt = Z.px[item.y] - x0;
    float* st = s_stage[w];
    for (uint32_t base = 0; base < cnt; base += kGatherBatch * kWave) {
        // unconditional loads (lanes past the item re-read its last arc; cnt >= 1), so
        // the compiler issues the pass's gathers together (masked: 839 against 826 us
        // per rmat23 sweep)
        uint32_t u[kGatherBatch];
#pragma unroll
        for (uint32_t j = 0; j < kGatherBatch; ++j) {
            const uint32_t i = base + j * kWave + lane;
            u[j] = __builtin_nontemporal_load(&Z.tci[x0 + min(i, cnt - 1)]);
        }
        float v[kGatherBatch];
#pragma unroll
        for (uint32_t j = 0; j < kGatherBatch; ++j) v[j] = P.contrib_in[u[j]];
#pragma unroll
        for (uint32_t j = 0; j < kGatherBatch; ++j) {
            const uint32_t i = base + j * kWave + lane;
            if (i < cnt) st[i] = v[j];
        }
    }
    wave_lds_sync();
    const uint32_t pieces = item.y - item.x;
    uint32_t G = pieces >= kWave ? 1u : kWave / pieces;
    G = G >= 64 ? 64 : (G >= 32 ? 32 : (G >= 16 ? 16 : (G >= 8 ? 8 : (G >= 4 ? 4 : (G >= 2 ? 2 : 1)))));
    const uint32_t per_round = kWave / G, rr = lane / G, gl = lane % G;
    for (uint32_t pb = 0; pb < pieces; pb += per_round) {
        // bounds and slot in one round of unconditional loads, as in pr_band_body
        const bool ok = pb + rr < pieces;
        const uint32_t q = ok ? item.x + pb + rr : item.x;
        const uint32_t a0 = Z.px[q], b0 = Z.px[q + 1], dst = Z.pdst[q];
        Kahan k;
        if (ok)
            for (uint32_t x = a0 - x0 + gl; x < b0 - x0; x += G) k.add(st[x]);
        float hi = k.s, lo = -k.c;
        comp_reduce(hi, lo, G);
        if (ok && gl == 0 && dst < Z.n_part) Z.part[dst] = make_float2(hi, lo);  // dst < n_part: always
    }
}

// One workgroup per band task {tile, first item, end item}: the tile's
// contributions (sources tail_begin + tile * band_tile onwards) are copied to
// LDS, then each wave takes the task's items w, w + kBandWaves, ... (the next
// item's descriptor loaded while the current one is summed) and G lanes per
// piece (G: the largest power of two with G * pieces <= 64) sum its arcs from
// the LDS tile, eight arc ids in flight per lane, each lane in arc order.
__device__ __forceinline__ void pr_band_body(const PrSweep& P, const PrTailDev& Z, const uint4* tasks,
                                             uint32_t tail_begin, uint32_t band_tile) {
    __shared__ float s_tile[kBandTile];
    const uint4 task = tasks[blockIdx.x];
    const uint32_t tb = tail_begin + task.x * band_tile;
    const uint32_t tcnt = min(band_tile, P.n - tb);  // the host keeps band_tile <= kBandTile and tb < n
    for (uint32_t i = threadIdx.x; i < tcnt; i += blockDim.x) s_tile[i] = P.contrib_in[tb + i];
    __syncthreads();
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    uint2 item = task.y + w < task.z ? Z.items[task.y + w] : make_uint2(0, 0);
    for (uint32_t it = task.y + w; it < task.z; it += kBandWaves) {
        const uint2 cur = item;
        item = it + kBandWaves < task.z ? Z.items[it + kBandWaves] : make_uint2(0, 0);
        const uint32_t pieces = cur.y - cur.x;
        uint32_t G = pieces >= kWave ? 1u : kWave / pieces;
        G = G >= 64 ? 64 : (G >= 32 ? 32 : (G >= 16 ? 16 : (G >= 8 ? 8 : (G >= 4 ? 4 : (G >= 2 ? 2 : 1)))));
        const uint32_t per_round = kWave / G, rr = lane / G, gl = lane % G;
        for (uint32_t pb = 0; pb < pieces; pb += per_round) {
            // the piece's bounds and partial slot in one round of unconditional loads (a lane
            // past the item's pieces re-reads the first piece and keeps nothing): 827 against
            // 808 us per rmat23 sweep with the loads masked and the slot read after the sum
            const bool ok = pb + rr < pieces;
            const uint32_t q = ok ? cur.x + pb + rr : cur.x;
            const uint32_t a0 = Z.px[q], b0 = Z.px[q + 1], dst = Z.pdst[q];
            const uint32_t a = ok ? a0 : 0u, b = ok ? b0 : 0u;
            Kahan k;
            for (uint32_t x = a + gl; x < b; x += 8 * G) {
                uint32_t u[8];
#pragma unroll
                for (uint32_t j = 0; j < 8; ++j) u[j] = Z.tci[min(x + j * G, b - 1)];
#pragma unroll
                for (uint32_t j = 0; j < 8; ++j) {
                    // a consistent plan keeps u in the tile: the clamp only keeps LDS reads in bounds
                    const float v = s_tile[min(u[j] - tb, kBandTile - 1)];
                    if (x + j * G < b) k.add(v);
                }
            }
            float hi = k.s, lo = -k.c;
            comp_reduce(hi, lo, G);
            if (ok && gl == 0 && dst < Z.n_part) Z.part[dst] = make_float2(hi, lo);  // dst < n_part: always
        }
    }
}

// Adds one (hi, lo) partial to a running TwoSum pair.
__device__ __forceinline__ void tail_add(float& H, float& L, float2 q) {
    float a, e;
    two_sum(H, q.x, a, e);
    H = a;
    L = L + q.y + e;
}

// Head chunk partials of split row j: lpart[first, first + count).
__device__ __forceinline__ uint2 tail_head(const PrSweep& P, const PrTailDev& Z, uint32_t j) {
    const uint32_t lr = Z.slr[j];
    if (lr == kNoHead) return make_uint2(0, 0);
    const uint2 e = P.stream.lrows[lr];
    return make_uint2(e.x, e.y & ~kLrowSplit);
}

// Split rows whose head and overflow partials number more than
// kMergeLanePartials (the list `big`) get a wave each: lane l adds partials l,
// l + 64, ... of the row's sequence (head chunks, tiles, overflow slots), then a
// TwoSum butterfly; every other split row gets a lane that adds its sequence in
// order.
__device__ __forceinline__ void pr_tail_merge_body(const PrSweep& P, const PrTailDev& Z) {
    const double addd = (1.0 - (double)P.damping) / (double)P.n + (double)P.damping * P.acc_in->dangling / (double)P.n;
    const float add = (float)addd, d = P.damping, inv_n = 1.0f / (float)P.n;
    const uint32_t lane = lane_id();
    const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nw = gridDim.x * blockDim.x / kWave;
    PrPart part;
    for (uint32_t i = gw; i < Z.n_big; i += nw) {
        const uint32_t j = Z.big[i];
        const uint2 h = tail_head(P, Z, j);
        const uint32_t o0 = Z.ovptr[j], no = Z.ovptr[j + 1] - o0;
        const uint32_t total = h.y + Z.T + no;
        float H = 0.f, L = 0.f;
        for (uint32_t q = lane; q < total; q += kWave) {
            float2 x;
            if (q < h.y) x = P.stream.lpart[h.x + q];
            else if (q < h.y + Z.T) x = Z.part[(size_t)(q - h.y) * Z.R + j];
            else x = Z.part[(size_t)Z.T * Z.R + o0 + (q - h.y - Z.T)];
            tail_add(H, L, x);
        }
        comp_reduce(H, L, kWave);
        if (lane == 0 && Z.srow[j] < P.n) pr_store(P, Z.srow[j], add + d * (H + L), part, inv_n);
    }
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < Z.R; j += gridDim.x * blockDim.x) {
        const uint2 h = tail_head(P, Z, j);
        const uint32_t o0 = Z.ovptr[j], o1 = Z.ovptr[j + 1];
        if (h.y + (o1 - o0) > kMergeLanePartials) continue;  // in `big`
        float H = 0.f, L = 0.f;
        for (uint32_t q = 0; q < h.y; ++q) tail_add(H, L, P.stream.lpart[h.x + q]);
        // tile partials eight at a time: their loads (clamped to the last tile, unconditional) are
        // issued before the adds, which keep tile order
        for (uint32_t t0 = 0; t0 < Z.T; t0 += 8) {
            float2 x[8];
#pragma unroll
            for (uint32_t i = 0; i < 8; ++i) x[i] = Z.part[(size_t)min(t0 + i, Z.T - 1) * Z.R + j];
#pragma unroll
            for (uint32_t i = 0; i < 8; ++i)
                if (t0 + i < Z.T) tail_add(H, L, x[i]);
        }
        for (uint32_t o = o0; o < o1; ++o) tail_add(H, L, Z.part[(size_t)Z.T * Z.R + o]);
        if (Z.srow[j] < P.n) pr_store(P, Z.srow[j], add + d * (H + L), part, inv_n);
    }
    write_block_partials(part, P.partial, P.partial_cap, P.partial_base + blockIdx.x);
}

}  // namespace ggpu
