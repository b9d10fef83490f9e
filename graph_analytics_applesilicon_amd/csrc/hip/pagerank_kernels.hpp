// PageRank pull kernels for gfx950.
//
// Definition (src/cpu/pagerank_cpu.hpp:4-8):
//   r_new[v] = (1-d)/n + d * ( sum_{u in in(v)} r[u]/outdeg[u] + dangling/n )
// One launch = one sweep over a set of vertex "sections" (the in-degree bins
// of gpu_pull_adaptive, the colour sets / natural lists of gpu_pull_gs, or
// all vertices). Each vertex's in-list sum is FP32 with Kahan compensation
// per lane (pagerank.metal:34-42) and a TwoSum-compensated combination across
// lanes (the reference's simd_sum drops the compensation, pagerank.metal:325).
// The same pass writes r_new, the next sweep's contributions r_new/outdeg and
// FP64 per-block partials of the dangling mass, the L1 change, the rank sum
// and the max relative change; the last block of the launch reduces them in
// a fixed order (deterministic) into FP64 accumulators.
#pragma once

#include "dev.hpp"

namespace ggpu {

constexpr int kMaxSections = 3;
constexpr uint32_t kStreamNnz = 2048;
constexpr uint32_t kStreamRows = 1024;
constexpr uint32_t kLongChunk = 2048;

enum SectionMode : uint32_t { SEC_LANE = 0, SEC_GROUP = 1, SEC_BLOCK = 2 };

struct PrSection {
    const uint32_t* list;  // vertex ids, or null: ids first .. first+count-1
    uint32_t first;
    uint32_t count;
    uint32_t mode;         // SectionMode
    uint32_t lanes;        // SEC_GROUP: lanes per vertex (power of two <= 64)
    uint32_t blk_begin;    // block range [blk_begin, blk_end) of this section
    uint32_t blk_end;
};

// A long row whose item count carries this bit is a HOT_TAIL split row's head:
// its chunk partials are always stored, and the tail merge (pr_tail.hpp), not
// pr_longfix_body, combines them with the row's tail partials.
constexpr uint32_t kLrowSplit = 0x80000000u;
// Set in a long item's fourth word (the long-row index) when the item is a split
// row's head chunk whose arc ids are stored as u16 in hci16[y, z), y even (every source
// of a head lies below the LDS copy's end, which is below 2^16); only the HOT
// sweep runs such items.
constexpr uint32_t kLiHead16 = 0x80000000u;

struct PrStreamLayout {
    const uint4* sitems;  // {first row, end row, first edge, end edge}
    uint32_t n_sitems;
    const uint4* litems;  // {row, first edge, end edge, long-row index}
    const uint16_t* hci16;  // the head chunks' arc ids (items flagged kLiHead16)
    uint32_t n_litems;
    const uint2* lrows;   // {first item, item count (| kLrowSplit)} per long row
    float2* lpart;        // per long item (hi, lo)
    uint32_t n_lrows;
};

// FP64 accumulators of one sweep.
struct PrAcc {
    double dangling;  // rank on dangling vertices after the sweep
    double delta;     // sum |r_new - r_old|
    double sum;       // sum r_new
    double maxrel;    // max |r_new - r_old| / max(r_old, 1/n)
};

struct PrSweep {
    const uint32_t* rp;
    const uint32_t* ci;
    const uint32_t* outdeg;
    const float* contrib_in;  // contributions read by this sweep
    float* contrib_out;       // contributions of r_new (== contrib_in when in place)
    const float* r_old;       // previous ranks (for delta / maxrel); may be null
    float* r_new;
    const PrAcc* acc_in;      // dangling of the state being pulled from
    double* partial;          // 4 * partial_cap doubles; this launch writes slots partial_base + block
    uint32_t partial_cap;
    uint32_t partial_base;
    uint32_t n;
    float damping;
    PrSection sec[kMaxSections];
    uint32_t nsec;
    PrStreamLayout stream;    // used when use_stream != 0 (instead of the sections)
    uint32_t use_stream;      // 1: row blocks (k_pr_sweep), 2: HOT layout (k_pr_sweep_hot)
    uint32_t stream_multi;    // some long row spans several chunks (needs pr_longfix_body)
};

struct Kahan {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// Knuth TwoSum: a + b == s + e exactly.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

// Combines (hi, lo) pairs across `width` lanes (aligned groups) with TwoSum.
__device__ __forceinline__ void comp_reduce(float& hi, float& lo, uint32_t width) {
    for (uint32_t off = width >> 1; off > 0; off >>= 1) {
        const float ohi = __shfl_xor(hi, (int)off, kWave);
        const float olo = __shfl_xor(lo, (int)off, kWave);
        float s, e;
        two_sum(hi, ohi, s, e);
        hi = s;
        lo = lo + olo + e;
    }
}

struct PrPart {
    double dang = 0, delta = 0, sum = 0, maxrel = 0;
};

__device__ __forceinline__ void pr_store(const PrSweep& P, uint32_t v, float val, PrPart& part, float inv_n) {
    P.r_new[v] = val;
    const uint32_t od = P.outdeg[v];
    P.contrib_out[v] = od ? val / (float)od : 0.0f;
    if (!od) part.dang += (double)val;
    part.sum += (double)val;
    if (P.r_old) {
        const float old = P.r_old[v];
        const float dlt = fabsf(val - old);
        part.delta += (double)dlt;
        const double rel = (double)dlt / (double)fmaxf(old, inv_n);
        part.maxrel = rel > part.maxrel ? rel : part.maxrel;
    }
}

// pr_store with the vertex's out-degree and previous rank already loaded (0 when
// the sweep keeps no previous ranks).
__device__ __forceinline__ void pr_store_pre(const PrSweep& P, uint32_t v, float val, uint32_t od, float old,
                                             PrPart& part, float inv_n) {
    P.r_new[v] = val;
    P.contrib_out[v] = od ? val / (float)od : 0.0f;
    if (!od) part.dang += (double)val;
    part.sum += (double)val;
    if (P.r_old) {
        const float dlt = fabsf(val - old);
        part.delta += (double)dlt;
        const double rel = (double)dlt / (double)fmaxf(old, inv_n);
        part.maxrel = rel > part.maxrel ? rel : part.maxrel;
    }
}

__device__ __forceinline__ float in_sum_lane(const PrSweep& P, uint32_t v) {
    Kahan k;
    const uint32_t b = P.rp[v], e = P.rp[v + 1];
    for (uint32_t x = b; x < e; ++x) k.add(P.contrib_in[P.ci[x]]);
    return k.s - k.c;
}

// ------------------------------------------------------------ stream layout
// gpu_pull_adaptive's default layout (the CSR-Adaptive scheme for SpMV):
//  * row blocks: consecutive vertices whose in-lists total <= kStreamNnz
//    entries. The block reads the col_idx span once, coalesced, and issues
//    every contrib gather of the span at once into LDS; then G lanes per row
//    (G = blockDim / rows, a power of two <= 64) sum that row's LDS slice with
//    Kahan and a TwoSum combine.
//  * long rows (> kStreamNnz entries): chunks of <= kLongChunk entries, one
//    block each; chunk partials (hi, lo) are combined in chunk order by the
//    last chunk to finish (deterministic).

__device__ __forceinline__ void block_comp_reduce(float& hi, float& lo, float* s_hi, float* s_lo) {
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave, nwv = blockDim.x / kWave;
    comp_reduce(hi, lo, kWave);
    __syncthreads();
    if (lane == 0) {
        s_hi[w] = hi;
        s_lo[w] = lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float H = s_hi[0], L = s_lo[0];
        for (uint32_t j = 1; j < nwv; ++j) {
            float a, e;
            two_sum(H, s_hi[j], a, e);
            H = a;
            L = L + s_lo[j] + e;
        }
        hi = H;
        lo = L;
    }
}

// Gathers contrib[ci[x0 + i]] for i in [0, cnt) issuing kGatherBatch
// independent col_idx loads per thread before their gathers, so each thread
// keeps that many memory requests in flight. col_idx is streamed once, with
// non-temporal loads (it must not evict the hot contributions from L2).
constexpr uint32_t kGatherBatch = 8;

template <class Sink>
__device__ __forceinline__ void gather_batched(const PrSweep& P, uint32_t x0, uint32_t cnt, Sink sink) {
    const uint32_t T = blockDim.x;
    for (uint32_t base = 0; base < cnt; base += kGatherBatch * T) {
        uint32_t u[kGatherBatch];
#pragma unroll
        for (uint32_t j = 0; j < kGatherBatch; ++j) {
            const uint32_t i = base + j * T + threadIdx.x;
            u[j] = i < cnt ? __builtin_nontemporal_load(&P.ci[x0 + i]) : 0u;
        }
        float v[kGatherBatch];
#pragma unroll
        for (uint32_t j = 0; j < kGatherBatch; ++j) {
            const uint32_t i = base + j * T + threadIdx.x;
            v[j] = i < cnt ? P.contrib_in[u[j]] : 0.0f;
        }
#pragma unroll
        for (uint32_t j = 0; j < kGatherBatch; ++j) {
            const uint32_t i = base + j * T + threadIdx.x;
            if (i < cnt) sink(i, v[j]);
        }
    }
}

// Items i < n_litems are long-row chunks, the rest row blocks; block b takes
// items b, b + gridDim, ... so every block gets a mix of equal-sized work.
__device__ __forceinline__ void pr_stream_body(const PrSweep& P, const PrStreamLayout& Lyt, float add, float d,
                                               float inv_n, PrPart& part) {
    __shared__ float s_val[kStreamNnz];
    __shared__ float s_hi[16], s_lo[16];
    const uint32_t T = blockDim.x;
    const uint32_t n_items = Lyt.n_litems + Lyt.n_sitems;
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        if (it < Lyt.n_litems) {
            const uint4 li = Lyt.litems[it];
            const uint2 lr = Lyt.lrows[li.w];  // loaded beside the gathers, not after the sum
            Kahan k;
            gather_batched(P, li.y, li.z - li.y, [&](uint32_t, float x) { k.add(x); });
            float hi = k.s, lo = -k.c;
            block_comp_reduce(hi, lo, s_hi, s_lo);
            if (threadIdx.x == 0) {
                if (lr.y == 1) pr_store(P, li.x, add + d * (hi + lo), part, inv_n);
                else Lyt.lpart[it] = make_float2(hi, lo);  // combined by pr_longfix_body
            }
            __syncthreads();
            continue;
        }
        const uint4 si = Lyt.sitems[it - Lyt.n_litems];
        const uint32_t r0 = si.x, r1 = si.y;  // r1 > r0
        const uint32_t rows = r1 - r0;
        uint32_t G = T / rows;
        G = G >= 64 ? 64 : (G >= 32 ? 32 : (G >= 16 ? 16 : (G >= 8 ? 8 : (G >= 4 ? 4 : (G >= 2 ? 2 : 1)))));
        const uint32_t per_round = T / G, rr = threadIdx.x / G, gl = threadIdx.x % G;
        // row bounds, out-degree and previous rank one round ahead, as in pr_hot_body
        uint32_t rc = min(r0 + rr, r1 - 1);
        uint32_t ra = P.rp[rc], rz = P.rp[rc + 1], od = P.outdeg[rc];
        float old = P.r_old ? P.r_old[rc] : 0.0f;
        const uint32_t x0 = si.z, x1 = si.w;  // rp[r0], rp[r1], in the item: no load round
        gather_batched(P, x0, x1 - x0, [&](uint32_t i, float x) { s_val[i] = x; });
        __syncthreads();
        for (uint32_t rb = 0; rb < rows; rb += per_round) {
            const uint32_t r = r0 + rb + rr;
            const bool ok = r < r1;
            const uint32_t ca = ra, cz = rz, cod = od;
            const float cold = old;
            rc = min(r + per_round, r1 - 1);
            ra = P.rp[rc];
            rz = P.rp[rc + 1];
            od = P.outdeg[rc];
            old = P.r_old ? P.r_old[rc] : 0.0f;
            Kahan k;
            if (ok)
                for (uint32_t x = ca - x0 + gl; x < cz - x0; x += G) k.add(s_val[x]);
            float hi = k.s, lo = -k.c;
            comp_reduce(hi, lo, G);
            if (ok && gl == 0) pr_store_pre(P, r, add + d * (hi + lo), cod, cold, part, inv_n);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------- hot-cached layout
// gpu_pull_adaptive's HOT layout (Jacobi sweeps only: the cache is filled once
// per sweep). One kHotThreads workgroup per CU runs the whole sweep. It first
// copies contributions [0, min(n, kHotCap)) into LDS; in the degree-ordered
// numbering those are the most gathered ones (on a skewed RMAT graph the top
// 30K of 8.4M vertices are the sources of ~45% of the arcs), so those gathers
// never enter the texture / L1 / L2 path. The work is wave-sized: row blocks
// of <= kWaveNnz in-list entries and long-row chunks of kWaveNnz entries,
// taken by the waves in turn, so no wave waits for another inside the sweep.
constexpr uint32_t kHotThreads = 1024;
constexpr uint32_t kHotWaves = kHotThreads / kWave;
constexpr uint32_t kWaveNnz = 512;
constexpr uint32_t kHotCap = 30720;  // 120 KiB of LDS, + 32 KiB of per-wave staging

// Orders a wave's LDS writes before its later LDS reads by other lanes (a
// wave's LDS operations execute in order; this keeps the compiler's too).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// gather_batched for one wave, with contributions below `hot` read from LDS.
template <class Sink>
__device__ __forceinline__ void gather_wave(const PrSweep& P, const float* s_hot, uint32_t hot, uint32_t x0,
                                            uint32_t cnt, Sink sink) {
    const uint32_t lane = lane_id();
    for (uint32_t base = 0; base < cnt; base += kGatherBatch * kWave) {
        uint32_t u[kGatherBatch];
#pragma unroll
        for (uint32_t j = 0; j < kGatherBatch; ++j) {
            const uint32_t i = base + j * kWave + lane;
            u[j] = i < cnt ? __builtin_nontemporal_load(&P.ci[x0 + i]) : 0u;
        }
        float v[kGatherBatch];
#pragma unroll
        for (uint32_t j = 0; j < kGatherBatch; ++j) {
            const uint32_t i = base + j * kWave + lane;
            float x = 0.0f;
            if (i < cnt) {
                if (u[j] < hot) x = s_hot[u[j]];
                else x = P.contrib_in[u[j]];
            }
            v[j] = x;
        }
#pragma unroll
        for (uint32_t j = 0; j < kGatherBatch; ++j) {
            const uint32_t i = base + j * kWave + lane;
            if (i < cnt) sink(i, v[j]);
        }
    }
}

// A split row's head chunk: u16 arc ids hci16[x0, x0 + cnt) (x0 even, the
// array padded to an even length), two per 32-bit load, every source read from
// the LDS copy (the clamp only keeps a read in bounds).
__device__ __forceinline__ void gather_head16(const uint16_t* hci16, const float* s_hot, uint32_t hot, uint32_t x0,
                                              uint32_t cnt, Kahan& k) {
    const uint32_t lane = lane_id();
    const uint32_t* w = reinterpret_cast<const uint32_t*>(hci16 + x0);
    const uint32_t pairs = (cnt + 1) / 2;
    for (uint32_t base = 0; base < pairs; base += kGatherBatch * kWave) {
        uint32_t u[kGatherBatch];
#pragma unroll
        for (uint32_t j = 0; j < kGatherBatch; ++j) {
            const uint32_t i = base + j * kWave + lane;
            u[j] = i < pairs ? __builtin_nontemporal_load(&w[i]) : 0u;
        }
#pragma unroll
        for (uint32_t j = 0; j < kGatherBatch; ++j) {
            const uint32_t i = base + j * kWave + lane;
            if (i < pairs) {
                k.add(s_hot[min(u[j] & 0xFFFFu, hot - 1)]);
                if (2 * i + 1 < cnt) k.add(s_hot[min(u[j] >> 16, hot - 1)]);
            }
        }
    }
}

__device__ __forceinline__ void pr_hot_body(const PrSweep& P, float add, float d, float inv_n, PrPart& part) {
    __shared__ float s_hot[kHotCap];
    __shared__ float s_stage[kHotWaves][kWaveNnz];
    const PrStreamLayout& Lyt = P.stream;
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
    const uint32_t hot = P.n < kHotCap ? P.n : kHotCap;
    for (uint32_t i = threadIdx.x; i < hot; i += blockDim.x) s_hot[i] = P.contrib_in[i];
    __syncthreads();
    float* st = s_stage[w];
    const uint32_t n_items = Lyt.n_litems + Lyt.n_sitems;
    for (uint32_t it = blockIdx.x * kHotWaves + w; it < n_items; it += gridDim.x * kHotWaves) {
        if (it < Lyt.n_litems) {
            const uint4 li = Lyt.litems[it];
            const uint32_t chunks = Lyt.lrows[li.w & ~kLiHead16].y;  // loaded beside the gathers, not after the sum
            Kahan k;
            if (li.w & kLiHead16) gather_head16(Lyt.hci16, s_hot, hot, li.y, li.z - li.y, k);
            else gather_wave(P, s_hot, hot, li.y, li.z - li.y, [&](uint32_t, float x) { k.add(x); });
            float hi = k.s, lo = -k.c;
            comp_reduce(hi, lo, kWave);
            if (lane == 0) {
                if (chunks == 1) pr_store(P, li.x, add + d * (hi + lo), part, inv_n);
                else Lyt.lpart[it] = make_float2(hi, lo);  // combined by pr_longfix_body
            }
            continue;
        }
        const uint4 si = Lyt.sitems[it - Lyt.n_litems];
        const uint32_t r0 = si.x, r1 = si.y;  // r1 > r0
        const uint32_t rows = r1 - r0;
        uint32_t G = rows >= kWave ? 1u : kWave / rows;
        G = G >= 64 ? 64 : (G >= 32 ? 32 : (G >= 16 ? 16 : (G >= 8 ? 8 : (G >= 4 ? 4 : (G >= 2 ? 2 : 1)))));
        const uint32_t per_round = kWave / G, rr = lane / G, gl = lane % G;
        // a row's bounds, out-degree and previous rank are loaded one round ahead (the
        // first round's beside the gathers; rows clamped to the block's last, loads
        // unconditional), so a round of row sums waits on no load of its own: 981 ->
        // 965 us per HOT sweep and 806 -> 791 us per HOT_TAIL sweep on rmat23
        uint32_t rc = min(r0 + rr, r1 - 1);
        uint32_t ra = P.rp[rc], rz = P.rp[rc + 1], od = P.outdeg[rc];
        float old = P.r_old ? P.r_old[rc] : 0.0f;
        const uint32_t x0 = si.z, x1 = si.w;  // rp[r0], rp[r1], in the item: no load round
        gather_wave(P, s_hot, hot, x0, x1 - x0, [&](uint32_t i, float x) { st[i] = x; });
        wave_lds_sync();
        for (uint32_t rb = 0; rb < rows; rb += per_round) {
            const uint32_t r = r0 + rb + rr;
            const bool ok = r < r1;
            const uint32_t ca = ra, cz = rz, cod = od;
            const float cold = old;
            rc = min(r + per_round, r1 - 1);
            ra = P.rp[rc];
            rz = P.rp[rc + 1];
            od = P.outdeg[rc];
            old = P.r_old ? P.r_old[rc] : 0.0f;
            Kahan k;
            if (ok)
                for (uint32_t x = ca - x0 + gl; x < cz - x0; x += G) k.add(st[x]);
            float hi = k.s, lo = -k.c;
            comp_reduce(hi, lo, G);
            if (ok && gl == 0) pr_store_pre(P, r, add + d * (hi + lo), cod, cold, part, inv_n);
        }
        wave_lds_sync();
    }
}

// Block-reduces the per-thread FP64 partials and stores them at slot `index`
// (quantity q at partial[q * cap + index]). No fences and no counters: a
// later kernel in the same stream reduces the slots.
__device__ __forceinline__ void write_block_partials(const PrPart& part, double* partial, uint32_t cap,
                                                     uint32_t index) {
    __shared__ double red[4][16];
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave, nwv = blockDim.x / kWave;
    const double a = wave_sum(part.dang), b = wave_sum(part.delta), c = wave_sum(part.sum), m = wave_max(part.maxrel);
    __syncthreads();
    if (lane == 0) {
        red[0][w] = a;
        red[1][w] = b;
        red[2][w] = c;
        red[3][w] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
        for (uint32_t j = 0; j < nwv; ++j) {
            t0 += red[0][j];
            t1 += red[1][j];
            t2 += red[2][j];
            t3 = red[3][j] > t3 ? red[3][j] : t3;
        }
        partial[index] = t0;
        partial[cap + index] = t1;
        partial[2 * cap + index] = t2;
        partial[3 * cap + index] = t3;
    }
}

// The sweep body. `add` = (1-d)/n + d*dangling/n, evaluated in FP64 once.
__device__ __forceinline__ void pr_sweep_body(const PrSweep& P) {
    const uint32_t b = blockIdx.x;
    const uint32_t lane = lane_id();
    const double dang_in = P.acc_in->dangling;
    const double addd = (1.0 - (double)P.damping) / (double)P.n + (double)P.damping * dang_in / (double)P.n;
    const float add = (float)addd;
    const float d = P.damping;
    const float inv_n = 1.0f / (float)P.n;
    PrPart part;
    __shared__ float s_hi[16], s_lo[16];

    if (P.use_stream) pr_stream_body(P, P.stream, add, d, inv_n, part);
    for (uint32_t si = 0; si < P.nsec; ++si) {
        const PrSection& S = P.sec[si];
        if (b < S.blk_begin || b >= S.blk_end) continue;
        const uint32_t sb = b - S.blk_begin, sg = S.blk_end - S.blk_begin;
        if (S.mode == SEC_LANE) {
            for (uint64_t i = (uint64_t)sb * blockDim.x + threadIdx.x; i < S.count; i += (uint64_t)sg * blockDim.x) {
                const uint32_t v = S.list ? S.list[i] : S.first + (uint32_t)i;
                const float s = in_sum_lane(P, v);
                pr_store(P, v, add + d * s, part, inv_n);
            }
        } else if (S.mode == SEC_GROUP) {
            const uint32_t L = S.lanes;
            const uint32_t gpb = blockDim.x / L;
            const uint32_t g = threadIdx.x / L, gl = threadIdx.x % L;
            // rounds are uniform per block, so every wave stays convergent
            for (uint64_t base = (uint64_t)sb * gpb; base < S.count; base += (uint64_t)sg * gpb) {
                const uint64_t i = base + g;
                const bool ok = i < S.count;
                const uint32_t v = ok ? (S.list ? S.list[i] : S.first + (uint32_t)i) : 0u;
                Kahan k;
                if (ok) {
                    const uint32_t e0 = P.rp[v], e1 = P.rp[v + 1];
                    for (uint32_t x = e0 + gl; x < e1; x += L) k.add(P.contrib_in[P.ci[x]]);
                }
                float hi = k.s, lo = -k.c;
                comp_reduce(hi, lo, L);
                if (ok && gl == 0) pr_store(P, v, add + d * (hi + lo), part, inv_n);
            }
        } else {  // SEC_BLOCK: one block per vertex
            const uint32_t nwv = blockDim.x / kWave, w = threadIdx.x / kWave;
            for (uint32_t i = sb; i < S.count; i += sg) {
                const uint32_t v = S.list ? S.list[i] : S.first + i;
                Kahan k;
                const uint32_t e0 = P.rp[v], e1 = P.rp[v + 1];
                for (uint32_t x = e0 + threadIdx.x; x < e1; x += blockDim.x) k.add(P.contrib_in[P.ci[x]]);
                float hi = k.s, lo = -k.c;
                comp_reduce(hi, lo, kWave);
                __syncthreads();
                if (lane == 0) {
                    s_hi[w] = hi;
                    s_lo[w] = lo;
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    float H = s_hi[0], Lo = s_lo[0];
                    for (uint32_t j = 1; j < nwv; ++j) {
                        float s, e;
                        two_sum(H, s_hi[j], s, e);
                        H = s;
                        Lo = Lo + s_lo[j] + e;
                    }
                    pr_store(P, v, add + d * (H + Lo), part, inv_n);
                }
            }
        }
    }

    write_block_partials(part, P.partial, P.partial_cap, P.partial_base + b);
}

// The HOT-layout sweep (its own kernel: its LDS footprint is the whole CU's).
__device__ __forceinline__ void pr_hot_sweep_body(const PrSweep& P) {
    const double addd =
        (1.0 - (double)P.damping) / (double)P.n + (double)P.damping * P.acc_in->dangling / (double)P.n;
    PrPart part;
    pr_hot_body(P, (float)addd, P.damping, 1.0f / (float)P.n, part);
    write_block_partials(part, P.partial, P.partial_cap, P.partial_base + blockIdx.x);
}

// Multi-chunk long rows: combine the chunk partials of each row in chunk order
// (wave per row; lane j takes chunks j, j+64, ...; then a TwoSum butterfly) and
// store the row. Runs after the sweep, so the partials are visible.
__device__ __forceinline__ void pr_longfix_body(const PrSweep& P) {
    const PrStreamLayout& Lyt = P.stream;
    const double addd = (1.0 - (double)P.damping) / (double)P.n + (double)P.damping * P.acc_in->dangling / (double)P.n;
    const float add = (float)addd, d = P.damping, inv_n = 1.0f / (float)P.n;
    const uint32_t lane = lane_id(), wpb = blockDim.x / kWave, wid = threadIdx.x / kWave;
    PrPart part;
    for (uint32_t row = blockIdx.x * wpb + wid; row < Lyt.n_lrows; row += gridDim.x * wpb) {
        const uint2 lr = Lyt.lrows[row];
        // single-chunk rows were stored by the sweep; split rows are the tail merge's (wave-uniform)
        if (lr.y < 2 || (lr.y & kLrowSplit)) continue;
        float H = 0.f, L = 0.f;
        for (uint32_t j = lane; j < lr.y; j += kWave) {
            const float2 q = Lyt.lpart[lr.x + j];
            float a, e;
            two_sum(H, q.x, a, e);
            H = a;
            L = L + q.y + e;
        }
        comp_reduce(H, L, kWave);
        if (lane == 0) pr_store(P, Lyt.litems[lr.x].x, add + d * (H + L), part, inv_n);
    }
    write_block_partials(part, P.partial, P.partial_cap, P.partial_base + blockIdx.x);
}

// One block: sums partials [0, count) of each quantity in index order.
__device__ __forceinline__ void pr_reduce_body(const double* partial, uint32_t cap, uint32_t count, PrAcc* out) {
    __shared__ double red[4][16];
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave, nwv = blockDim.x / kWave;
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    for (uint32_t j = threadIdx.x; j < count; j += blockDim.x) {
        t0 += partial[j];
        t1 += partial[cap + j];
        t2 += partial[2 * cap + j];
        const double m3 = partial[3 * cap + j];
        t3 = m3 > t3 ? m3 : t3;
    }
    t0 = wave_sum(t0);
    t1 = wave_sum(t1);
    t2 = wave_sum(t2);
    t3 = wave_max(t3);
    if (lane == 0) {
        red[0][w] = t0;
        red[1][w] = t1;
        red[2][w] = t2;
        red[3][w] = t3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double T0 = 0, T1 = 0, T2 = 0, T3 = 0;
        for (uint32_t j = 0; j < nwv; ++j) {
            T0 += red[0][j];
            T1 += red[1][j];
            T2 += red[2][j];
            T3 = red[3][j] > T3 ? red[3][j] : T3;
        }
        out->dangling = T0;
        out->delta = T1;
        out->sum = T2;
        out->maxrel = T3;
    }
}

}  // namespace ggpu
