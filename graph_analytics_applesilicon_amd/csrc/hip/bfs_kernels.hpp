// BFS level kernels for gfx950 (wave64). Shared by the host-driven level loop,
// the hipGraph level loop and the persistent kernel: every phase is a
// __device__ function over a (block, grid) range so the three drivers run the
// same code.
//
// Reference contracts kept (src/shaders/bfs.metal):
//   stats words   frontier / examined / found / next-edges / 3 bin sizes (:7-21), u64 here
//   claim         relaxed load, then CAS INF -> level+1 (:84-93, :177-181); HIP's CAS is strong
//   edge expand   chunk of tg*ept edges per workgroup, owner = largest i with off[i] <= e (:296-357)
//   binned        lane / wave / workgroup per vertex by degree, discoveries appended to the
//                 next level's bin queues (:191-291)
//   bottom-up     unvisited vertex stops at its first frontier neighbour; one wave assembles
//                 one bitmap word without atomics (:390-424) -- 64 lanes = one u64 word here
#pragma once

#include "dev.hpp"
#include "level_policy.hpp"

namespace ggpu {

enum StatWord : int { S_FRONTIER = 0, S_EXAMINED = 1, S_FOUND = 2, S_NEXT_EDGES = 3, S_BIN0 = 4 };
constexpr int kStatWords = 8;
// Level counters are striped: block b adds its totals (reduced in registers
// and LDS first) to stripe b % kStripes, and readers sum the stripes. One
// device-scope atomic per block per counter, spread over kStripes addresses,
// instead of one per wave on a single address (same-address atomics
// serialize across the XCDs). Queue tails stay single counters but are
// reserved once per block and round.
constexpr int kStripes = 32;
constexpr int kSlotWords = kStripes * kStatWords;
// conv[]: append counters. [0] collect-to-queue, [1..3] collect-to-bins,
// [kTailQueue] edge-expand queue tail, [kTailBin0..+2] binned queue tails.
constexpr int kTailQueue = 4;
constexpr int kTailBin0 = 5;
constexpr int kConvWords = 8;
constexpr int kStatSlots = 4;     // ring of per-level stats slots
constexpr int kOwnerStage = 2048; // LDS-staged owner range of the edge expand
constexpr int kScanItems = 16;    // elements per thread of the tile scan


struct BfsDev {
    const uint32_t* rp;
    const uint32_t* ci;
    uint32_t* dist;
    uint32_t* q[2];
    uint32_t* deg[2];      // degrees next to queue entries; scanned in place into offsets
    uint32_t* bins[2][3];  // binned frontier queues [current/next][small/medium/large]
    uint64_t* bm[2];       // frontier bitmaps (bit v%64 of word v/64)
    uint64_t* visited;     // visited bitmap, valid while consecutive levels are bottom-up
    uint32_t* tile;        // scan tile sums
    unsigned long long* stats;  // kStatSlots x kStatWords
    unsigned long long* conv;   // conversion append counters [4]
    const uint4* heavy;    // bottom-up heavy items {v, first edge, end edge, degree}
    uint32_t n_heavy;
    uint32_t heavy_min;    // degree from which a vertex is scanned by the heavy pass
    uint32_t n;
    uint32_t nw;           // bitmap words
    uint64_t m;
};

// dist only moves from INF to a final value, so a plain (possibly stale) load
// that sees a value is exact; a stale INF just costs a failing CAS.
// Adds per-thread counter values, block-reduced, to the block's stripe.
template <int K>
__device__ __forceinline__ void block_flush(unsigned long long* slot, const int (&word)[K],
                                            const unsigned long long (&val)[K]) {
    __shared__ unsigned long long s_red[K][16];
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave, nwv = blockDim.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const unsigned long long t = wave_sum(val[k]);
        if (lane == 0) s_red[k][w] = t;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        unsigned long long t = 0;
        for (uint32_t i = 0; i < nwv; ++i) t += s_red[threadIdx.x][i];
        if (t) atomicAdd(&slot[(blockIdx.x % kStripes) * kStatWords + word[threadIdx.x]], t);
    }
    __syncthreads();
}

// Block-wide append: every thread of the block calls it (uniform control
// flow); one atomic per block on `counter`. Returns the slot of threads with
// `want` set.
__device__ __forceinline__ unsigned long long block_append(bool want, unsigned long long* counter) {
    __shared__ uint32_t s_cnt[16];
    __shared__ unsigned long long s_base[17];
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave, nwv = blockDim.x / kWave;
    const uint64_t mask = __ballot(want);
    if (lane == 0) s_cnt[w] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (uint32_t i = 0; i < nwv; ++i) {
            s_base[1 + i] = run;
            run += s_cnt[i];
        }
        s_base[0] = run ? atomicAdd(counter, run) : 0ull;
    }
    __syncthreads();
    return s_base[0] + s_base[1 + w] + rank_in(mask);
}

__device__ __forceinline__ bool try_claim(uint32_t* dist, uint32_t u, uint32_t value) {
    if (dist[u] != GGPU_INF) return false;
    return atomicCAS(&dist[u], GGPU_INF, value) == GGPU_INF;
}

// Largest i in [lo, hi] with off[i] <= e.
__device__ __forceinline__ uint32_t find_owner(const uint32_t* off, uint32_t lo, uint32_t hi, uint32_t e) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------- init
__device__ __forceinline__ void ph_init(const BfsDev& d, uint32_t src, uint32_t block, uint32_t grid) {
    const uint64_t stride = (uint64_t)grid * blockDim.x;
    for (uint64_t v = (uint64_t)block * blockDim.x + threadIdx.x; v < d.n; v += stride)
        d.dist[v] = v == src ? 0u : GGPU_INF;
}

// ---------------------------------------------------------------- scan
// In-place exclusive scan of vals[0..size) in tiles of blockDim*kScanItems:
// pass 1 writes tile sums, pass 2 adds the sum of all earlier tiles.
__device__ __forceinline__ void ph_scan_tiles(uint32_t* vals, uint32_t size, uint32_t* tile, uint32_t block,
                                              uint32_t grid) {
    __shared__ uint32_t red[16];
    const uint32_t span = blockDim.x * kScanItems;
    const uint32_t ntiles = (size + span - 1) / span;
    for (uint32_t t = block; t < ntiles; t += grid) {
        const uint32_t base = t * span;
        uint32_t s = 0;
        for (uint32_t j = threadIdx.x; j < span; j += blockDim.x) {
            const uint32_t i = base + j;
            if (i < size) s += vals[i];
        }
        s = block_sum(s, red);
        if (threadIdx.x == 0) tile[t] = s;
    }
}

__device__ __forceinline__ void ph_scan_apply(uint32_t* vals, uint32_t size, const uint32_t* tile, uint32_t block,
                                              uint32_t grid) {
    __shared__ uint32_t red[16];
    __shared__ uint32_t wave_tot[16];
    const uint32_t span = blockDim.x * kScanItems;
    const uint32_t ntiles = (size + span - 1) / span;
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    for (uint32_t t = block; t < ntiles; t += grid) {
        // carry-in: sum of the tile sums before t
        uint32_t c = 0;
        for (uint32_t j = threadIdx.x; j < t; j += blockDim.x) c += tile[j];
        const uint32_t carry = block_sum(c, red);
        // this thread's kScanItems consecutive elements
        const uint32_t first = t * span + threadIdx.x * kScanItems;
        uint32_t x[kScanItems];
        uint32_t own = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const uint32_t i = first + k;
            x[k] = i < size ? vals[i] : 0u;
            own += x[k];
        }
        const uint32_t incl = wave_inclusive_scan(own);
        if (lane == kWave - 1) wave_tot[w] = incl;
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t i = 0; i < w; ++i) before += wave_tot[i];
        uint32_t run = carry + before + incl - own;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const uint32_t i = first + k;
            if (i < size) vals[i] = run;
            run += x[k];
        }
        __syncthreads();  // wave_tot reuse
        (void)nw;
    }
}

// -------------------------------------------------- edge-balanced top-down
// Expands `total` frontier edges: workgroup chunk c covers edges
// [c*tg*ept, (c+1)*tg*ept); thread j of the chunk takes edges j, j+tg, ...
// (coalesced col_idx). The chunk's owner range and the owners' list starts
// are staged in LDS, so each edge resolves its owner by an LDS binary search.
__device__ __forceinline__ void ph_td_edge(const BfsDev& d, uint32_t cur, uint32_t level, uint32_t size, uint32_t total,
                                           uint32_t ept, unsigned long long* stats, uint32_t block, uint32_t grid) {
    __shared__ uint32_t s_off[kOwnerStage];
    __shared__ uint32_t s_base[kOwnerStage];
    __shared__ uint32_t s_range[2];
    const uint32_t* q = d.q[cur];
    const uint32_t* off = d.deg[cur];
    uint32_t* nq = d.q[cur ^ 1];
    uint32_t* ndeg = d.deg[cur ^ 1];
    const uint32_t chunk = blockDim.x * ept;
    unsigned long long next_edges = 0;
    for (uint64_t c0l = (uint64_t)block * chunk; c0l < total; c0l += (uint64_t)grid * chunk) {
        const uint32_t c0 = (uint32_t)c0l;
        const uint32_t c1 = (uint32_t)min<uint64_t>(c0l + chunk, total);
        if (threadIdx.x == 0) {
            const uint32_t lo = find_owner(off, 0, size - 1, c0);
            s_range[0] = lo;
            s_range[1] = find_owner(off, lo, size - 1, c1 - 1);
        }
        __syncthreads();
        const uint32_t lo = s_range[0], hi = s_range[1];
        const uint32_t span = hi - lo + 1;
        const bool staged = span <= (uint32_t)kOwnerStage;
        if (staged)
            for (uint32_t j = threadIdx.x; j < span; j += blockDim.x) {
                s_off[j] = off[lo + j];
                s_base[j] = d.rp[q[lo + j]];
            }
        __syncthreads();
        for (uint32_t k = 0; k < ept; ++k) {
            const uint32_t e = c0 + k * blockDim.x + threadIdx.x;
            const bool has = e < c1;
            uint32_t u = 0;
            if (has) {
                if (staged) {
                    uint32_t a = 0, b = span - 1;
                    while (a < b) {
                        const uint32_t mid = (a + b + 1) >> 1;
                        if (s_off[mid] <= e) a = mid;
                        else b = mid - 1;
                    }
                    u = d.ci[s_base[a] + (e - s_off[a])];
                } else {
                    const uint32_t i = find_owner(off, lo, hi, e);
                    u = d.ci[d.rp[q[i]] + (e - off[i])];
                }
            }
            const bool claimed = has && try_claim(d.dist, u, level + 1);
            const uint32_t udeg = claimed ? d.rp[u + 1] - d.rp[u] : 0u;
            const unsigned long long pos = block_append(claimed, &d.conv[kTailQueue]);
            if (claimed) {
                nq[pos] = u;
                ndeg[pos] = udeg;
            }
            next_edges += udeg;
        }
        __syncthreads();  // LDS reuse by the next chunk
    }
    const int words[1] = {S_NEXT_EDGES};
    const unsigned long long vals[1] = {next_edges};
    block_flush<1>(stats, words, vals);
}

// ---------------------------------------------------------- binned top-down
__device__ __forceinline__ void bin_visit(const BfsDev& d, uint32_t nxt, bool has, uint32_t e, uint32_t level,
                                          uint32_t small_max, uint32_t large_min, unsigned long long& next_edges) {
    const uint32_t u = has ? d.ci[e] : 0u;
    const bool claimed = has && try_claim(d.dist, u, level + 1);
    if (__ballot(claimed) == 0) return;  // wave-uniform
    const uint32_t udeg = claimed ? d.rp[u + 1] - d.rp[u] : 0u;
    const uint32_t b = udeg < small_max ? 0u : (udeg < large_min ? 1u : 2u);
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k) {
        const bool mine = claimed && b == k;
        const unsigned long long pos = wave_append(mine, &d.conv[kTailBin0 + k]);
        if (mine) d.bins[nxt][k][pos] = u;
    }
    next_edges += udeg;
}

// Bin counts cnt[3]. All blocks stride over the large bin (workgroup per
// vertex), then the medium bin (wave per vertex), then the small bin (lane
// per vertex, the wave expanding its 64 short lists together).
__device__ __forceinline__ void ph_td_binned(const BfsDev& d, uint32_t cur, uint32_t level, uint32_t small_max,
                                             uint32_t large_min, uint32_t cnt0, uint32_t cnt1, uint32_t cnt2,
                                             unsigned long long* stats, uint32_t block, uint32_t grid) {
    const uint32_t nxt = cur ^ 1;
    const uint32_t lane = lane_id();
    unsigned long long next_edges = 0;
    for (uint32_t i = block; i < cnt2; i += grid) {
        const uint32_t v = d.bins[cur][2][i];
        const uint32_t b = d.rp[v], e = d.rp[v + 1];
        for (uint32_t base = b; base < e; base += blockDim.x)
            bin_visit(d, nxt, base + threadIdx.x < e, base + threadIdx.x, level, small_max, large_min, next_edges);
    }
    const uint32_t wpb = blockDim.x / kWave, wid = threadIdx.x / kWave;
    for (uint32_t i = block * wpb + wid; i < cnt1; i += grid * wpb) {
        const uint32_t v = d.bins[cur][1][i];
        const uint32_t b = d.rp[v], e = d.rp[v + 1];
        for (uint32_t base = b; base < e; base += kWave)
            bin_visit(d, nxt, base + lane < e, base + lane, level, small_max, large_min, next_edges);
    }
    for (uint32_t w0 = (block * wpb + wid) * kWave; w0 < cnt0; w0 += grid * blockDim.x) {
        const uint32_t idx = w0 + lane;
        const bool active = idx < cnt0;
        const uint32_t v = active ? d.bins[cur][0][idx] : 0u;
        const uint32_t begin = active ? d.rp[v] : 0u;
        const uint32_t dg = active ? d.rp[v + 1] - begin : 0u;
        const uint32_t incl = wave_inclusive_scan(dg);
        const uint32_t pre = incl - dg;
        const uint32_t tot = __shfl(incl, kWave - 1, kWave);
        for (uint32_t base = 0; base < tot; base += kWave) {
            const uint32_t j = base + lane;
            // owner = largest lane l with pre[l] <= j (6 shuffle steps for 64 lanes)
            uint32_t owner = 0;
#pragma unroll
            for (uint32_t step = 32; step > 0; step >>= 1) {
                const uint32_t cand = owner + step;
                if (__shfl(pre, (int)cand, kWave) <= j) owner = cand;
            }
            const uint32_t ob = __shfl(begin, (int)owner, kWave);
            const uint32_t op = __shfl(pre, (int)owner, kWave);
            bin_visit(d, nxt, j < tot, ob + (j - op), level, small_max, large_min, next_edges);
        }
    }
    const int words[1] = {S_NEXT_EDGES};
    const unsigned long long vals[1] = {next_edges};
    block_flush<1>(stats, words, vals);
}

// ---------------------------------------------------------------- bottom-up
// Light pass: lane v%64 of the wave owning word v/64, for vertices of degree
// < heavy_min (a short serial scan). `visited` marks vertices with a final
// distance; the wave writes next[w] and visited[w] whole (no atomics).
// Heavier vertices are left to ph_bu_heavy, which runs after this pass.
__device__ __forceinline__ void ph_bu(const BfsDev& d, const uint64_t* front, uint64_t* next, uint32_t level,
                                      unsigned long long* stats, uint32_t block, uint32_t grid) {
    const uint32_t lane = lane_id();
    const uint32_t wpb = blockDim.x / kWave, wid = threadIdx.x / kWave;
    unsigned long long t_found = 0, t_examined = 0, t_next = 0;
    for (uint32_t w = block * wpb + wid; w < d.nw; w += grid * wpb) {
        const uint64_t vis = d.visited[w];
        const uint32_t v = w * kWave + lane;
        bool found = false;
        uint32_t examined = 0, ndeg = 0;
        if (vis != ~0ull && v < d.n && !((vis >> lane) & 1ull)) {
            const uint32_t b = d.rp[v], e = d.rp[v + 1];
            if (e - b < d.heavy_min) {
                for (uint32_t x = b; x < e; ++x) {
                    ++examined;
                    const uint32_t u = d.ci[x];
                    if ((front[u >> 6] >> (u & 63)) & 1ull) {
                        found = true;
                        break;
                    }
                }
                if (found) {
                    d.dist[v] = level + 1;
                    ndeg = e - b;
                }
            }
        }
        const uint64_t mask = __ballot(found);
        if (lane == 0) {
            next[w] = mask;
            if (mask) d.visited[w] = vis | mask;
        }
        t_found += found;
        t_examined += examined;
        t_next += ndeg;
    }
    const int words[3] = {S_FOUND, S_EXAMINED, S_NEXT_EDGES};
    const unsigned long long vals[3] = {t_found, t_examined, t_next};
    block_flush<3>(stats, words, vals);
}

// Heavy pass: one wave per item (a vertex of degree >= heavy_min, or one
// <= 4096-edge chunk of a longer list). The wave checks 64 neighbours per round
// and stops at the first round with a frontier hit (ballot). The first chunk
// to set the vertex's bit in `next` (atomicOr) records it.
__device__ __forceinline__ void ph_bu_heavy(const BfsDev& d, const uint64_t* front, uint64_t* next, uint32_t level,
                                            unsigned long long* stats, uint32_t block, uint32_t grid) {
    const uint32_t lane = lane_id();
    const uint32_t wpb = blockDim.x / kWave, wid = threadIdx.x / kWave;
    unsigned long long t_found = 0, t_examined = 0, t_next = 0;
    for (uint32_t it = block * wpb + wid; it < d.n_heavy; it += grid * wpb) {
        const uint4 item = d.heavy[it];
        const uint32_t v = item.x, x0 = item.y, x1 = item.z;
        const uint64_t bit = 1ull << (v & 63);
        if (d.visited[v >> 6] & bit) continue;  // wave-uniform
        uint32_t examined = 0;
        bool hit_any = false;
        for (uint32_t base = x0; base < x1; base += kWave) {
            const uint32_t x = base + lane;
            bool hit = false;
            if (x < x1) {
                const uint32_t u = d.ci[x];
                hit = (front[u >> 6] >> (u & 63)) & 1ull;
            }
            const uint64_t m = __ballot(hit);
            if (m) {
                examined += (uint32_t)__ffsll((long long)m);
                hit_any = true;
                break;
            }
            examined += min(kWave, (int)(x1 - base));
        }
        if (lane == 0) {
            if (hit_any) {
                const uint64_t old = atomicOr((unsigned long long*)&next[v >> 6], (unsigned long long)bit);
                if (!(old & bit)) {
                    d.dist[v] = level + 1;
                    atomicOr((unsigned long long*)&d.visited[v >> 6], (unsigned long long)bit);
                    t_found += 1;
                    t_next += item.w;
                }
            }
            t_examined += examined;
        }
    }
    const int words[3] = {S_FOUND, S_EXAMINED, S_NEXT_EDGES};
    const unsigned long long vals[3] = {t_found, t_examined, t_next};
    block_flush<3>(stats, words, vals);
}

// ------------------------------------------------------------ conversions
// dist -> frontier bitmap (dist == level) and visited bitmap (dist != INF).
__device__ __forceinline__ void ph_dist_to_bitmap(const BfsDev& d, uint64_t* front, uint32_t level, uint32_t block,
                                                  uint32_t grid) {
    const uint32_t lane = lane_id();
    const uint32_t wpb = blockDim.x / kWave, wid = threadIdx.x / kWave;
    for (uint32_t w = block * wpb + wid; w < d.nw; w += grid * wpb) {
        const uint32_t v = w * kWave + lane;
        const uint32_t dv = v < d.n ? d.dist[v] : GGPU_INF;
        const uint64_t f = __ballot(dv == level);
        const uint64_t s = __ballot(dv != GGPU_INF);
        if (lane == 0) {
            front[w] = f;
            d.visited[w] = s;
        }
    }
}

// bitmap -> queue (+ degrees); lane per word, one atomic per wave.
__device__ __forceinline__ void ph_bitmap_to_queue(const BfsDev& d, const uint64_t* bm, uint32_t* q, uint32_t* deg,
                                                   unsigned long long* counter, uint32_t block, uint32_t grid) {
    const uint64_t stride = (uint64_t)grid * blockDim.x;
    for (uint64_t w0 = (uint64_t)block * blockDim.x + (threadIdx.x & ~(kWave - 1)); w0 < d.nw; w0 += stride) {
        const uint64_t w = w0 + lane_id();
        uint64_t word = w < d.nw ? bm[w] : 0ull;
        const uint32_t c = (uint32_t)__popcll(word);
        const uint32_t incl = wave_inclusive_scan(c);
        const uint32_t tot = __shfl(incl, kWave - 1, kWave);
        unsigned long long base = 0;
        if (lane_id() == 0 && tot) base = atomicAdd(counter, (unsigned long long)tot);
        base = __shfl(base, 0, kWave);
        unsigned long long at = base + incl - c;
        while (word) {
            const uint32_t v = (uint32_t)(w * kWave + (uint64_t)__builtin_ctzll(word));
            word &= word - 1;
            q[at] = v;
            deg[at] = d.rp[v + 1] - d.rp[v];
            ++at;
        }
    }
}

// dist == level -> queue (+ degrees), or -> degree bins when to_bins.
__device__ __forceinline__ void ph_dist_collect(const BfsDev& d, uint32_t level, bool to_bins, uint32_t dst,
                                                uint32_t small_max, uint32_t large_min, uint32_t block, uint32_t grid) {
    const uint64_t stride = (uint64_t)grid * blockDim.x;
    for (uint64_t v0 = (uint64_t)block * blockDim.x + (threadIdx.x & ~(kWave - 1)); v0 < d.n; v0 += stride) {
        const uint64_t v = v0 + lane_id();
        const bool f = v < d.n && d.dist[v] == level;
        const uint32_t dg = f ? d.rp[v + 1] - d.rp[v] : 0u;
        if (!to_bins) {
            const unsigned long long pos = wave_append(f, &d.conv[0]);
            if (f) {
                d.q[dst][pos] = (uint32_t)v;
                d.deg[dst][pos] = dg;
            }
        } else {
            const uint32_t b = dg < small_max ? 0u : (dg < large_min ? 1u : 2u);
#pragma unroll
            for (uint32_t k = 0; k < 3; ++k) {
                const bool mine = f && b == k;
                const unsigned long long pos = wave_append(mine, &d.conv[1 + k]);
                if (mine) d.bins[dst][k][pos] = (uint32_t)v;
            }
        }
    }
}

}  // namespace ggpu
