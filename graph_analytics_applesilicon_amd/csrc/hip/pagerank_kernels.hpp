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
    PrAcc* acc_out;           // this sweep's totals (set when first, else accumulated)
    double* partial;          // 4 * partial_cap doubles
    uint32_t* ticket;
    uint32_t partial_cap;
    uint32_t n;
    float damping;
    uint32_t first;           // first launch of the sweep: overwrite acc_out
    PrSection sec[kMaxSections];
    uint32_t nsec;
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

__device__ __forceinline__ float in_sum_lane(const PrSweep& P, uint32_t v) {
    Kahan k;
    const uint32_t b = P.rp[v], e = P.rp[v + 1];
    for (uint32_t x = b; x < e; ++x) k.add(P.contrib_in[P.ci[x]]);
    return k.s - k.c;
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

    // ---- block partials, then the last block reduces all partials in order
    __shared__ double red[4][16];
    __shared__ bool last;
    double a = wave_sum(part.dang), bb = wave_sum(part.delta), c = wave_sum(part.sum), mx = wave_max(part.maxrel);
    const uint32_t w = threadIdx.x / kWave, nwv = blockDim.x / kWave;
    if (lane == 0) {
        red[0][w] = a;
        red[1][w] = bb;
        red[2][w] = c;
        red[3][w] = mx;
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
        P.partial[b] = t0;
        P.partial[P.partial_cap + b] = t1;
        P.partial[2 * P.partial_cap + b] = t2;
        P.partial[3 * P.partial_cap + b] = t3;
        __threadfence();
        const uint32_t t = atomicAdd(P.ticket, 1u);
        last = t == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    for (uint32_t j = threadIdx.x; j < gridDim.x; j += blockDim.x) {
        t0 += P.partial[j];
        t1 += P.partial[P.partial_cap + j];
        t2 += P.partial[2 * P.partial_cap + j];
        const double m3 = P.partial[3 * P.partial_cap + j];
        t3 = m3 > t3 ? m3 : t3;
    }
    t0 = wave_sum(t0);
    t1 = wave_sum(t1);
    t2 = wave_sum(t2);
    t3 = wave_max(t3);
    __syncthreads();
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
        PrAcc* o = P.acc_out;
        if (P.first) {
            o->dangling = T0;
            o->delta = T1;
            o->sum = T2;
            o->maxrel = T3;
        } else {
            o->dangling += T0;
            o->delta += T1;
            o->sum += T2;
            o->maxrel = T3 > o->maxrel ? T3 : o->maxrel;
        }
        *P.ticket = 0;  // ready for the next launch
    }
}

}  // namespace ggpu
