// Personalized PageRank as an n x B FP32 block (no reference counterpart).
//
// Column b restarts at seeds[b]:
//   r_new[v][b] = d * sum_{u in in(v)} r[u][b] / outdeg[u]
//               + [v == seeds[b]] * ((1 - d) + d * dangling_b)
// with dangling_b the column's rank on dangling vertices (FP64), so every
// column keeps total mass 1 (networkx.pagerank with personalization {s: 1}).
//
// Layout: vertex-major, row v = B consecutive floats, so every in-edge is one
// row gather of 4*B bytes that serves all columns, and the in-lists are read
// once per sweep for the whole block. Columns run in passes of Bw lanes,
// Bw = min(64, B rounded up to a power of two).
//
// Work assignment (large_min and medium_lanes of the options; ppr_autotune
// picks them, and the workgroup size, by timing):
//  * heavy vertices, in-degree >= large_min: the in-list is cut into chunks of
//    large_min entries and k_ppr_chunks gives every chunk a wave, which writes
//    per-column (hi, lo) partial sums; the sweep adds a heavy vertex's chunk
//    partials in chunk order, so no vertex's list is walked by a single wave.
//  * light vertices: L = medium_lanes lanes per vertex (clamped to [Bw, 64]);
//    lane (j, b) of the group sums column b over the entries j, j + L/Bw, ...
//    and the group's sums merge by a butterfly; 64/L vertices per wave.
// Per-lane Kahan sums, compensated (hi, lo) merges; FP64 per-column dangling
// mass and L1 change through per-workgroup partials added in workgroup order,
// so a run is deterministic for a given assignment.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <limits>

#include "graph.hpp"
#include "pagerank_kernels.hpp"
#include "pr_shared.hpp"

namespace ggpu {

constexpr uint32_t kPprMaxB = 256;
constexpr uint32_t kPprBlocks = 2048;  // workgroup partial slots of the sweep
constexpr uint32_t kPprNoHeavy = std::numeric_limits<uint32_t>::max();

struct PprArgs {
    const uint32_t* rp;
    const uint32_t* ci;
    const uint32_t* od;
    const uint32_t* seeds;
    const float* c_in;
    float* c_out;
    float* r_new;
    const float* r_old;    // null: no L1 change
    const double* dang_in; // B
    double* partial;       // (B + 1) * kPprBlocks
    const uint2* chunks;
    const uint2* heavy;
    float2* cpart;         // n_chunks * B
    uint32_t n, B, Bw, L;
    uint32_t heavy_min;    // in-degree at which a vertex is heavy
    uint32_t n_heavy, n_groups, n_chunks;
    float damping;
};

__global__ void k_ppr_init(float* r, float* c, const uint32_t* od, const uint32_t* seeds, uint32_t n, uint32_t B) {
    const uint64_t total = (uint64_t)n * B;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)(i / B), b = (uint32_t)(i % B);
        const bool on = seeds[b] == v;
        r[i] = on ? 1.0f : 0.0f;
        const uint32_t d = od[v];
        c[i] = (on && d) ? 1.0f / (float)d : 0.0f;
    }
}

// (hi, lo) += (x_hi, x_lo): TwoSum of the leading parts, errors folded into
// the trailing part, renormalised. Symmetric in its operands, so both lanes of
// a butterfly pair end with the same pair.
__device__ __forceinline__ void hl_add(float& hi, float& lo, float x_hi, float x_lo) {
    float s, e;
    two_sum(hi, x_hi, s, e);
    e += lo + x_lo;
    hi = s + e;
    lo = e - (hi - s);
}

// Kahan sum of c[u * B + b] over the list entries x = x0, x0 + E, ... < x1,
// U row gathers in flight.
template <int U>
__device__ __forceinline__ void ppr_gather(const uint32_t* __restrict__ ci, const float* __restrict__ c, uint32_t B,
                                           uint32_t x0, uint32_t x1, uint32_t E, uint32_t b, Kahan& k) {
    uint32_t x = x0;
    for (; x + (U - 1) * E < x1; x += U * E) {
        uint32_t u[U];
#pragma unroll
        for (int i = 0; i < U; ++i) u[i] = ci[x + i * E];
        float a[U];
#pragma unroll
        for (int i = 0; i < U; ++i) a[i] = c[(size_t)u[i] * B + b];
#pragma unroll
        for (int i = 0; i < U; ++i) k.add(a[i]);
    }
    for (; x < x1; x += E) k.add(c[(size_t)ci[x] * B + b]);
}

// Merges the (hi, lo) sums of the lanes that differ in bits [Bw, L) of the lane id.
__device__ __forceinline__ void ppr_merge_group(float& hi, float& lo, uint32_t Bw, uint32_t L) {
    for (uint32_t off = Bw; off < L; off <<= 1) {
        const float oh = __shfl_xor(hi, (int)off, kWave);
        const float ol = __shfl_xor(lo, (int)off, kWave);
        hl_add(hi, lo, oh, ol);
    }
}

// One wave per chunk of a heavy in-list: per-column partial sums. NP column
// passes (NP = ceil(B / 64) for blocks wider than 64, else 1).
template <int NP>
__global__ void k_ppr_chunks(PprArgs A) {
    const uint32_t lane = lane_id();
    const uint32_t wpb = blockDim.x / kWave;
    const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint32_t Bw = A.Bw, E = kWave / Bw;
    const uint32_t j = lane / Bw, bl = lane % Bw;
    for (uint32_t c = blockIdx.x * wpb + wid; c < A.n_chunks; c += gridDim.x * wpb) {
        const uint2 r = A.chunks[c];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const uint32_t b = p * Bw + bl;
            Kahan k;
            ppr_gather<8>(A.ci, A.c_in, A.B, r.x + j, r.y, E, std::min(b, A.B - 1), k);
            float hi = k.s, lo = -k.c;
            ppr_merge_group(hi, lo, Bw, kWave);
            if (j == 0 && b < A.B) A.cpart[(size_t)c * A.B + b] = make_float2(hi, lo);
        }
    }
}

template <int NP>
__global__ void k_ppr_sweep(PprArgs A) {
    const uint32_t lane = lane_id();
    const uint32_t wpb = blockDim.x / kWave;
    const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint32_t Bw = A.Bw, L = A.L, E = L / Bw, G = kWave / L;
    const uint32_t g = lane / L, j = (lane % L) / Bw, bl = lane % Bw;
    const float d = A.damping;
    double dang[NP];
    double delta = 0;
    float restart[NP];
    uint32_t seed[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        dang[p] = 0;
        const uint32_t b = p * Bw + bl;
        seed[p] = b < A.B ? A.seeds[b] : GGPU_INF;
        restart[p] = b < A.B ? (float)((1.0 - (double)d) + (double)d * A.dang_in[b]) : 0.0f;
    }
    // the new rank of column b at v from its in-sum
    auto finish = [&](uint32_t v, int p, uint32_t b, uint32_t odv, float sum) {
        float val = d * sum;
        if (seed[p] == v) val += restart[p];
        const size_t at = (size_t)v * A.B + b;
        A.r_new[at] = val;
        A.c_out[at] = odv ? val / (float)odv : 0.0f;
        if (!odv) dang[p] += (double)val;
        if (A.r_old) delta += (double)fabsf(val - A.r_old[at]);
    };
    const uint32_t items = A.n_heavy + A.n_groups;
    for (uint32_t it = blockIdx.x * wpb + wid; it < items; it += gridDim.x * wpb) {
        if (it < A.n_heavy) {
            // heavy vertex: its chunk partials in chunk order, lane b = column
            const uint2 h = A.heavy[it];
            const uint32_t c1 = A.heavy[it + 1].y;
            const uint32_t v = h.x, odv = A.od[v];
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                const uint32_t b = p * Bw + lane;
                if (lane < Bw && b < A.B) {
                    float hi = 0.f, lo = 0.f;
                    for (uint32_t c = h.y; c < c1; ++c) {
                        const float2 q = A.cpart[(size_t)c * A.B + b];
                        hl_add(hi, lo, q.x, q.y);
                    }
                    finish(v, p, b, odv, hi + lo);
                }
            }
            continue;
        }
        const uint32_t v = (it - A.n_heavy) * G + g;
        uint32_t x0 = 0, x1 = 0, odv = 0;
        bool live = v < A.n;
        if (live) {
            x0 = A.rp[v];
            x1 = A.rp[v + 1];
            odv = A.od[v];
            live = x1 - x0 < A.heavy_min;
        }
        if (!live) x1 = x0;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const uint32_t b = p * Bw + bl;
            Kahan k;
            ppr_gather<8>(A.ci, A.c_in, A.B, x0 + j, x1, E, std::min(b, A.B - 1), k);
            float hi = k.s, lo = -k.c;
            ppr_merge_group(hi, lo, Bw, L);
            if (live && j == 0 && b < A.B) finish(v, p, b, odv, hi + lo);
        }
    }
    // workgroup partials: per column dangling mass (lanes of one column position
    // and all waves, in fixed order), then the L1 change
    __shared__ double s_d[16][kWave];
    __shared__ double s_x[16];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        __syncthreads();
        s_d[wid][lane] = dang[p];
        __syncthreads();
        if (wid == 0 && lane < Bw) {
            double t = 0;
            for (uint32_t w = 0; w < wpb; ++w)
                for (uint32_t l = lane; l < kWave; l += Bw) t += s_d[w][l];
            const uint32_t b = p * Bw + lane;
            if (b < A.B) A.partial[(size_t)blockIdx.x * (A.B + 1) + b] = t;
        }
    }
    const double dx = wave_sum(delta);
    __syncthreads();
    if (lane == 0) s_x[wid] = dx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (uint32_t w = 0; w < wpb; ++w) t += s_x[w];
        A.partial[(size_t)blockIdx.x * (A.B + 1) + A.B] = t;
    }
}

// Rows back to the caller's numbering: dst row v = src row id_of[v].
__global__ void k_ppr_unpermute(const float* src, float* dst, const uint32_t* id_of, uint32_t n, uint32_t B) {
    const uint64_t total = (uint64_t)n * B;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)(i / B), b = (uint32_t)(i % B);
        dst[i] = src[(size_t)id_of[v] * B + b];
    }
}

// Column sums of the sweep's workgroup partials: a wave per column (B + 1
// columns: dangling mass per seed, then the L1 change), lane j adding
// workgroups j, j + 64, ... and a fixed butterfly combining the lanes.
__global__ void k_ppr_reduce(const double* partial, uint32_t blocks, uint32_t B, double* dang_out, double* delta_out) {
    const uint32_t col = blockIdx.x, lane = threadIdx.x;
    double t = 0;
    for (uint32_t blk = lane; blk < blocks; blk += kWave) t += partial[(size_t)blk * (B + 1) + col];
    t = wave_sum(t);
    if (lane == 0) {
        if (col < B) dang_out[col] = t;
        else *delta_out = t;
    }
}

}  // namespace ggpu

namespace ggpu {

namespace {

uint32_t pow2_at_least(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Heavy in-lists of the given threshold, cut into chunks (host, once per threshold).
void ppr_build_heavy(PprState& S, const PprView& V, uint32_t heavy_min) {
    if (S.heavy_built && S.heavy_min == heavy_min && S.heavy_space == V.space) return;
    std::vector<uint2> hv, ch;
    if (heavy_min != kPprNoHeavy) {
        for (uint32_t v = 0; v < V.n; ++v) {
            const uint32_t e0 = V.rp_host[v], e1 = V.rp_host[v + 1];
            if (e1 - e0 < heavy_min) continue;
            hv.push_back(make_uint2(v, (uint32_t)ch.size()));
            for (uint32_t e = e0; e < e1; e += std::min(heavy_min, e1 - e))
                ch.push_back(make_uint2(e, e + std::min(heavy_min, e1 - e)));
        }
    }
    hv.push_back(make_uint2(0, (uint32_t)ch.size()));
    S.heavy.alloc(hv.size());
    S.chunks.alloc(std::max<size_t>(ch.size(), 1));
    hipStream_t s = V.ctx->stream;
    GGPU_HIP_TRY(hipMemcpyAsync(S.heavy.p, hv.data(), hv.size() * sizeof(uint2), hipMemcpyHostToDevice, s));
    if (!ch.empty())
        GGPU_HIP_TRY(hipMemcpyAsync(S.chunks.p, ch.data(), ch.size() * sizeof(uint2), hipMemcpyHostToDevice, s));
    GGPU_HIP_TRY(hipStreamSynchronize(s));
    S.n_heavy = (uint32_t)hv.size() - 1;
    S.n_chunks = (uint32_t)ch.size();
    S.heavy_min = heavy_min;
    S.heavy_space = V.space;
    S.heavy_built = true;
}

template <int NP>
void ppr_occupancy_np(PprState& S, uint32_t tg) {
    int a = 0, b = 0;
    GGPU_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_ppr_sweep<NP>, (int)tg, 0));
    GGPU_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_ppr_chunks<NP>, (int)tg, 0));
    S.occ_sweep = std::max(a, 1);
    S.occ_chunks = std::max(b, 1);
}

// Resident workgroups per CU of the two kernels at this shape (cached).
void ppr_occupancy(PprState& S, uint32_t tg, uint32_t np) {
    const int key = (int)(tg * 8 + np);
    if (S.occ_tg == key) return;
    switch (np) {
        case 1: ppr_occupancy_np<1>(S, tg); break;
        case 2: ppr_occupancy_np<2>(S, tg); break;
        case 3: ppr_occupancy_np<3>(S, tg); break;
        default: ppr_occupancy_np<4>(S, tg); break;
    }
    S.occ_tg = key;
}

template <int NP>
void ppr_launch(const PprArgs& A, uint32_t grid, uint32_t cgrid, uint32_t tg, hipStream_t s) {
    if (A.n_chunks) {
        k_ppr_chunks<NP><<<cgrid, tg, 0, s>>>(A);
        GGPU_HIP_TRY(hipGetLastError());
    }
    k_ppr_sweep<NP><<<grid, tg, 0, s>>>(A);
    GGPU_HIP_TRY(hipGetLastError());
}

}  // namespace

ggpu_pr_result ppr_run(PrEngine& e, const uint32_t* seeds_in, uint32_t B, const ggpu_pr_opts& o) {
    PprState& S = ppr_state(e);
    if (B == 0 || B > kPprMaxB) throw ArgError("PPR block width must be in [1, 256]");
    if (o.tg < 64 || o.tg > 1024 || o.tg % 64) throw ArgError("workgroup size must be a multiple of 64 in [64, 1024]");
    if (o.iters < 0) throw ArgError("iters must be >= 0");
    if (o.large_min < 64) throw ArgError("PPR large_min (heavy in-degree and chunk length) must be >= 64");
    if (o.medium_lanes == 0 || o.medium_lanes > 64 || (o.medium_lanes & (o.medium_lanes - 1)))
        throw ArgError("medium_lanes must be a power of two in [1, 64]");
    const uint32_t n = ppr_view(e, GGPU_PR_ORDER_GIVEN).n;
    for (uint32_t b = 0; b < B; ++b)
        if (seeds_in[b] >= n) throw ArgError("seed out of range");
    const PprView V = ppr_view(e, o.order);
    std::vector<uint32_t> seeds(seeds_in, seeds_in + B);
    if (V.id_of)
        for (auto& x : seeds) x = V.id_of[x];
    const size_t cells = (size_t)n * B;
    hipStream_t s = V.ctx->stream;
    const uint32_t Bw = std::min<uint32_t>(pow2_at_least(B), kWave);
    const uint32_t L = std::max(Bw, o.medium_lanes);
    const uint32_t np = (B + Bw - 1) / Bw;
    ppr_build_heavy(S, V, o.large_min);
    ppr_occupancy(S, o.tg, np);
    S.r.alloc(2 * cells);
    S.c.alloc(2 * cells);
    S.part.alloc((size_t)(kPprMaxB + 1) * kPprBlocks);
    S.dang.alloc(2 * kPprMaxB);
    S.delta.alloc(2);
    S.seeds.alloc(kPprMaxB);
    S.cpart.alloc(std::max<size_t>((size_t)S.n_chunks * B, 1));
    while (S.ev.size() < 2) {
        hipEvent_t ev;
        GGPU_HIP_TRY(hipEventCreate(&ev));
        S.ev.push_back(ev);
    }
    S.B = B;
    std::vector<double> dang0(B);
    for (uint32_t b = 0; b < B; ++b) dang0[b] = V.od_host[seeds[b]] == 0 ? 1.0 : 0.0;
    GGPU_HIP_TRY(hipMemcpyAsync(S.seeds.p, seeds.data(), 4ull * B, hipMemcpyHostToDevice, s));
    GGPU_HIP_TRY(hipMemcpyAsync(S.dang.p, dang0.data(), 8ull * B, hipMemcpyHostToDevice, s));
    GGPU_HIP_TRY(hipEventRecord(S.ev[0], s));
    const uint32_t ib = (uint32_t)std::min<uint64_t>((cells + 255) / 256, 8192);
    k_ppr_init<<<std::max(1u, ib), 256, 0, s>>>(S.r.p, S.c.p, V.od, S.seeds.p, n, B);
    GGPU_HIP_TRY(hipGetLastError());
    const bool tol_mode = o.tol > 0;
    const uint32_t wpb = o.tg / kWave;
    const uint32_t groups = (uint32_t)(((uint64_t)n + kWave / L - 1) / (kWave / L));
    const uint64_t cus = (uint64_t)std::max(V.ctx->cus, 1);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>({((uint64_t)S.n_heavy + groups + wpb - 1) / wpb, cus * S.occ_sweep, kPprBlocks}));
    const uint32_t cgrid =
        (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(((uint64_t)S.n_chunks + wpb - 1) / wpb, cus * S.occ_chunks));
    ggpu_pr_result res{};
    res.last_delta = -1;
    res.last_max_rel = -1;
    auto* pin = (double*)ctx_pinned(V.ctx, 4096);
    int k = 0;
    for (; k < o.iters; ++k) {
        const int c = k & 1, nx = c ^ 1;
        PprArgs A{};
        A.rp = V.rp;
        A.ci = V.ci;
        A.od = V.od;
        A.seeds = S.seeds.p;
        A.c_in = S.c.p + (size_t)c * cells;
        A.c_out = S.c.p + (size_t)nx * cells;
        A.r_old = tol_mode ? S.r.p + (size_t)c * cells : nullptr;
        A.r_new = S.r.p + (size_t)nx * cells;
        A.dang_in = S.dang.p + (size_t)c * kPprMaxB;
        A.partial = S.part.p;
        A.chunks = S.chunks.p;
        A.heavy = S.heavy.p;
        A.cpart = S.cpart.p;
        A.n = n;
        A.B = B;
        A.Bw = Bw;
        A.L = L;
        A.heavy_min = o.large_min;
        A.n_heavy = S.n_heavy;
        A.n_groups = groups;
        A.n_chunks = S.n_chunks;
        A.damping = o.damping;
        switch (np) {
            case 1: ppr_launch<1>(A, grid, cgrid, o.tg, s); break;
            case 2: ppr_launch<2>(A, grid, cgrid, o.tg, s); break;
            case 3: ppr_launch<3>(A, grid, cgrid, o.tg, s); break;
            default: ppr_launch<4>(A, grid, cgrid, o.tg, s); break;
        }
        k_ppr_reduce<<<B + 1, kWave, 0, s>>>(S.part.p, grid, B, S.dang.p + (size_t)nx * kPprMaxB, S.delta.p);
        GGPU_HIP_TRY(hipGetLastError());
        if (!tol_mode) continue;
        GGPU_HIP_TRY(hipMemcpyAsync(pin, S.delta.p, 8, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
        res.last_delta = pin[0];
        if (pin[0] < o.tol) {
            ++k;
            break;
        }
    }
    GGPU_HIP_TRY(hipEventRecord(S.ev[1], s));
    GGPU_HIP_TRY(hipStreamSynchronize(s));
    float ms = 0;
    GGPU_HIP_TRY(hipEventElapsedTime(&ms, S.ev[0], S.ev[1]));
    res.gpu_ms = ms;
    res.iters = k;
    S.slot = k & 1;
    S.order = o.order;
    S.have = true;
    return res;
}

// Times every candidate (heavy threshold, lanes per light vertex, workgroup
// size) for `iters` fixed sweeps of this block and keeps the fastest in o.
// The ranks of the last candidate run are left in the block; every candidate
// computes the same sums up to the order of FP32 additions.
void ppr_autotune(PrEngine& e, const uint32_t* seeds, uint32_t B, int iters, ggpu_pr_opts& o) {
    if (iters <= 0) throw ArgError("iters must be >= 1");
    if (B == 0 || B > kPprMaxB) throw ArgError("PPR block width must be in [1, 256]");
    const uint32_t Bw = std::min<uint32_t>(pow2_at_least(B), kWave);
    std::vector<uint32_t> lanes;
    for (uint32_t l = Bw; l <= kWave; l <<= 1) lanes.push_back(l);
    const uint32_t heavy[] = {64, 128, 256, 1024};
    const uint32_t tgs[] = {256, 1024};
    std::vector<ggpu_pr_opts> cands;
    for (uint32_t ord : {GGPU_PR_ORDER_GIVEN, GGPU_PR_ORDER_DEGREE})
        for (uint32_t h : heavy)
            for (uint32_t l : lanes)
                for (uint32_t tg : tgs) {
                    ggpu_pr_opts t = o;
                    t.tol = 0;
                    t.iters = iters;
                    t.order = ord;
                    t.large_min = h;
                    t.medium_lanes = l;
                    t.tg = tg;
                    cands.push_back(t);
                }
    double best = -1;
    ggpu_pr_opts pick = o;
    for (const ggpu_pr_opts& t : cands) {
        ggpu_pr_opts w = t;
        w.iters = 1;
        ppr_run(e, seeds, B, w);  // first run of a shape: buffers, heavy lists
        const double ms = ppr_run(e, seeds, B, t).gpu_ms;
        if (best < 0 || ms < best) {
            best = ms;
            pick = t;
        }
    }
    o.order = pick.order;
    o.large_min = pick.large_min;
    o.medium_lanes = pick.medium_lanes;
    o.tg = pick.tg;
}

void ppr_fetch(PrEngine& e, float* out) {
    PprState& S = ppr_state(e);
    if (!S.have) throw ArgError("no PPR result yet: call ggpu_ppr_run first", GGPU_ESTATE);
    const PprView V = ppr_view(e, S.order);
    const size_t cells = (size_t)V.n * S.B;
    const float* src = S.r.p + (size_t)S.slot * cells;
    if (V.id_of) {
        if (!S.id_of_ready) {
            S.id_of.alloc(std::max(V.n, 1u));
            GGPU_HIP_TRY(hipMemcpy(S.id_of.p, V.id_of, 4ull * V.n, hipMemcpyHostToDevice));
            S.id_of_ready = true;
        }
        float* dst = S.r.p + (size_t)(S.slot ^ 1) * cells;
        const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((cells + 255) / 256, 8192));
        k_ppr_unpermute<<<grid, 256, 0, V.ctx->stream>>>(src, dst, S.id_of.p, V.n, S.B);
        GGPU_HIP_TRY(hipGetLastError());
        GGPU_HIP_TRY(hipStreamSynchronize(V.ctx->stream));
        src = dst;
    }
    GGPU_HIP_TRY(hipMemcpy(out, src, 4 * cells, hipMemcpyDeviceToHost));
}

}  // namespace ggpu
