// Personalized PageRank as an n x B FP32 block (no reference counterpart).
//
// Column b restarts at seeds[b]:
//   r_new[v][b] = d * sum_{u in in(v)} r[u][b] / outdeg[u]
//               + [v == seeds[b]] * ((1 - d) + d * dangling_b)
// with dangling_b the column's rank on dangling vertices (FP64), so every
// column keeps total mass 1 (networkx.pagerank with personalization {s: 1}).
//
// Layout: vertex-major, row v = B consecutive floats. A wave owns one
// destination vertex and lane b one column, so every in-edge is one
// coalesced row gather of 4*B bytes and the in-list is read once for all
// columns. Per-lane Kahan sums; FP64 per-column dangling and L1 change.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "graph.hpp"
#include "pagerank_kernels.hpp"
#include "pr_shared.hpp"

namespace ggpu {

constexpr uint32_t kPprMaxB = 256;
constexpr uint32_t kPprBlocks = 2048;

struct PprArgs {
    const uint32_t* rp;
    const uint32_t* ci;
    const uint32_t* od;
    const uint32_t* seeds;
    const float* c_in;
    float* c_out;
    float* r_new;
    const float* r_old;   // null: no L1 change
    const double* dang_in;  // B
    double* dang_out;       // B
    double* delta_out;      // 1
    double* partial;        // (B + 1) * kPprBlocks
    uint32_t n, B;
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

__global__ void k_ppr_sweep(PprArgs A) {
    const uint32_t lane = lane_id();
    const uint32_t wpb = blockDim.x / kWave, wid = threadIdx.x / kWave;
    const uint32_t chunks = (A.B + kWave - 1) / kWave;
    const float d = A.damping;
    double dang[kPprMaxB / kWave] = {0, 0, 0, 0};
    double delta = 0;
    float restart[kPprMaxB / kWave];
    uint32_t seed[kPprMaxB / kWave];
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint32_t b = c * kWave + lane;
        seed[c] = b < A.B ? A.seeds[b] : GGPU_INF;
        restart[c] = b < A.B ? (float)((1.0 - (double)d) + (double)d * A.dang_in[b]) : 0.0f;
    }
    for (uint32_t v = blockIdx.x * wpb + wid; v < A.n; v += gridDim.x * wpb) {
        const uint32_t e0 = A.rp[v], e1 = A.rp[v + 1];
        const uint32_t odv = A.od[v];
        for (uint32_t c = 0; c < chunks; ++c) {
            const uint32_t b = c * kWave + lane;
            const bool act = b < A.B;
            Kahan k;
            uint32_t x = e0;
            // four row gathers in flight per wave
            for (; x + 4 <= e1; x += 4) {
                const uint32_t u0 = A.ci[x], u1 = A.ci[x + 1], u2 = A.ci[x + 2], u3 = A.ci[x + 3];
                if (act) {
                    const float a0 = A.c_in[(size_t)u0 * A.B + b];
                    const float a1 = A.c_in[(size_t)u1 * A.B + b];
                    const float a2 = A.c_in[(size_t)u2 * A.B + b];
                    const float a3 = A.c_in[(size_t)u3 * A.B + b];
                    k.add(a0);
                    k.add(a1);
                    k.add(a2);
                    k.add(a3);
                }
            }
            for (; x < e1; ++x) {
                const uint32_t u = A.ci[x];
                if (act) k.add(A.c_in[(size_t)u * A.B + b]);
            }
            if (!act) continue;
            float val = d * (k.s - k.c);
            if (seed[c] == v) val += restart[c];
            const size_t at = (size_t)v * A.B + b;
            A.r_new[at] = val;
            A.c_out[at] = odv ? val / (float)odv : 0.0f;
            if (!odv) dang[c] += (double)val;
            if (A.r_old) delta += (double)fabsf(val - A.r_old[at]);
        }
    }
    // block partials: per column dangling (waves combined in LDS), total delta
    __shared__ double s_d[16][kWave];
    __shared__ double s_x[16];
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint32_t b = c * kWave + lane;
        __syncthreads();
        s_d[wid][lane] = dang[c];
        __syncthreads();
        if (wid == 0) {
            double t = 0;
            for (uint32_t w = 0; w < wpb; ++w) t += s_d[w][lane];
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

// One block: column sums of the sweep's block partials, in block order.
__global__ void k_ppr_reduce(const double* partial, uint32_t blocks, uint32_t B, double* dang_out, double* delta_out) {
    for (uint32_t col = threadIdx.x; col <= B; col += blockDim.x) {
        double t = 0;
        for (uint32_t blk = 0; blk < blocks; ++blk) t += partial[(size_t)blk * (B + 1) + col];
        if (col < B) dang_out[col] = t;
        else *delta_out = t;
    }
}

}  // namespace ggpu

namespace ggpu {

ggpu_pr_result ppr_run(PrEngine& e, const uint32_t* seeds, uint32_t B, const ggpu_pr_opts& o) {
    PprView V = ppr_view(e);
    PprState& S = ppr_state(e);
    if (B == 0 || B > kPprMaxB) throw ArgError("PPR block width must be in [1, 256]");
    if (o.tg < 64 || o.tg > 1024 || o.tg % 64) throw ArgError("workgroup size must be a multiple of 64 in [64, 1024]");
    if (o.iters < 0) throw ArgError("iters must be >= 0");
    for (uint32_t b = 0; b < B; ++b)
        if (seeds[b] >= V.n) throw ArgError("seed out of range");
    const uint32_t n = V.n;
    const size_t cells = (size_t)n * B;
    hipStream_t s = V.ctx->stream;
    S.r.alloc(2 * cells);
    S.c.alloc(2 * cells);
    S.part.alloc((size_t)(kPprMaxB + 1) * kPprBlocks);
    S.dang.alloc(2 * kPprMaxB);
    S.delta.alloc(2);
    S.seeds.alloc(kPprMaxB);
    while (S.ev.size() < 2) {
        hipEvent_t ev;
        GGPU_HIP_TRY(hipEventCreate(&ev));
        S.ev.push_back(ev);
    }
    S.B = B;
    std::vector<double> dang0(B);
    for (uint32_t b = 0; b < B; ++b) dang0[b] = V.od_host[seeds[b]] == 0 ? 1.0 : 0.0;
    GGPU_HIP_TRY(hipMemcpyAsync(S.seeds.p, seeds, 4ull * B, hipMemcpyHostToDevice, s));
    GGPU_HIP_TRY(hipMemcpyAsync(S.dang.p, dang0.data(), 8ull * B, hipMemcpyHostToDevice, s));
    GGPU_HIP_TRY(hipEventRecord(S.ev[0], s));
    const uint32_t ib = (uint32_t)std::min<uint64_t>((cells + 255) / 256, 8192);
    k_ppr_init<<<std::max(1u, ib), 256, 0, s>>>(S.r.p, S.c.p, V.od, S.seeds.p, n, B);
    GGPU_HIP_TRY(hipGetLastError());
    const bool tol_mode = o.tol > 0;
    const uint32_t wpb = o.tg / kWave;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + wpb - 1) / wpb, kPprBlocks));
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
        A.dang_out = S.dang.p + (size_t)nx * kPprMaxB;
        A.delta_out = S.delta.p;
        A.partial = S.part.p;
        A.n = n;
        A.B = B;
        A.damping = o.damping;
        k_ppr_sweep<<<grid, o.tg, 0, s>>>(A);
        GGPU_HIP_TRY(hipGetLastError());
        k_ppr_reduce<<<1, 256, 0, s>>>(S.part.p, grid, B, A.dang_out, A.delta_out);
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
    S.have = true;
    return res;
}

void ppr_fetch(PrEngine& e, float* out) {
    PprState& S = ppr_state(e);
    if (!S.have) throw ArgError("no PPR result yet: call ggpu_ppr_run first", GGPU_ESTATE);
    PprView V = ppr_view(e);
    const size_t cells = (size_t)V.n * S.B;
    GGPU_HIP_TRY(hipMemcpy(out, S.r.p + (size_t)S.slot * cells, 4 * cells, hipMemcpyDeviceToHost));
}

}  // namespace ggpu
