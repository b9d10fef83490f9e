// PageRank backend: buffers, sweep scheduling and the C-ABI.
//
//   GpuPageRank ctor        src/metal/pagerank_gpu.mm:30-68     -> PrEngine()
//   prepareAdaptive         :70-86                              -> prepare_adaptive (host binning)
//   prepareGS               :88-192                             -> gs_order (host, libggpu_host)
//   encodeIteration / run   :205-407                            -> jacobi_sweep / gs_sweep / run
//   ranks()                 :194-203 (maps colour relabeling back) -> fetch_ranks
//
// Jacobi (scalar / vector / adaptive): ONE launch per sweep. The pull writes
// r_new and the next contributions and reduces the FP64 dangling mass, L1
// change and sum in the same pass (the reference needs contrib + reduce +
// pull + diff + reduce dispatches).
// In place (gpu_pull_gs, --inplace): contributions are updated as soon as a
// vertex is done, so later vertices of the sweep read new values; after the
// sweep a finish pass renormalizes to sum 1 (pagerank.metal:85-117), computes
// the FP64 L1 change against the previous normalized ranks and rebuilds the
// contributions and the dangling mass for the next sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

#include "graph.hpp"
#include "pagerank_kernels.hpp"
#include "pr_shared.hpp"

namespace ggpu {

constexpr uint32_t kSecBlocks = 2048;             // block cap per section (grid-stride beyond)
constexpr uint32_t kPartialCap = 3 * kSecBlocks;  // FP64 partials per quantity

// ------------------------------------------------------------------ kernels
__global__ void k_pr_sweep(PrSweep P) { pr_sweep_body(P); }

// Shared tail: block partials, last block reduces into acc_out (overwrite).
__device__ __forceinline__ void block_finish(PrPart part, double* partial, uint32_t* ticket, PrAcc* acc_out) {
    __shared__ double red[4][16];
    __shared__ bool last;
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave, nwv = blockDim.x / kWave;
    double a = wave_sum(part.dang), b = wave_sum(part.delta), c = wave_sum(part.sum), m = wave_max(part.maxrel);
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
        partial[blockIdx.x] = t0;
        partial[kPartialCap + blockIdx.x] = t1;
        partial[2 * kPartialCap + blockIdx.x] = t2;
        partial[3 * kPartialCap + blockIdx.x] = t3;
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    for (uint32_t j = threadIdx.x; j < gridDim.x; j += blockDim.x) {
        t0 += partial[j];
        t1 += partial[kPartialCap + j];
        t2 += partial[2 * kPartialCap + j];
        const double q = partial[3 * kPartialCap + j];
        t3 = q > t3 ? q : t3;
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
        acc_out->dangling = T0;
        acc_out->delta = T1;
        acc_out->sum = T2;
        acc_out->maxrel = T3;
        *ticket = 0;
    }
}

// r = 1/n, contrib = r/outdeg; acc_out = {dangling, 0, sum, 0}.
__global__ void k_pr_init(float* r, float* contrib, const uint32_t* outdeg, uint32_t n, double* partial,
                          uint32_t* ticket, PrAcc* acc_out) {
    PrPart part;
    const float r0 = 1.0f / (float)n;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        r[v] = r0;
        const uint32_t od = outdeg[v];
        contrib[v] = od ? r0 / (float)od : 0.0f;
        if (!od) part.dang += (double)r0;
        part.sum += (double)r0;
    }
    block_finish(part, partial, ticket, acc_out);
}

// After an in-place sweep: r_new *= 1/sum (sum from acc_sweep), L1 change
// against the previous normalized ranks, new contributions and dangling mass.
__global__ void k_pr_finish(float* r_new, const float* r_old, float* contrib, const uint32_t* outdeg, uint32_t n,
                            const PrAcc* acc_sweep, double* partial, uint32_t* ticket, PrAcc* acc_out) {
    PrPart part;
    const float scale = (float)(1.0 / acc_sweep->sum);
    const float inv_n = 1.0f / (float)n;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const float x = r_new[v] * scale;
        r_new[v] = x;
        const uint32_t od = outdeg[v];
        contrib[v] = od ? x / (float)od : 0.0f;
        if (!od) part.dang += (double)x;
        part.sum += (double)x;
        const float old = r_old[v];
        const float dl = fabsf(x - old);
        part.delta += (double)dl;
        const double rel = (double)dl / (double)fmaxf(old, inv_n);
        part.maxrel = rel > part.maxrel ? rel : part.maxrel;
    }
    block_finish(part, partial, ticket, acc_out);
}

// ------------------------------------------------------------------ engine
struct PrEngine {
    ggpu_ctx* ctx;
    uint32_t n;
    uint64_t m;
    Csr in_host;
    std::vector<uint32_t> od_host;
    DBuf<uint32_t> rp, ci, od;
    DBuf<float> r, contrib;  // 2n each
    DBuf<double> partial;
    DBuf<uint32_t> ticket;
    DBuf<PrAcc> acc;         // [0],[1] Jacobi ping-pong / GS state, [2] GS sweep totals
    // adaptive bins: [small | medium | large], vertex ids ascending within a bin
    DBuf<uint32_t> bins;
    uint32_t bin_count[3] = {0, 0, 0};
    uint32_t bin_small = 0, bin_large = 0;
    bool bins_ready = false;
    // Gauss-Seidel order
    GsOrder gs;
    bool gs_ready = false;
    uint32_t gs_pair_k = 0;
    DBuf<uint32_t> gs_rp, gs_ci, gs_od, gs_lists;
    // result
    int result_slot = 0;      // which half of r holds the last ranks
    bool result_relabeled = false;
    bool have_result = false;
    std::vector<hipEvent_t> ev;
    // personalized PageRank block (ppr.hip)
    PprState ppr;

    PrEngine(ggpu_ctx* c, uint64_t n_, uint64_t m_, const uint32_t* hrp, const uint32_t* hci, const uint32_t* hod)
        : ctx(c), n((uint32_t)n_), m(m_) {
        if (n_ >= GGPU_INF) throw ArgError("graph too large for 32-bit vertex ids");
        if (m_ >= (1ull << 32)) throw ArgError("graph has 2^32 or more arcs");
        in_host = from_arrays(n_, m_, hrp, hci, GGPU_FLAG_SORTED);
        od_host.assign(hod, hod + n_);
        GGPU_HIP_TRY(hipSetDevice(ctx->device));
        rp.alloc(n + 1);
        ci.alloc(m ? m : 1);
        od.alloc(n ? n : 1);
        r.alloc(2ull * (n ? n : 1));
        contrib.alloc(2ull * (n ? n : 1));
        partial.alloc(4ull * kPartialCap);
        ticket.alloc(1);
        acc.alloc(4);
        hipStream_t s = ctx->stream;
        GGPU_HIP_TRY(hipMemcpyAsync(rp.p, hrp, 4ull * (n + 1), hipMemcpyHostToDevice, s));
        if (m) GGPU_HIP_TRY(hipMemcpyAsync(ci.p, hci, 4ull * m, hipMemcpyHostToDevice, s));
        if (n) GGPU_HIP_TRY(hipMemcpyAsync(od.p, hod, 4ull * n, hipMemcpyHostToDevice, s));
        GGPU_HIP_TRY(hipMemsetAsync(ticket.p, 0, 4, s));
        GGPU_HIP_TRY(hipMemsetAsync(acc.p, 0, 4 * sizeof(PrAcc), s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
    }
    ~PrEngine() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    hipEvent_t event(size_t i) {
        while (ev.size() <= i) {
            hipEvent_t e;
            GGPU_HIP_TRY(hipEventCreate(&e));
            ev.push_back(e);
        }
        return ev[i];
    }

    void prepare_adaptive(uint32_t small_max, uint32_t large_min) {
        if (bins_ready && bin_small == small_max && bin_large == large_min) return;
        std::vector<uint32_t> lists[3];
        for (uint32_t v = 0; v < n; ++v) {
            const uint32_t d = in_host.deg(v);
            lists[d < small_max ? 0 : (d < large_min ? 1 : 2)].push_back(v);
        }
        std::vector<uint32_t> all;
        all.reserve(n);
        for (int b = 0; b < 3; ++b) {
            bin_count[b] = (uint32_t)lists[b].size();
            all.insert(all.end(), lists[b].begin(), lists[b].end());
        }
        bins.alloc(n ? n : 1);
        if (n) GGPU_HIP_TRY(hipMemcpy(bins.p, all.data(), 4ull * n, hipMemcpyHostToDevice));
        bin_small = small_max;
        bin_large = large_min;
        bins_ready = true;
    }

    void prepare_gs(uint32_t pair_k) {
        if (gs_ready && gs_pair_k == pair_k) return;
        gs = gs_order(in_host, od_host.data(), pair_k);
        gs_pair_k = pair_k;
        if (gs.colored) {
            const Csr& R = gs.relabeled_in;
            gs_rp.alloc(n + 1);
            gs_ci.alloc(R.m ? R.m : 1);
            gs_od.alloc(n ? n : 1);
            GGPU_HIP_TRY(hipMemcpy(gs_rp.p, R.rp.data(), 4ull * (n + 1), hipMemcpyHostToDevice));
            if (R.m) GGPU_HIP_TRY(hipMemcpy(gs_ci.p, R.ci.data(), 4ull * R.m, hipMemcpyHostToDevice));
            if (n) GGPU_HIP_TRY(hipMemcpy(gs_od.p, gs.relabeled_outdeg.data(), 4ull * n, hipMemcpyHostToDevice));
        } else {
            std::vector<uint32_t> both(gs.main_list);
            both.insert(both.end(), gs.deferred_list.begin(), gs.deferred_list.end());
            gs_lists.alloc(n ? n : 1);
            if (n) GGPU_HIP_TRY(hipMemcpy(gs_lists.p, both.data(), 4ull * n, hipMemcpyHostToDevice));
        }
        gs_ready = true;
    }

    static uint32_t blocks(uint64_t items, uint32_t per_block) {
        const uint64_t b = (items + per_block - 1) / per_block;
        return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, kSecBlocks));
    }

    // Builds the sections of one launch over [list or range]; returns the grid.
    uint32_t add_section(PrSweep& P, const uint32_t* list, uint32_t first, uint32_t count, uint32_t mode,
                         uint32_t lanes, uint32_t tg, uint32_t grid) {
        if (count == 0) return grid;
        PrSection& S = P.sec[P.nsec++];
        S.list = list;
        S.first = first;
        S.count = count;
        S.mode = mode;
        S.lanes = lanes;
        uint32_t nb;
        if (mode == SEC_LANE) nb = blocks(count, tg);
        else if (mode == SEC_GROUP) nb = blocks(count, tg / lanes);
        else nb = blocks(count, 1);
        S.blk_begin = grid;
        S.blk_end = grid + nb;
        return grid + nb;
    }

    PrSweep base_sweep(bool relabeled, float damping) {
        PrSweep P{};
        P.rp = relabeled ? gs_rp.p : rp.p;
        P.ci = relabeled ? gs_ci.p : ci.p;
        P.outdeg = relabeled ? gs_od.p : od.p;
        P.partial = partial.p;
        P.ticket = ticket.p;
        P.partial_cap = kPartialCap;
        P.n = n;
        P.damping = damping;
        return P;
    }

    void launch(PrSweep& P, uint32_t grid, uint32_t tg) {
        if (P.nsec == 0) return;
        k_pr_sweep<<<grid, tg, 0, ctx->stream>>>(P);
        GGPU_HIP_TRY(hipGetLastError());
    }

    void init(const uint32_t* outdeg, PrAcc* acc_out) {
        k_pr_init<<<blocks(n, 256), 256, 0, ctx->stream>>>(r.p, contrib.p, outdeg, n, partial.p, ticket.p, acc_out);
        GGPU_HIP_TRY(hipGetLastError());
    }

    // One Jacobi sweep k (contrib[k&1] -> contrib[(k+1)&1], r[k&1] -> r[(k+1)&1]).
    void jacobi_sweep(int variant, const ggpu_pr_opts& o, int k, bool want_delta) {
        const int c = k & 1, nx = c ^ 1;
        PrSweep P = base_sweep(false, o.damping);
        P.contrib_in = contrib.p + (size_t)c * n;
        P.contrib_out = contrib.p + (size_t)nx * n;
        P.r_old = want_delta ? r.p + (size_t)c * n : nullptr;
        P.r_new = r.p + (size_t)nx * n;
        P.acc_in = acc.p + c;
        P.acc_out = acc.p + nx;
        P.first = 1;
        uint32_t grid = 0;
        const uint32_t tg = o.tg;
        if (variant == GGPU_PR_SCALAR) grid = add_section(P, nullptr, 0, n, SEC_LANE, 1, tg, grid);
        else if (variant == GGPU_PR_VECTOR) grid = add_section(P, nullptr, 0, n, SEC_GROUP, kWave, tg, grid);
        else {
            grid = add_section(P, bins.p + bin_count[0] + bin_count[1], 0, bin_count[2], SEC_BLOCK, 1, tg, grid);
            grid = add_section(P, bins.p + bin_count[0], 0, bin_count[1], SEC_GROUP, o.medium_lanes, tg, grid);
            grid = add_section(P, bins.p, 0, bin_count[0], SEC_LANE, 1, tg, grid);
        }
        launch(P, grid, tg);
    }

    // One in-place sweep (gpu_pull_gs, or a Jacobi variant with in_place) + finish.
    void inplace_sweep(int variant, const ggpu_pr_opts& o, int k) {
        const bool relabeled = variant == GGPU_PR_GS && gs.colored;
        const int c = k & 1, nx = c ^ 1;
        float* rin = r.p + (size_t)c * n;
        float* rout = r.p + (size_t)nx * n;
        auto fresh = [&](uint32_t first) {
            PrSweep P = base_sweep(relabeled, o.damping);
            P.contrib_in = contrib.p;
            P.contrib_out = contrib.p;
            P.r_old = nullptr;
            P.r_new = rout;
            P.acc_in = acc.p + 0;
            P.acc_out = acc.p + 2;
            P.first = first;
            return P;
        };
        const uint32_t tg = o.tg;
        uint32_t launches = 0;
        if (variant == GGPU_PR_GS && gs.colored) {
            // one lane-per-vertex launch per colour set, in order
            for (size_t s = 0; s + 1 < gs.set_begin.size(); ++s) {
                PrSweep P = fresh(launches == 0);
                const uint32_t a = gs.set_begin[s], b = gs.set_begin[s + 1];
                const uint32_t grid = add_section(P, nullptr, a, b - a, SEC_LANE, 1, tg, 0);
                launch(P, grid, tg);
                ++launches;
            }
        } else if (variant == GGPU_PR_GS) {
            // natural order, wave per vertex; deferred pair partners in a second launch
            const uint32_t na = (uint32_t)gs.main_list.size(), nb = (uint32_t)gs.deferred_list.size();
            PrSweep P = fresh(1);
            uint32_t grid = add_section(P, gs_lists.p, 0, na, SEC_GROUP, kWave, tg, 0);
            launch(P, grid, tg);
            ++launches;
            if (nb) {
                PrSweep Q = fresh(0);
                grid = add_section(Q, gs_lists.p + na, 0, nb, SEC_GROUP, kWave, tg, 0);
                launch(Q, grid, tg);
            }
        } else {
            PrSweep P = fresh(1);
            uint32_t grid = 0;
            if (variant == GGPU_PR_SCALAR) grid = add_section(P, nullptr, 0, n, SEC_LANE, 1, tg, grid);
            else if (variant == GGPU_PR_VECTOR) grid = add_section(P, nullptr, 0, n, SEC_GROUP, kWave, tg, grid);
            else {
                grid = add_section(P, bins.p + bin_count[0] + bin_count[1], 0, bin_count[2], SEC_BLOCK, 1, tg, grid);
                grid = add_section(P, bins.p + bin_count[0], 0, bin_count[1], SEC_GROUP, o.medium_lanes, tg, grid);
                grid = add_section(P, bins.p, 0, bin_count[0], SEC_LANE, 1, tg, grid);
            }
            launch(P, grid, tg);
        }
        k_pr_finish<<<blocks(n, 256), 256, 0, ctx->stream>>>(rout, rin, contrib.p, relabeled ? gs_od.p : od.p, n,
                                                              acc.p + 2, partial.p, ticket.p, acc.p + 0);
        GGPU_HIP_TRY(hipGetLastError());
    }

    ggpu_pr_result run(int variant, const ggpu_pr_opts& o) {
        if (o.tg < 64 || o.tg > 1024 || o.tg % 64) throw ArgError("workgroup size must be a multiple of 64 in [64, 1024]");
        if (variant < GGPU_PR_SCALAR || variant > GGPU_PR_GS) throw ArgError("unknown PageRank variant");
        const uint32_t L = o.medium_lanes;
        if (L == 0 || L > 64 || (L & (L - 1))) throw ArgError("medium_lanes must be a power of two in [1, 64]");
        if (o.iters < 0) throw ArgError("iters must be >= 0");
        if (variant == GGPU_PR_ADAPTIVE) prepare_adaptive(o.small_max, o.large_min);
        if (variant == GGPU_PR_GS && !gs_ready) prepare_gs(2);
        const bool inplace = variant == GGPU_PR_GS || o.in_place;
        if (inplace && variant == GGPU_PR_ADAPTIVE) prepare_adaptive(o.small_max, o.large_min);
        ggpu_pr_result res{};
        res.last_delta = -1;
        res.last_max_rel = -1;
        have_result = false;
        result_relabeled = variant == GGPU_PR_GS && gs.colored;
        if (n == 0) return res;
        hipStream_t s = ctx->stream;
        const bool tol_mode = o.tol > 0;
        GGPU_HIP_TRY(hipEventRecord(event(0), s));
        init(result_relabeled ? gs_od.p : od.p, acc.p + 0);
        auto* pin = (PrAcc*)ctx_pinned(ctx, 4096);
        double best = 1e300;
        int since_best = 0;
        int k = 0;
        for (; k < o.iters; ++k) {
            const PrAcc* out;
            if (inplace) {
                inplace_sweep(variant, o, k);
                out = acc.p + 0;
            } else {
                jacobi_sweep(variant, o, k, tol_mode);
                out = acc.p + ((k + 1) & 1);
            }
            if (!tol_mode) continue;
            // tol mode: one synchronization per sweep to read the FP64 change
            GGPU_HIP_TRY(hipMemcpyAsync(pin, out, sizeof(PrAcc), hipMemcpyDeviceToHost, s));
            GGPU_HIP_TRY(hipStreamSynchronize(s));
            res.iters = k + 1;
            res.last_delta = pin->delta;
            res.last_max_rel = pin->maxrel;
            res.dangling = pin->dangling;
            res.sum = pin->sum;
            if (res.last_delta < o.tol && (o.tol_rel <= 0 || res.last_max_rel < o.tol_rel)) {
                ++k;
                break;
            }
            // stagnation guard (pagerank_gpu.mm:515-533): 5 sweeps without a new best
            if (res.last_delta < best) {
                best = res.last_delta;
                since_best = 0;
            } else if (++since_best >= 5) {
                res.floor_stop = 1;
                ++k;
                break;
            }
        }
        GGPU_HIP_TRY(hipEventRecord(event(1), s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
        float ms = 0;
        GGPU_HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        res.gpu_ms = ms;
        res.iters = k;
        // ranks live in r[k & 1]: Jacobi writes r[(k+1)&1] in sweep k; in place too
        result_slot = k & 1;
        if (!tol_mode) {
            const PrAcc* out = inplace ? acc.p + 0 : acc.p + (k & 1);
            GGPU_HIP_TRY(hipMemcpy(pin, out, sizeof(PrAcc), hipMemcpyDeviceToHost));
            res.dangling = pin->dangling;
            res.sum = pin->sum;
        }
        have_result = true;
        return res;
    }

    void fetch_ranks(float* out) {
        if (!have_result) throw ArgError("no PageRank result yet: call ggpu_pr_run first", GGPU_ESTATE);
        const float* src = r.p + (size_t)result_slot * n;
        if (!result_relabeled) {
            GGPU_HIP_TRY(hipMemcpy(out, src, 4ull * n, hipMemcpyDeviceToHost));
            return;
        }
        std::vector<float> tmp(n);
        GGPU_HIP_TRY(hipMemcpy(tmp.data(), src, 4ull * n, hipMemcpyDeviceToHost));
        for (uint32_t v = 0; v < n; ++v) out[v] = tmp[gs.new_id[v]];
    }

    // Time `iters` Jacobi sweeps of gpu_pull_adaptive for each candidate.
    void autotune(int iters, ggpu_pr_opts& o) {
        const uint32_t smalls[] = {4, 8, 16, 32};
        const uint32_t larges[] = {256, 1024, 4096};
        const uint32_t lanes[] = {8, 16, 32, 64};
        float best = 1e30f;
        ggpu_pr_opts bo = o;
        for (uint32_t sm : smalls)
            for (uint32_t lg : larges)
                for (uint32_t L : lanes) {
                    if (L > 4 * sm * 2 && L > 16) continue;  // wide groups only pay off for long lists
                    ggpu_pr_opts t = o;
                    t.small_max = sm;
                    t.large_min = lg;
                    t.medium_lanes = L;
                    t.iters = iters;
                    t.tol = 0;
                    t.in_place = 0;
                    run(GGPU_PR_ADAPTIVE, t);  // warm-up (also builds the bins)
                    ggpu_pr_result rr = run(GGPU_PR_ADAPTIVE, t);
                    if (rr.gpu_ms < best) {
                        best = (float)rr.gpu_ms;
                        bo = t;
                    }
                }
        o.small_max = bo.small_max;
        o.large_min = bo.large_min;
        o.medium_lanes = bo.medium_lanes;
    }
};

}  // namespace ggpu

using namespace ggpu;

struct ggpu_pr {
    ggpu_ctx* ctx;
    PrEngine* eng;
};

namespace ggpu {
PprView ppr_view(PrEngine& e) { return PprView{e.ctx, e.n, e.rp.p, e.ci.p, e.od.p, e.od_host.data()}; }
PprState& ppr_state(PrEngine& e) { return e.ppr; }
}  // namespace ggpu

extern "C" {

void ggpu_pr_opts_default(ggpu_pr_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->damping = 0.85f;
    o->iters = 20;
    o->tol = 0;
    o->tol_rel = 0;
    o->tg = 256;
    o->small_max = 16;
    o->large_min = 256;
    o->in_place = 0;
    o->medium_lanes = 64;
}

int ggpu_pr_parse_variant(const char* name, int* variant) {
    if (!name || !variant) return GGPU_EINVAL;
    const std::string s = name;
    if (s == "gpu_pull_scalar") *variant = GGPU_PR_SCALAR;
    else if (s == "gpu_pull_vector") *variant = GGPU_PR_VECTOR;
    else if (s == "gpu_pull_adaptive") *variant = GGPU_PR_ADAPTIVE;
    else if (s == "gpu_pull_gs") *variant = GGPU_PR_GS;
    else return GGPU_EINVAL;
    return GGPU_OK;
}

int ggpu_pr_create(ggpu_ctx* ctx, uint64_t n, uint64_t m, const uint32_t* in_row_ptr, const uint32_t* in_col_idx,
                   const uint32_t* outdeg, ggpu_pr** out) {
    if (!ctx || !out || !in_row_ptr || (m && !in_col_idx) || (n && !outdeg))
        return set_error(ctx, GGPU_EINVAL, "null argument");
    *out = nullptr;
    PrEngine* e = nullptr;
    int rc = guarded(ctx, [&] { e = new PrEngine(ctx, n, m, in_row_ptr, in_col_idx, outdeg); });
    if (rc != GGPU_OK) return rc;
    *out = new ggpu_pr{ctx, e};
    return GGPU_OK;
}

void ggpu_pr_destroy(ggpu_pr* p) {
    if (!p) return;
    (void)hipStreamSynchronize(p->ctx->stream);
    delete p->eng;
    delete p;
}

int ggpu_pr_prepare_adaptive(ggpu_pr* p, uint32_t small_max, uint32_t large_min) {
    if (!p) return GGPU_EINVAL;
    return guarded(p->ctx, [&] { p->eng->prepare_adaptive(small_max, large_min); });
}

int ggpu_pr_prepare_gs(ggpu_pr* p, uint32_t pair_max_outdeg) {
    if (!p) return GGPU_EINVAL;
    return guarded(p->ctx, [&] { p->eng->prepare_gs(pair_max_outdeg); });
}

int ggpu_pr_gs_info(const ggpu_pr* p, int* colored, uint64_t* sets, uint64_t* deferred) {
    if (!p) return GGPU_EINVAL;
    const PrEngine& e = *p->eng;
    if (colored) *colored = e.gs_ready && e.gs.colored;
    if (sets) *sets = e.gs.colored && !e.gs.set_begin.empty() ? e.gs.set_begin.size() - 1 : 0;
    if (deferred) *deferred = e.gs.deferred_list.size();
    return GGPU_OK;
}

int ggpu_pr_run(ggpu_pr* p, int variant, const ggpu_pr_opts* opt, ggpu_pr_result* res) {
    if (!p) return GGPU_EINVAL;
    ggpu_pr_opts o;
    ggpu_pr_opts_default(&o);
    if (opt) o = *opt;
    return guarded(p->ctx, [&] {
        GGPU_HIP_TRY(hipSetDevice(p->ctx->device));
        ggpu_pr_result r = p->eng->run(variant, o);
        if (res) *res = r;
    });
}

int ggpu_pr_fetch_ranks(ggpu_pr* p, float* out) {
    if (!p || !out) return GGPU_EINVAL;
    return guarded(p->ctx, [&] { p->eng->fetch_ranks(out); });
}

int ggpu_pr_autotune(ggpu_pr* p, int iters, ggpu_pr_opts* opt) {
    if (!p || !opt) return GGPU_EINVAL;
    return guarded(p->ctx, [&] {
        if (iters <= 0) throw ArgError("iters must be >= 1");
        p->eng->autotune(iters, *opt);
    });
}

int ggpu_ppr_run(ggpu_pr* p, const uint32_t* seeds, uint32_t B, const ggpu_pr_opts* opt, ggpu_pr_result* res) {
    if (!p || !seeds) return GGPU_EINVAL;
    ggpu_pr_opts o;
    ggpu_pr_opts_default(&o);
    if (opt) o = *opt;
    return guarded(p->ctx, [&] {
        GGPU_HIP_TRY(hipSetDevice(p->ctx->device));
        ggpu_pr_result r = ppr_run(*p->eng, seeds, B, o);
        if (res) *res = r;
    });
}

int ggpu_ppr_fetch(ggpu_pr* p, float* out) {
    if (!p || !out) return GGPU_EINVAL;
    return guarded(p->ctx, [&] { ppr_fetch(*p->eng, out); });
}

}  // extern "C"
