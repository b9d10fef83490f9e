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
#include <memory>
#include <numeric>

#include "ggpu_hip_testing.h"
#include "graph.hpp"
#include "pagerank_kernels.hpp"
#include "pr_shared.hpp"
#include "pr_tail.hpp"
#include "pr_tail_plan.hpp"
#include "pr_tiled.hpp"
#include "pr_tiling.hpp"

namespace ggpu {

constexpr uint32_t kSecBlocks = 2048;             // block cap per section (grid-stride beyond)
constexpr uint32_t kStreamBlocks = 8192;          // block cap of a stream-layout sweep
constexpr uint32_t kPartialCap = 1u << 17;        // FP64 partial slots per quantity (all launches of a sweep)
constexpr uint32_t kPushHeavy = 256;              // gpu_push: out-degree from which a wave pushes a source
// HOT_TAIL geometry: the split rows' heads are the sources of the HOT sweep's
// LDS copy (below kHotCap); kBandTiles band tiles of kBandTile sources follow
// (64 KiB of contributions each, copied to LDS by k_pr_band), then tail tiles
// of kTailTile sources (2 MiB of contributions each; one XCD's L2 is 4 MiB).
constexpr uint32_t kTailBegin = kHotCap;
constexpr uint32_t kBandTiles = 30;
constexpr uint32_t kTailTile = 1u << 19;
constexpr uint32_t kTailMergeBlocks = 16384;       // block cap of the tail merge (grid-stride beyond)
// arcs per chunk of a split row's head (LDS gathers only; a wave sums a chunk in
// passes of kWaveNnz): 512 / 1024 / 2048 / 4096 -> 772 / 759 / 762 / 759 us per
// HOT_TAIL sweep on rmat23 (fewer chunks, partials and merge reads)
constexpr uint32_t kHeadChunk = 1024;

// ------------------------------------------------------------------ kernels
__global__ void k_pr_sweep(PrSweep P) { pr_sweep_body(P); }
__global__ void __launch_bounds__(kHotThreads) k_pr_sweep_hot(PrSweep P) { pr_hot_sweep_body(P); }

__global__ void k_pr_longfix(PrSweep P) { pr_longfix_body(P); }
__global__ void __launch_bounds__(kTileThreads) k_pr_sweep_tiled(PrSweep P, PrTiledDev Z) { pr_tiled_body(P, Z); }
__global__ void __launch_bounds__(256) k_pr_tiled_merge(PrSweep P, PrTiledDev Z) { pr_tiled_merge_body(P, Z); }
__global__ void __launch_bounds__(kTailWaves * kWave) k_pr_tail(PrSweep P, PrTailDev Z, uint32_t item_base) {
    pr_tail_body(P, Z, item_base);
}
// two workgroups per CU (their LDS tiles and registers allow it)
__global__ void __launch_bounds__(kBandWaves * kWave, 2)
    k_pr_band(PrSweep P, PrTailDev Z, const uint4* tasks, uint32_t tail_begin, uint32_t band_tile) {
    pr_band_body(P, Z, tasks, tail_begin, band_tile);
}
__global__ void __launch_bounds__(256) k_pr_tail_merge(PrSweep P, PrTailDev Z) { pr_tail_merge_body(P, Z); }
__global__ void k_pr_reduce(const double* partial, uint32_t cap, uint32_t count, PrAcc* out) {
    pr_reduce_body(partial, cap, count, out);
}

// r = 1/n, contrib = r/outdeg; block partials of {dangling, 0, sum, 0}.
__global__ void k_pr_init(float* r, float* contrib, const uint32_t* outdeg, uint32_t n, double* partial,
                          uint32_t cap) {
    PrPart part;
    const float r0 = 1.0f / (float)n;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        r[v] = r0;
        const uint32_t od = outdeg[v];
        contrib[v] = od ? r0 / (float)od : 0.0f;
        if (!od) part.dang += (double)r0;
        part.sum += (double)r0;
    }
    write_block_partials(part, partial, cap, blockIdx.x);
}

// gpu_push (pagerank.metal pr_push / pr_push_finalize): acc[t] gathers the
// contributions of t's in-neighbours through FP64 atomic adds (the hardware
// f64 add on this coarse-grained device memory), so the order in which the
// adds land moves the sum by ~1e-16 relative and needs no compensation: a
// hub's 1e5 adds of near-equal FP32 terms round the same way every time, and
// neither an FP32 accumulator nor an FP32 error accumulator beside it keeps
// that below 1e-6.
__device__ __forceinline__ void push_add(double* acc, uint32_t t, float x) { unsafeAtomicAdd(&acc[t], (double)x); }

// Sources of out-degree >= kPushHeavy are pushed by a wave each (lanes stride
// the out-list), the others by a lane each; zero contributions push nothing.
__global__ void k_pr_push(const uint32_t* orp, const uint32_t* oci, const float* contrib, double* acc, uint32_t n,
                          const uint32_t* heavy, uint32_t n_heavy) {
    const uint32_t lane = lane_id(), wpb = blockDim.x / kWave;
    const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    for (uint32_t h = blockIdx.x * wpb + wid; h < n_heavy; h += gridDim.x * wpb) {
        const uint32_t u = heavy[h];
        const float c = contrib[u];
        if (c == 0.0f) continue;
        for (uint32_t e = orp[u] + lane; e < orp[u + 1]; e += kWave) push_add(acc, oci[e], c);
    }
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        const uint32_t e0 = orp[u], e1 = orp[u + 1];
        if (e1 - e0 >= kPushHeavy) continue;
        const float c = contrib[u];
        if (c == 0.0f) continue;
        for (uint32_t e = e0; e < e1; ++e) push_add(acc, oci[e], c);
    }
}

// r_new = (1-d)/n + d * (acc + dangling/n), stored like the pull sweeps store
// (contributions, FP64 partials); the accumulators are cleared.
__global__ void k_pr_push_finish(PrSweep P, double* acc) {
    const double addd = (1.0 - (double)P.damping) / (double)P.n + (double)P.damping * P.acc_in->dangling / (double)P.n;
    const float add = (float)addd, d = P.damping, inv_n = 1.0f / (float)P.n;
    PrPart part;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < P.n; v += (uint64_t)gridDim.x * blockDim.x) {
        const float val = add + d * (float)acc[v];
        acc[v] = 0.0;
        pr_store(P, (uint32_t)v, val, part, inv_n);
    }
    write_block_partials(part, P.partial, P.partial_cap, P.partial_base + blockIdx.x);
}

// After an in-place sweep: r_new *= 1/sum (sum from acc_sweep), L1 change
// against the previous normalized ranks, new contributions and dangling mass.
__global__ void k_pr_finish(float* r_new, const float* r_old, float* contrib, const uint32_t* outdeg, uint32_t n,
                            const PrAcc* acc_sweep, double* partial, uint32_t cap) {
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
    write_block_partials(part, partial, cap, blockIdx.x);
}

// ------------------------------------------------------------------ engine
// Row-block items of the stream layout over rows [r_begin, r_end).
struct StreamLayout {
    uint32_t r_begin = 0, r_end = 0;
    DBuf<uint4> sitems;
    DBuf<uint2> lrows;
    DBuf<uint4> litems;
    DBuf<float2> lpart;
    DBuf<uint16_t> h16;  // HOT_TAIL: the split rows' head arc ids (items flagged kLiHead16)
    uint32_t ns = 0, nl = 0, nlr = 0;
    bool multi = false;  // some long row has more than one chunk
};

// The TILED layout's plan (pr_tiling.hpp) on the device.
struct TiledLayout {
    uint32_t min_indeg = 0, R = 0, K = 0, groups = 0;
    uint64_t dense_arcs = 0, cold_arcs = 0, untiled_arcs = 0;
    DBuf<uint4> items, segs;
    DBuf<TileTask> tasks;
    DBuf<uint2> wg;
    DBuf<uint16_t> tci, ent_len;
    DBuf<uint32_t> cci, ent_row, trow, optr, oidx;
    DBuf<float2> part, ovf;
    PrTiledDev dev() const {
        PrTiledDev Z{};
        Z.items = items.p;
        Z.tasks = tasks.p;
        Z.wg_tasks = wg.p;
        Z.segs = segs.p;
        Z.tci = tci.p;
        Z.cci = cci.p;
        Z.ent_row = ent_row.p;
        Z.ent_len = ent_len.p;
        Z.part = part.p;
        Z.ovf = ovf.p;
        Z.optr = optr.p;
        Z.oidx = oidx.p;
        Z.trow = trow.p;
        Z.R = R;
        Z.K = K;
        return Z;
    }
};

// The HOT_TAIL layout of one numbering: the HOT items with the split rows'
// heads as flagged long rows (stream), the tail plan (pr_tail_plan.hpp) on the
// device, the band tasks, and the partial array.
struct TailLayout {
    uint32_t min_indeg = 0, tail_begin = 0, tile = 0, band_end = 0, band_tile = 0;
    StreamLayout stream;
    uint32_t R = 0, T = 0, n_items = 0, n_big = 0;
    uint32_t band_tiles = 0, n_tasks = 0, tail_item0 = 0;  // items [0, tail_item0) are the band tiles'
    uint64_t tail_arcs = 0, n_part = 0;
    DBuf<uint4> tasks;  // band tasks {tile, first item, end item, 0}
    DBuf<uint32_t> tci, px, pdst, srow, slr, ovptr, big;
    DBuf<uint2> items;
    DBuf<float2> part;
    PrTailDev dev() const {
        PrTailDev Z{};
        Z.tci = tci.p;
        Z.px = px.p;
        Z.pdst = pdst.p;
        Z.items = items.p;
        Z.n_items = n_items;
        Z.part = part.p;
        Z.n_part = n_part;
        Z.srow = srow.p;
        Z.slr = slr.p;
        Z.ovptr = ovptr.p;
        Z.big = big.p;
        Z.n_big = n_big;
        Z.R = R;
        Z.T = T;
        return Z;
    }
};

template <class T, class H>
static void upload_vec(DBuf<T>& d, const std::vector<H>& h) {
    static_assert(sizeof(T) == sizeof(H), "host and device records differ in size");
    d.alloc(std::max<size_t>(h.size(), 1));
    // an empty list keeps one zeroed placeholder (clamped loads read element 0 and use it as an index)
    if (!h.empty()) GGPU_HIP_TRY(hipMemcpy(d.p, h.data(), h.size() * sizeof(H), hipMemcpyHostToDevice));
    else GGPU_HIP_TRY(hipMemset(d.p, 0, sizeof(T)));
}

// One vertex numbering of the in-CSR with its device copy and work layouts.
// Spaces: the caller's ids; descending out-degree (hot contributions packed
// into few cache lines); the Gauss-Seidel colour order (colour sets are
// contiguous id ranges); the Gauss-Seidel natural order with deferred pair
// partners moved to the end. id_of[caller id] = id in the space.
enum SpaceId { SP_GIVEN = 0, SP_DEGREE = 1, SP_COLOR = 2, SP_NATURAL = 3, SP_COUNT = 4 };

struct Space {
    bool built = false;
    Csr in;
    std::vector<uint32_t> od;
    std::vector<uint32_t> id_of;  // empty: identity
    std::vector<std::pair<uint32_t, uint32_t>> ranges;  // GS: consecutive update ranges
    DBuf<uint32_t> rp, ci, odd;
    std::vector<std::unique_ptr<StreamLayout>> stream;  // one per range
    std::unique_ptr<StreamLayout> wstream;              // HOT layout: wave-sized items over all rows
    std::unique_ptr<TiledLayout> tiled;                 // TILED layout
    std::unique_ptr<TailLayout> tail;                   // HOT_TAIL layout
    DBuf<uint32_t> bins;
    uint32_t bin_count[3] = {0, 0, 0};
    // gpu_push: out-arcs (the transpose of `in`), sources of out-degree >= kPushHeavy
    DBuf<uint32_t> orp, oci, oheavy;
    uint32_t n_oheavy = 0;
    bool out_ready = false;
    uint32_t bin_small = 0, bin_large = 0;
    bool bins_ready = false;
};

struct PrEngine {
    ggpu_ctx* ctx;
    uint32_t n;
    uint64_t m;
    Space sp[SP_COUNT];
    DBuf<float> r, contrib;  // 2n each
    DBuf<double> pacc;       // gpu_push: FP64 accumulators acc[n] (zero between sweeps)
    DBuf<double> partial;
    uint32_t npart = 0;      // partial slots written since the last reduction
    DBuf<PrAcc> acc;         // [0],[1] Jacobi ping-pong / in-place state, [2] in-place sweep totals
    // HOT_TAIL geometry (test hook overrides; tail_begin and tile both 0: the defaults)
    uint32_t tail_begin_set = 0, tail_tile_set = 0, band_end_set = 0, band_tile_set = 0;
    // Gauss-Seidel decision (made on the caller's numbering)
    bool gs_ready = false, gs_colored = false;
    uint32_t gs_pair_k = 0;
    int natural_base = -1;
    uint64_t gs_deferred = 0;
    // result
    int result_slot = 0;
    int result_space = SP_GIVEN;
    bool have_result = false;
    std::vector<hipEvent_t> ev;
    // personalized PageRank block (ppr.hip)
    PprState ppr;

    PrEngine(ggpu_ctx* c, uint64_t n_, uint64_t m_, const uint32_t* hrp, const uint32_t* hci, const uint32_t* hod)
        : ctx(c), n((uint32_t)n_), m(m_) {
        if (n_ >= GGPU_INF) throw ArgError("graph too large for 32-bit vertex ids");
        if (m_ >= (1ull << 32)) throw ArgError("graph has 2^32 or more arcs");
        GGPU_HIP_TRY(hipSetDevice(ctx->device));
        Space& g = sp[SP_GIVEN];
        g.in = from_arrays(n_, m_, hrp, hci, GGPU_FLAG_SORTED);
        g.od.assign(hod, hod + n_);
        upload(g);
        r.alloc(2ull * (n ? n : 1));
        contrib.alloc(2ull * (n ? n : 1));
        partial.alloc(4ull * kPartialCap);
        acc.alloc(4);
        hipStream_t s = ctx->stream;
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

    void upload(Space& S) {
        S.rp.alloc(n + 1);
        S.ci.alloc(S.in.m ? S.in.m : 1);
        S.odd.alloc(n ? n : 1);
        GGPU_HIP_TRY(hipMemcpy(S.rp.p, S.in.rp.data(), 4ull * (n + 1), hipMemcpyHostToDevice));
        // a graph without arcs keeps one placeholder arc, source 0: the TILED sweep's
        // clamped loads read arc 0 for empty rows and gather the contribution it names
        if (S.in.m) GGPU_HIP_TRY(hipMemcpy(S.ci.p, S.in.ci.data(), 4ull * S.in.m, hipMemcpyHostToDevice));
        else GGPU_HIP_TRY(hipMemset(S.ci.p, 0, 4));
        if (n) GGPU_HIP_TRY(hipMemcpy(S.odd.p, S.od.data(), 4ull * n, hipMemcpyHostToDevice));
        if (S.ranges.empty()) S.ranges.push_back({0u, n});
        S.built = true;
    }

    // Relabels `base` by perm (perm[base id] = new id) into space `dst`.
    void derive(int dst, const Space& base, const std::vector<uint32_t>& perm) {
        Space& S = sp[dst];
        S.in = permute(base.in, perm);
        S.od.assign(n, 0);
        for (uint32_t v = 0; v < n; ++v) S.od[perm[v]] = base.od[v];
        S.id_of.resize(n);
        for (uint32_t v = 0; v < n; ++v) S.id_of[v] = perm[base.id_of.empty() ? v : base.id_of[v]];
        S.stream.clear();
        S.wstream.reset();
        S.tiled.reset();
        S.tail.reset();
        S.bins_ready = false;
        S.out_ready = false;
    }

    // gpu_push's out-arcs of space S: the transpose of its in-CSR.
    void ensure_out(Space& S) {
        if (S.out_ready) return;
        const Csr out = transpose(S.in);
        S.orp.alloc(n + 1);
        S.oci.alloc(out.m ? out.m : 1);
        GGPU_HIP_TRY(hipMemcpy(S.orp.p, out.rp.data(), 4ull * (n + 1), hipMemcpyHostToDevice));
        if (out.m) GGPU_HIP_TRY(hipMemcpy(S.oci.p, out.ci.data(), 4ull * out.m, hipMemcpyHostToDevice));
        else GGPU_HIP_TRY(hipMemset(S.oci.p, 0, 4));
        std::vector<uint32_t> heavy;
        for (uint32_t v = 0; v < n; ++v)
            if (out.deg(v) >= kPushHeavy) heavy.push_back(v);
        S.oheavy.alloc(std::max<size_t>(heavy.size(), 1));
        if (!heavy.empty())
            GGPU_HIP_TRY(hipMemcpy(S.oheavy.p, heavy.data(), 4ull * heavy.size(), hipMemcpyHostToDevice));
        S.n_oheavy = (uint32_t)heavy.size();
        S.out_ready = true;
    }

    // One gpu_push sweep: atomics along the out-arcs, then the stores and
    // FP64 partials of every vertex (the accumulators are cleared for the next).
    void push_sweep(Space& S, const ggpu_pr_opts& o, int k, bool want_delta) {
        const int c = k & 1, nx = c ^ 1;
        PrSweep P = base_sweep(S, o.damping);
        P.contrib_in = contrib.p + (size_t)c * n;
        P.contrib_out = contrib.p + (size_t)nx * n;
        P.r_old = want_delta ? r.p + (size_t)c * n : nullptr;
        P.r_new = r.p + (size_t)nx * n;
        P.acc_in = acc.p + c;
        const uint32_t tg = o.tg;
        const uint32_t grid = blocks(std::max<uint64_t>(n, (uint64_t)S.n_oheavy * kWave), tg);
        k_pr_push<<<grid, tg, 0, ctx->stream>>>(S.orp.p, S.oci.p, P.contrib_in, pacc.p, n, S.oheavy.p, S.n_oheavy);
        GGPU_HIP_TRY(hipGetLastError());
        const uint32_t g = blocks(n, 256);
        if (npart + g > kPartialCap) throw ArgError("too many blocks in one sweep");
        P.partial_base = npart;
        npart += g;
        k_pr_push_finish<<<g, 256, 0, ctx->stream>>>(P, pacc.p);
        GGPU_HIP_TRY(hipGetLastError());
        reduce(acc.p + nx);
    }

    Space& space(int id) {
        Space& S = sp[id];
        if (S.built) return S;
        if (id == SP_DEGREE) {
            // descending out-degree, ties by id: the most gathered contributions first
            const Space& g = sp[SP_GIVEN];
            std::vector<uint32_t> order(n);
            std::iota(order.begin(), order.end(), 0u);
            std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return g.od[a] > g.od[b]; });
            std::vector<uint32_t> perm(n);
            for (uint32_t i = 0; i < n; ++i) perm[order[i]] = i;
            derive(SP_DEGREE, g, perm);
            upload(S);
            return S;
        }
        throw ArgError("space not prepared");
    }

    // Row blocks of <= nnz_cap entries and <= rows_cap rows over rows [ra, rb);
    // rows longer than nnz_cap become chunks of <= chunk entries.
    static std::unique_ptr<StreamLayout> make_stream(const Space& S, uint32_t ra, uint32_t rb, uint32_t nnz_cap,
                                                     uint32_t rows_cap, uint32_t chunk) {
        auto L = std::make_unique<StreamLayout>();
        L->r_begin = ra;
        L->r_end = rb;
        std::vector<uint4> sitems, litems;
        std::vector<uint2> lrows;
        uint32_t cur0 = ra, rows = 0, nnz = 0;
        auto close = [&](uint32_t end) {
            if (rows) sitems.push_back(make_uint4(cur0, end, S.in.rp[cur0], S.in.rp[end]));
            rows = 0;
            nnz = 0;
        };
        for (uint32_t v = ra; v < rb; ++v) {
            const uint32_t dg = S.in.deg(v);
            if (dg > nnz_cap) {
                close(v);
                const uint32_t first = (uint32_t)litems.size();
                for (uint32_t x = S.in.rp[v]; x < S.in.rp[v + 1]; x += chunk)
                    litems.push_back(make_uint4(v, x, std::min(x + chunk, S.in.rp[v + 1]), (uint32_t)lrows.size()));
                lrows.push_back(make_uint2(first, (uint32_t)litems.size() - first));
                continue;
            }
            if (rows && (nnz + dg > nnz_cap || rows == rows_cap)) close(v);
            if (!rows) cur0 = v;
            ++rows;
            nnz += dg;
        }
        close(rb);
        L->ns = (uint32_t)sitems.size();
        L->nl = (uint32_t)litems.size();
        L->nlr = (uint32_t)lrows.size();
        L->sitems.alloc(L->ns + 1);
        L->litems.alloc(L->nl + 1);
        L->lrows.alloc(L->nlr + 1);
        L->lpart.alloc(L->nl + 1);
        for (const uint2& lr : lrows) L->multi |= lr.y > 1;
        if (L->ns) GGPU_HIP_TRY(hipMemcpy(L->sitems.p, sitems.data(), L->ns * sizeof(uint4), hipMemcpyHostToDevice));
        if (L->nl) GGPU_HIP_TRY(hipMemcpy(L->litems.p, litems.data(), L->nl * sizeof(uint4), hipMemcpyHostToDevice));
        if (L->nlr) GGPU_HIP_TRY(hipMemcpy(L->lrows.p, lrows.data(), L->nlr * sizeof(uint2), hipMemcpyHostToDevice));
        return L;
    }

    void build_stream(Space& S) {
        if (S.stream.size() == S.ranges.size()) return;
        S.stream.clear();
        for (auto [ra, rb] : S.ranges) S.stream.push_back(make_stream(S, ra, rb, kStreamNnz, kStreamRows, kLongChunk));
    }

    void build_wave_stream(Space& S) {
        if (!S.wstream) S.wstream = make_stream(S, 0, n, kWaveNnz, kWaveNnz, kWaveNnz);
    }

    // The TILED plan of space S for rows of in-degree >= min_indeg.
    void build_tiled(Space& S, uint32_t min_indeg) {
        if (S.tiled && S.tiled->min_indeg == min_indeg) return;
        TiledParams prm;
        prm.min_indeg = min_indeg;
        prm.tile = kTileSrc;
        prm.nnz_cap = kTileNnz;
        prm.rows_cap = kTileRows;
        prm.slots = kTileSlots;
        prm.groups = (uint32_t)std::max(ctx->cus, 1);
        const TiledPlan plan = plan_tiled(S.in, prm);
        auto T = std::make_unique<TiledLayout>();
        T->min_indeg = min_indeg;
        T->R = plan.R;
        T->K = plan.K;
        T->groups = prm.groups;
        T->dense_arcs = plan.dense_arcs;
        T->cold_arcs = plan.cold_arcs;
        T->untiled_arcs = plan.untiled_arcs;
        upload_vec(T->items, plan.items);
        upload_vec(T->tasks, plan.tasks);
        upload_vec(T->wg, plan.wg);
        upload_vec(T->segs, plan.segs);
        upload_vec(T->tci, plan.tci);
        upload_vec(T->cci, plan.cci);
        upload_vec(T->ent_row, plan.ent_row);
        upload_vec(T->ent_len, plan.ent_len);
        upload_vec(T->trow, plan.trow);
        upload_vec(T->optr, plan.optr);
        upload_vec(T->oidx, plan.oidx);
        // partials of rows without arcs from a tile stay zero: cleared once here
        const size_t np = (size_t)(plan.K + 1) * plan.R;
        T->part.alloc(std::max<size_t>(np, 1));
        T->ovf.alloc(std::max<size_t>(plan.ovf_row.size(), 1));
        GGPU_HIP_TRY(hipMemset(T->part.p, 0, std::max<size_t>(np, 1) * sizeof(float2)));
        GGPU_HIP_TRY(hipDeviceSynchronize());
        S.tiled = std::move(T);
    }

    // The HOT_TAIL layout of space S for split rows of in-degree >= min_indeg:
    // the HOT items (make_stream's wave-sized blocks), except that a split row's
    // head [rp[v], split) becomes long-row chunks flagged kLrowSplit (their
    // partials are always stored), and the tail plan with its partial array.
    void build_tail(Space& S, uint32_t min_indeg) {
        const bool dflt = !tail_begin_set && !tail_tile_set;
        const uint32_t tb = dflt ? kTailBegin : tail_begin_set;
        const uint32_t tile = dflt ? kTailTile : tail_tile_set;
        const uint32_t bt = dflt ? kBandTile : band_tile_set;
        const uint32_t be = dflt ? kTailBegin + kBandTiles * kBandTile : (bt ? band_end_set : 0u);
        if (S.tail && S.tail->min_indeg == min_indeg && S.tail->tail_begin == tb && S.tail->tile == tile &&
            S.tail->band_end == be && S.tail->band_tile == bt)
            return;
        TailParams prm;
        prm.min_indeg = min_indeg;
        prm.tail_begin = tb;
        prm.tile = tile;
        prm.nnz_cap = kWaveNnz;
        prm.pieces_cap = kWaveNnz;
        prm.band_end = be;
        prm.band_tile = bt;
        const TailPlan plan = plan_tail(S.in, prm);
        auto T = std::make_unique<TailLayout>();
        T->min_indeg = min_indeg;
        T->tail_begin = tb;
        T->tile = tile;
        T->band_end = be;
        T->band_tile = bt;
        T->R = plan.R;
        T->T = plan.T;
        T->n_items = (uint32_t)plan.items.size();
        T->tail_arcs = plan.tci.size();
        std::vector<uint4> sitems, litems;
        std::vector<uint2> lrows;
        std::vector<uint32_t> slr(plan.R, kNoHead);
        uint32_t cur0 = 0, rows = 0, nnz = 0, j = 0;
        bool multi = false;
        auto close = [&](uint32_t end) {
            if (rows) sitems.push_back(make_uint4(cur0, end, S.in.rp[cur0], S.in.rp[end]));
            rows = 0;
            nnz = 0;
        };
        // heads with every source below the LDS copy's end keep their arc ids as u16
        static_assert(kHotCap <= 65536, "head arc ids must fit in 16 bits");
        const bool head16 = tb <= std::min<uint32_t>(n, kHotCap);
        std::vector<uint16_t> h16;
        auto chunks = [&](uint32_t v, uint32_t b, uint32_t e, uint32_t flag) {
            const uint32_t first = (uint32_t)litems.size();
            const uint32_t cl = flag ? kHeadChunk : kWaveNnz;
            if (lrows.size() >= kLiHead16) throw ArgError("too many long rows for the HOT_TAIL layout");
            for (uint32_t x = b; x < e; x += cl) {
                const uint32_t xe = std::min(x + cl, e);
                if (flag && head16) {
                    if (h16.size() & 1) h16.push_back(0);  // chunks start at even offsets (32-bit pair loads)
                    const uint32_t o = (uint32_t)h16.size();
                    for (uint32_t y = x; y < xe; ++y) h16.push_back((uint16_t)S.in.ci[y]);
                    litems.push_back(make_uint4(v, o, o + (xe - x), (uint32_t)lrows.size() | kLiHead16));
                } else {
                    litems.push_back(make_uint4(v, x, xe, (uint32_t)lrows.size()));
                }
            }
            const uint32_t cnt = (uint32_t)litems.size() - first;
            lrows.push_back(make_uint2(first, cnt | flag));
            return cnt;
        };
        for (uint32_t v = 0; v < n; ++v) {
            const uint32_t b = S.in.rp[v], dg = S.in.deg(v);
            if (j < plan.R && plan.srow[j] == v) {
                close(v);
                if (plan.split[j] > b) {
                    slr[j] = (uint32_t)lrows.size();
                    chunks(v, b, plan.split[j], kLrowSplit);
                }
                ++j;
                continue;
            }
            if (dg > kWaveNnz) {
                close(v);
                multi |= chunks(v, b, b + dg, 0u) > 1;
                continue;
            }
            if (rows && (nnz + dg > kWaveNnz || rows == kWaveNnz)) close(v);
            if (!rows) cur0 = v;
            ++rows;
            nnz += dg;
        }
        close(n);
        StreamLayout& L = T->stream;
        L.r_begin = 0;
        L.r_end = n;
        L.ns = (uint32_t)sitems.size();
        L.nl = (uint32_t)litems.size();
        L.nlr = (uint32_t)lrows.size();
        L.multi = multi;
        upload_vec(L.sitems, sitems);
        upload_vec(L.litems, litems);
        upload_vec(L.lrows, lrows);
        if (h16.size() & 1) h16.push_back(0);  // a chunk's last pair load stays in the array
        upload_vec(L.h16, h16);
        L.lpart.alloc(L.nl + 1);
        // split rows with many head or overflow partials are merged by a wave
        std::vector<uint32_t> big;
        for (uint32_t r = 0; r < plan.R; ++r) {
            const uint32_t heads = slr[r] == kNoHead ? 0u : lrows[slr[r]].y & ~kLrowSplit;
            if (heads + (plan.ovptr[r + 1] - plan.ovptr[r]) > kMergeLanePartials) big.push_back(r);
        }
        T->n_big = (uint32_t)big.size();
        // band tasks: runs of <= kBandTaskItems consecutive items of one band tile
        // (items are tile-major, so the band tiles' items come first)
        std::vector<uint4> tasks;
        if (plan.R) {
            T->band_tiles = tail_band_tiles(prm, n);
            auto tile_of_item = [&](uint32_t i) { return tail_tile_of(prm, plan.tci[plan.px[plan.items[i][0]]]); };
            uint32_t i = 0;
            while (i < T->n_items && tile_of_item(i) < T->band_tiles) {
                const uint32_t t = tile_of_item(i);
                uint32_t e = i + 1;
                while (e < T->n_items && e - i < kBandTaskItems && tile_of_item(e) == t) ++e;
                tasks.push_back(make_uint4(t, i, e, 0));
                i = e;
            }
            T->tail_item0 = i;
        }
        T->n_tasks = (uint32_t)tasks.size();
        upload_vec(T->tasks, tasks);
        upload_vec(T->tci, plan.tci);
        upload_vec(T->px, plan.px);
        upload_vec(T->pdst, plan.pdst);
        upload_vec(T->items, plan.items);
        upload_vec(T->srow, plan.srow);
        upload_vec(T->slr, slr);
        upload_vec(T->ovptr, plan.ovptr);
        upload_vec(T->big, big);
        // tile partials of rows without arcs in a tile stay zero: cleared once here
        const size_t np = (size_t)plan.T * plan.R + plan.ovptr.back();
        T->n_part = np;
        T->part.alloc(std::max<size_t>(np, 1));
        GGPU_HIP_TRY(hipMemset(T->part.p, 0, std::max<size_t>(np, 1) * sizeof(float2)));
        GGPU_HIP_TRY(hipDeviceSynchronize());
        S.tail = std::move(T);
    }

    void build_bins(Space& S, uint32_t small_max, uint32_t large_min) {
        if (S.bins_ready && S.bin_small == small_max && S.bin_large == large_min) return;
        std::vector<uint32_t> lists[3];
        for (uint32_t v = 0; v < n; ++v) {
            const uint32_t d = S.in.deg(v);
            lists[d < small_max ? 0 : (d < large_min ? 1 : 2)].push_back(v);
        }
        std::vector<uint32_t> all;
        all.reserve(n);
        for (int b = 0; b < 3; ++b) {
            S.bin_count[b] = (uint32_t)lists[b].size();
            all.insert(all.end(), lists[b].begin(), lists[b].end());
        }
        S.bins.alloc(n ? n : 1);
        if (n) GGPU_HIP_TRY(hipMemcpy(S.bins.p, all.data(), 4ull * n, hipMemcpyHostToDevice));
        S.bin_small = small_max;
        S.bin_large = large_min;
        S.bins_ready = true;
    }

    void prepare_adaptive(uint32_t small_max, uint32_t large_min) { build_bins(sp[SP_GIVEN], small_max, large_min); }

    // Gauss-Seidel order (pagerank_gpu.mm:88-192): colour sets if they cover
    // >= 95% of the vertices, else natural order with pair deferral. The
    // natural order is built on `base` (the caller's or the degree order).
    void prepare_gs(uint32_t pair_k, int base_id = SP_GIVEN) {
        if (gs_ready && gs_pair_k == pair_k && (gs_colored || natural_base == base_id)) return;
        Space& g = sp[SP_GIVEN];
        if (!gs_ready || gs_pair_k != pair_k) {
            GsOrder o = gs_order(g.in, g.od.data(), pair_k);
            gs_colored = o.colored;
            gs_pair_k = pair_k;
            gs_ready = true;
            if (o.colored) {
                Space& C = sp[SP_COLOR];
                C.in = std::move(o.relabeled_in);
                C.od = std::move(o.relabeled_outdeg);
                C.id_of = std::move(o.new_id);
                C.ranges.clear();
                for (size_t i = 0; i + 1 < o.set_begin.size(); ++i)
                    C.ranges.push_back({o.set_begin[i], o.set_begin[i + 1]});
                C.stream.clear();
                C.bins_ready = false;
                upload(C);
                gs_deferred = 0;
                natural_base = -1;
                return;
            }
            natural_base = -1;
        }
        if (gs_colored) return;
        Space& B = space(base_id);
        GsOrder o = gs_order(B.in, B.od.data(), pair_k, false);
        std::vector<uint32_t> perm(n);
        uint32_t k = 0;
        for (uint32_t v : o.main_list) perm[v] = k++;
        const uint32_t split = k;
        for (uint32_t v : o.deferred_list) perm[v] = k++;
        derive(SP_NATURAL, B, perm);
        Space& N = sp[SP_NATURAL];
        N.ranges.clear();
        N.ranges.push_back({0u, split});
        if (split < n) N.ranges.push_back({split, n});
        upload(N);
        gs_deferred = o.deferred_list.size();
        natural_base = base_id;
    }

    static uint32_t blocks(uint64_t items, uint32_t per_block) {
        const uint64_t b = (items + per_block - 1) / per_block;
        return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, kSecBlocks));
    }

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

    // Stream-layout launch geometry: long-chunk blocks first, then row blocks.
    uint32_t use_stream(PrSweep& P, StreamLayout& L, uint32_t mode = 1) {
        PrStreamLayout& Y = P.stream;
        Y.sitems = L.sitems.p;
        Y.n_sitems = L.ns;
        Y.litems = L.litems.p;
        Y.n_litems = L.nl;
        Y.hci16 = L.h16.p;
        Y.lrows = L.lrows.p;
        Y.lpart = L.lpart.p;
        Y.n_lrows = L.nlr;
        P.use_stream = mode;
        P.stream_multi = L.multi;
        if (mode == 2) {  // one workgroup per CU, kHotWaves waves each
            const uint64_t want = ((uint64_t)L.ns + L.nl + kHotWaves - 1) / kHotWaves;
            return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)std::max(ctx->cus, 1)));
        }
        return std::max<uint32_t>(1, std::min<uint32_t>(L.ns + L.nl, kStreamBlocks));
    }

    PrSweep base_sweep(const Space& S, float damping) {
        PrSweep P{};
        P.rp = S.rp.p;
        P.ci = S.ci.p;
        P.outdeg = S.odd.p;
        P.partial = partial.p;
        P.partial_cap = kPartialCap;
        P.n = n;
        P.damping = damping;
        return P;
    }

    void launch(PrSweep& P, uint32_t grid, uint32_t tg) {
        if (P.nsec == 0 && !P.use_stream) return;
        if (npart + grid > kPartialCap) throw ArgError("too many blocks in one sweep");
        P.partial_base = npart;
        npart += grid;
        if (P.use_stream == 2) k_pr_sweep_hot<<<grid, kHotThreads, 0, ctx->stream>>>(P);
        else k_pr_sweep<<<grid, tg, 0, ctx->stream>>>(P);
        GGPU_HIP_TRY(hipGetLastError());
        if (P.use_stream && P.stream_multi) {
            const uint32_t g = blocks(P.stream.n_lrows, 256 / kWave);
            P.partial_base = npart;
            npart += g;
            k_pr_longfix<<<g, 256, 0, ctx->stream>>>(P);
            GGPU_HIP_TRY(hipGetLastError());
        }
    }

    // Reduces the partial slots written since the last reduction into *out.
    void reduce(PrAcc* out) {
        k_pr_reduce<<<1, 1024, 0, ctx->stream>>>(partial.p, kPartialCap, npart, out);
        GGPU_HIP_TRY(hipGetLastError());
        npart = 0;
    }

    // Sections of a Jacobi-variant sweep over all rows of S.
    uint32_t jacobi_layout(PrSweep& P, Space& S, int variant, const ggpu_pr_opts& o) {
        const uint32_t tg = o.tg;
        uint32_t grid = 0;
        if (variant == GGPU_PR_SCALAR) return add_section(P, nullptr, 0, n, SEC_LANE, 1, tg, grid);
        if (variant == GGPU_PR_VECTOR) return add_section(P, nullptr, 0, n, SEC_GROUP, kWave, tg, grid);
        if (o.layout == GGPU_PR_LAYOUT_HOT && P.contrib_in != P.contrib_out) {
            build_wave_stream(S);
            return use_stream(P, *S.wstream, 2);
        }
        if (o.layout != GGPU_PR_LAYOUT_BINS) {  // STREAM, and in-place HOT
            build_stream(S);
            return use_stream(P, *S.stream[0]);
        }
        build_bins(S, o.small_max, o.large_min);
        grid = add_section(P, S.bins.p + S.bin_count[0] + S.bin_count[1], 0, S.bin_count[2], SEC_BLOCK, 1, tg, grid);
        grid = add_section(P, S.bins.p + S.bin_count[0], 0, S.bin_count[1], SEC_GROUP, o.medium_lanes, tg, grid);
        return add_section(P, S.bins.p, 0, S.bin_count[0], SEC_LANE, 1, tg, grid);
    }

    // One Jacobi sweep k (contrib[k&1] -> contrib[(k+1)&1], r[k&1] -> r[(k+1)&1]).
    void jacobi_sweep(Space& S, int variant, const ggpu_pr_opts& o, int k, bool want_delta) {
        const int c = k & 1, nx = c ^ 1;
        PrSweep P = base_sweep(S, o.damping);
        P.contrib_in = contrib.p + (size_t)c * n;
        P.contrib_out = contrib.p + (size_t)nx * n;
        P.r_old = want_delta ? r.p + (size_t)c * n : nullptr;
        P.r_new = r.p + (size_t)nx * n;
        P.acc_in = acc.p + c;
        if (variant == GGPU_PR_ADAPTIVE && o.layout == GGPU_PR_LAYOUT_TILED) {
            // the tile sweep (partials and the untiled rows), then the merge of the tiled rows
            build_tiled(S, o.large_min);
            const TiledLayout& T = *S.tiled;
            const PrTiledDev Z = T.dev();
            const uint32_t gm = blocks(T.R, 256);
            if (npart + T.groups + gm > kPartialCap) throw ArgError("too many blocks in one sweep");
            P.partial_base = npart;
            npart += T.groups;
            k_pr_sweep_tiled<<<T.groups, kTileThreads, 0, ctx->stream>>>(P, Z);
            GGPU_HIP_TRY(hipGetLastError());
            P.partial_base = npart;
            npart += gm;
            k_pr_tiled_merge<<<gm, 256, 0, ctx->stream>>>(P, Z);
            GGPU_HIP_TRY(hipGetLastError());
            reduce(acc.p + nx);
            return;
        }
        if (variant == GGPU_PR_ADAPTIVE && o.layout == GGPU_PR_LAYOUT_HOT_TAIL) {
            // the HOT sweep (split rows' heads as partials), the tail tiles, then the merge of the split rows
            build_tail(S, o.large_min);
            TailLayout& T = *S.tail;
            launch(P, use_stream(P, T.stream, 2), o.tg);
            if (T.R) {
                const PrTailDev Z = T.dev();
                if (T.n_tasks) {
                    k_pr_band<<<T.n_tasks, kBandWaves * kWave, 0, ctx->stream>>>(P, Z, T.tasks.p, T.tail_begin,
                                                                                 T.band_tile);
                    GGPU_HIP_TRY(hipGetLastError());
                }
                if (T.n_items > T.tail_item0) {
                    const uint32_t g = (T.n_items - T.tail_item0 + kTailWaves - 1) / kTailWaves;  // one item per wave
                    k_pr_tail<<<g, kTailWaves * kWave, 0, ctx->stream>>>(P, Z, T.tail_item0);
                    GGPU_HIP_TRY(hipGetLastError());
                }
                // a lane per split row, and a wave per row of the `big` list (no wave
                // merges several hub rows one after another)
                const uint64_t want = std::max<uint64_t>((T.R + 255) / 256, ((uint64_t)T.n_big + 3) / 4);
                const uint32_t gm = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, kTailMergeBlocks));
                if (npart + gm > kPartialCap) throw ArgError("too many blocks in one sweep");
                P.partial_base = npart;
                npart += gm;
                k_pr_tail_merge<<<gm, 256, 0, ctx->stream>>>(P, Z);
                GGPU_HIP_TRY(hipGetLastError());
            }
            reduce(acc.p + nx);
            return;
        }
        const uint32_t grid = jacobi_layout(P, S, variant, o);
        launch(P, grid, o.tg);
        reduce(acc.p + nx);
    }

    // One in-place sweep (gpu_pull_gs, or a Jacobi variant with in_place) + finish.
    void inplace_sweep(Space& S, int variant, const ggpu_pr_opts& o, int k) {
        const int c = k & 1, nx = c ^ 1;
        float* rin = r.p + (size_t)c * n;
        float* rout = r.p + (size_t)nx * n;
        auto fresh = [&](uint32_t) {
            PrSweep P = base_sweep(S, o.damping);
            P.contrib_in = contrib.p;
            P.contrib_out = contrib.p;
            P.r_old = nullptr;
            P.r_new = rout;
            P.acc_in = acc.p + 0;
            return P;
        };
        const uint32_t tg = o.tg;
        if (variant == GGPU_PR_GS) {
            // ranges in order: colour sets, or natural order then deferred partners
            const bool rowblocks = o.layout != GGPU_PR_LAYOUT_BINS;  // HOT needs fixed contributions
            if (rowblocks) build_stream(S);
            for (size_t i = 0; i < S.ranges.size(); ++i) {
                PrSweep P = fresh(i == 0);
                const auto [a, b] = S.ranges[i];
                uint32_t grid;
                if (rowblocks) grid = use_stream(P, *S.stream[i]);
                else if (gs_colored) grid = add_section(P, nullptr, a, b - a, SEC_LANE, 1, tg, 0);
                else grid = add_section(P, nullptr, a, b - a, SEC_GROUP, kWave, tg, 0);
                launch(P, grid, tg);
            }
        } else {
            PrSweep P = fresh(1);
            const uint32_t grid = jacobi_layout(P, S, variant, o);
            launch(P, grid, tg);
        }
        reduce(acc.p + 2);  // sweep totals (sum for the renormalization)
        const uint32_t g = blocks(n, 256);
        k_pr_finish<<<g, 256, 0, ctx->stream>>>(rout, rin, contrib.p, S.odd.p, n, acc.p + 2, partial.p, kPartialCap);
        GGPU_HIP_TRY(hipGetLastError());
        npart = g;
        reduce(acc.p + 0);
    }

    ggpu_pr_result run(int variant, const ggpu_pr_opts& o) {
        if (o.tg < 64 || o.tg > 1024 || o.tg % 64) throw ArgError("workgroup size must be a multiple of 64 in [64, 1024]");
        if (variant < GGPU_PR_SCALAR || variant > GGPU_PR_PUSH) throw ArgError("unknown PageRank variant");
        if (variant == GGPU_PR_PUSH && o.in_place) throw ArgError("gpu_push runs Jacobi sweeps only (in_place = 0)");
        const uint32_t L = o.medium_lanes;
        if (L == 0 || L > 64 || (L & (L - 1))) throw ArgError("medium_lanes must be a power of two in [1, 64]");
        if (o.iters < 0) throw ArgError("iters must be >= 0");
        if (o.layout > GGPU_PR_LAYOUT_HOT_TAIL) throw ArgError("unknown PageRank layout");
        if (o.layout == GGPU_PR_LAYOUT_TILED && (o.large_min == 0 || o.large_min > kTileNnz))
            throw ArgError("TILED needs 1 <= large_min <= " + std::to_string(kTileNnz) + " (the in-degree of a tiled row)");
        if (o.order > GGPU_PR_ORDER_DEGREE) throw ArgError("unknown PageRank vertex order");
        const int base_id = o.order == GGPU_PR_ORDER_DEGREE ? SP_DEGREE : SP_GIVEN;
        int sid = base_id;
        if (variant == GGPU_PR_GS) {
            prepare_gs(gs_ready ? gs_pair_k : 2, base_id);
            sid = gs_colored ? SP_COLOR : SP_NATURAL;
        }
        Space& S = space(sid);
        const bool inplace = variant == GGPU_PR_GS || o.in_place;
        ggpu_pr_result res{};
        res.last_delta = -1;
        res.last_max_rel = -1;
        have_result = false;
        result_space = sid;
        if (n == 0) return res;
        hipStream_t s = ctx->stream;
        const bool tol_mode = o.tol > 0;
        GGPU_HIP_TRY(hipEventRecord(event(0), s));
        npart = 0;
        if (variant == GGPU_PR_PUSH) {
            ensure_out(S);
            pacc.alloc(n);
            GGPU_HIP_TRY(hipMemsetAsync(pacc.p, 0, 8ull * n, s));
        }
        const uint32_t gi = blocks(n, 256);
        k_pr_init<<<gi, 256, 0, s>>>(r.p, contrib.p, S.odd.p, n, partial.p, kPartialCap);
        GGPU_HIP_TRY(hipGetLastError());
        npart = gi;
        reduce(acc.p + 0);
        auto* pin = (PrAcc*)ctx_pinned(ctx, 4096);
        double best = 1e300;
        int since_best = 0;
        int k = 0;
        for (; k < o.iters; ++k) {
            const PrAcc* out;
            if (inplace) {
                inplace_sweep(S, variant, o, k);
                out = acc.p + 0;
            } else {
                if (variant == GGPU_PR_PUSH) push_sweep(S, o, k, tol_mode);
                else jacobi_sweep(S, variant, o, k, tol_mode);
                out = acc.p + ((k + 1) & 1);
            }
            if (!tol_mode) continue;
            // tol mode: one synchronization per sweep to read the FP64 change
            GGPU_HIP_TRY(hipMemcpyAsync(pin, out, sizeof(PrAcc), hipMemcpyDeviceToHost, s));
            GGPU_HIP_TRY(hipStreamSynchronize(s));
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
        result_slot = k & 1;  // sweep k writes r[(k+1)&1]
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
        const Space& S = sp[result_space];
        if (S.id_of.empty()) {
            GGPU_HIP_TRY(hipMemcpy(out, src, 4ull * n, hipMemcpyDeviceToHost));
            return;
        }
        std::vector<float> tmp(n);
        GGPU_HIP_TRY(hipMemcpy(tmp.data(), src, 4ull * n, hipMemcpyDeviceToHost));
        for (uint32_t v = 0; v < n; ++v) out[v] = tmp[S.id_of[v]];
    }

    // Times `iters` Jacobi sweeps of gpu_pull_adaptive per candidate layout /
    // order / bin shape and keeps the fastest (ranks do not depend on it).
    void autotune(int iters, ggpu_pr_opts& o) {
        struct Cand { uint32_t layout, order, sm, lg, L; };
        std::vector<Cand> cands = {{GGPU_PR_LAYOUT_STREAM, GGPU_PR_ORDER_GIVEN, 0, 0, 64},
                                   {GGPU_PR_LAYOUT_STREAM, GGPU_PR_ORDER_DEGREE, 0, 0, 64},
                                   {GGPU_PR_LAYOUT_HOT, GGPU_PR_ORDER_GIVEN, 0, 0, 64},
                                   {GGPU_PR_LAYOUT_HOT, GGPU_PR_ORDER_DEGREE, 0, 0, 64},
                                   {GGPU_PR_LAYOUT_TILED, GGPU_PR_ORDER_DEGREE, 0, 64, 64},
                                   {GGPU_PR_LAYOUT_TILED, GGPU_PR_ORDER_DEGREE, 0, 128, 64},
                                   {GGPU_PR_LAYOUT_TILED, GGPU_PR_ORDER_DEGREE, 0, 256, 64},
                                   {GGPU_PR_LAYOUT_HOT_TAIL, GGPU_PR_ORDER_DEGREE, 0, 64, 64},
                                   {GGPU_PR_LAYOUT_HOT_TAIL, GGPU_PR_ORDER_DEGREE, 0, 256, 64}};
        for (uint32_t order : {(uint32_t)GGPU_PR_ORDER_GIVEN, (uint32_t)GGPU_PR_ORDER_DEGREE})
            for (uint32_t sm : {4u, 16u})
                for (uint32_t lg : {256u, 4096u})
                    for (uint32_t L : {16u, 64u}) cands.push_back({GGPU_PR_LAYOUT_BINS, order, sm, lg, L});
        float best = 1e30f;
        ggpu_pr_opts bo = o;
        for (const Cand& c : cands) {
            ggpu_pr_opts t = o;
            t.layout = c.layout;
            t.order = c.order;
            if (c.layout == GGPU_PR_LAYOUT_BINS) {
                t.small_max = c.sm;
                t.large_min = c.lg;
                t.medium_lanes = c.L;
            } else if (c.layout == GGPU_PR_LAYOUT_TILED || c.layout == GGPU_PR_LAYOUT_HOT_TAIL) {
                t.large_min = c.lg;
            }
            t.iters = iters;
            t.tol = 0;
            t.in_place = 0;
            run(GGPU_PR_ADAPTIVE, t);  // warm-up (also builds the layout)
            const ggpu_pr_result rr = run(GGPU_PR_ADAPTIVE, t);
            if (rr.gpu_ms < best) {
                best = (float)rr.gpu_ms;
                bo = t;
            }
        }
        o.small_max = bo.small_max;
        o.large_min = bo.large_min;
        o.medium_lanes = bo.medium_lanes;
        o.layout = bo.layout;
        o.order = bo.order;
    }
};

}  // namespace ggpu

using namespace ggpu;

struct ggpu_pr {
    ggpu_ctx* ctx;
    PrEngine* eng;
};

namespace ggpu {
PprView ppr_view(PrEngine& e, uint32_t order) {
    if (order > GGPU_PR_ORDER_DEGREE) throw ArgError("unknown vertex order");
    const int id = order == GGPU_PR_ORDER_DEGREE ? SP_DEGREE : SP_GIVEN;
    const Space& g = e.space(id);
    return PprView{e.ctx, e.n, g.rp.p, g.ci.p, g.odd.p, g.od.data(), g.in.rp.data(),
                   g.id_of.empty() ? nullptr : g.id_of.data(), id};
}
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
    o->layout = GGPU_PR_LAYOUT_STREAM;
    o->order = GGPU_PR_ORDER_GIVEN;
}

int ggpu_pr_parse_variant(const char* name, int* variant) {
    if (!name || !variant) return GGPU_EINVAL;
    const std::string s = name;
    if (s == "gpu_pull_scalar") *variant = GGPU_PR_SCALAR;
    else if (s == "gpu_pull_vector") *variant = GGPU_PR_VECTOR;
    else if (s == "gpu_pull_adaptive") *variant = GGPU_PR_ADAPTIVE;
    else if (s == "gpu_pull_gs") *variant = GGPU_PR_GS;
    else if (s == "gpu_push") *variant = GGPU_PR_PUSH;
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
    if (colored) *colored = e.gs_ready && e.gs_colored;
    if (sets) *sets = e.gs_ready && e.gs_colored ? e.sp[SP_COLOR].ranges.size() : 0;
    if (deferred) *deferred = e.gs_deferred;
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

int ggpu_ppr_autotune(ggpu_pr* p, const uint32_t* seeds, uint32_t B, int iters, ggpu_pr_opts* opt) {
    if (!p || !seeds || !opt) return GGPU_EINVAL;
    return guarded(p->ctx, [&] {
        GGPU_HIP_TRY(hipSetDevice(p->ctx->device));
        ppr_autotune(*p->eng, seeds, B, iters, *opt);
    });
}

int ggpu_ppr_fetch(ggpu_pr* p, float* out) {
    if (!p || !out) return GGPU_EINVAL;
    return guarded(p->ctx, [&] { ppr_fetch(*p->eng, out); });
}

int ggpu_test_pr_tail_geometry(ggpu_pr* p, uint32_t tail_begin, uint32_t tile, uint32_t band_end,
                               uint32_t band_tile) {
    if (!p) return GGPU_EINVAL;
    const bool dflt = !tail_begin && !tile;
    if (!dflt && (!tile || band_tile > kBandTile || (band_tile && band_end <= tail_begin))) return GGPU_EINVAL;
    p->eng->tail_begin_set = tail_begin;
    p->eng->tail_tile_set = tile;
    p->eng->band_end_set = band_end;
    p->eng->band_tile_set = band_tile;
    return GGPU_OK;
}

int ggpu_test_pr_tail_info(const ggpu_pr* p, uint32_t order, uint64_t* stats) {
    if (!p || !stats || order > GGPU_PR_ORDER_DEGREE) return GGPU_EINVAL;
    const Space& S = p->eng->sp[order == GGPU_PR_ORDER_DEGREE ? SP_DEGREE : SP_GIVEN];
    for (int i = 0; i < 8; ++i) stats[i] = 0;
    if (!S.tail) return GGPU_OK;
    const TailLayout& T = *S.tail;
    stats[0] = T.R;
    stats[1] = T.T;
    stats[2] = T.tail_arcs;
    stats[3] = T.n_items;
    stats[4] = T.n_big;
    stats[5] = T.stream.nl;
    stats[6] = T.band_tiles;
    stats[7] = T.n_tasks;
    return GGPU_OK;
}

}  // extern "C"
