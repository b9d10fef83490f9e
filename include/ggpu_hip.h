/* ggpu_hip.h -- C-ABI of the MI355X (gfx950) BFS / PageRank backend.
 *
 * This is the drop-in boundary that replaces the reference's Metal host layer:
 *   gpu::MetalContext   src/metal/context.hpp:21-59      -> ggpu_ctx_*
 *   gpu::GpuBfs         src/metal/bfs_gpu.hpp:35-105     -> ggpu_bfs_*
 *   gpu::GpuBfsOptions  src/metal/bfs_gpu.hpp:23-33      -> ggpu_bfs_opts
 *   gpu::GpuPageRank    src/metal/pagerank_gpu.hpp:46-127-> ggpu_pr_*
 *   gpu::GpuPrOptions   src/metal/pagerank_gpu.hpp:27-36 -> ggpu_pr_opts
 *   gpu::GpuPrResult    src/metal/pagerank_gpu.hpp:38-44 -> ggpu_pr_result
 * plus entry points with no reference counterpart: batched multi-source BFS,
 * personalized PageRank blocks, the learned direction policy and the
 * autotuner (each says so below).
 *
 * Conventions: plain pointers and sizes only; every call returns a GGPU_*
 * status (0 = ok) and ggpu_last_error(ctx) describes the last failure on
 * that context. Inputs are copied to device memory (HBM) at create time;
 * results stay on the device until fetched, so result transfers are outside
 * any timed run. One host thread drives one context. */
#ifndef GGPU_HIP_H
#define GGPU_HIP_H

#include "ggpu_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ggpu_ctx ggpu_ctx;
typedef struct ggpu_bfs ggpu_bfs;
typedef struct ggpu_pr ggpu_pr;

/* ------------------------------------------------------------- context */
/* MetalContext() (context.mm:31-54): picks device `device`, creates one stream. */
int ggpu_ctx_create(int device, ggpu_ctx** out);
void ggpu_ctx_destroy(ggpu_ctx* ctx);
const char* ggpu_last_error(const ggpu_ctx* ctx);
/* Device name, compute-unit count and HBM bytes. */
int ggpu_device_info(ggpu_ctx* ctx, char* name, size_t name_len, int* cus, uint64_t* hbm_bytes);
/* Block until all work on the context's stream is done. */
int ggpu_ctx_sync(ggpu_ctx* ctx);

/* ----------------------------------------------------------------- BFS */
/* Variants (bfs_gpu.hpp:48-66). */
#define GGPU_BFS_EDGE 0   /* gpu_edge:   degree scan + edge-balanced top-down every level */
#define GGPU_BFS_BINNED 1 /* gpu_binned: thread / wave / workgroup per frontier vertex by degree */
#define GGPU_BFS_DO 2     /* gpu_do:     direction-optimizing (top-down edge-balanced + bitmap bottom-up) */
#define GGPU_BFS_HYBRID 3 /* hybrid:     gpu_do with levels whose work < hybrid_t on the CPU */
#define GGPU_BFS_VERTEX 4 /* gpu_vertex: lane per vertex every level, plain stores (bfs.metal:45-79) */
#define GGPU_BFS_NAIVE 5  /* gpu_naive:  lane per queued vertex, CAS claims, one append atomic per claim (bfs.metal:95-126) */

/* How the level loop is driven (no reference counterpart except HOST). */
#define GGPU_LOOP_HOST 0       /* one host synchronization per level (the reference protocol) */
#define GGPU_LOOP_GRAPH 1      /* device-side decisions; K levels per captured hipGraph launch */
#define GGPU_LOOP_PERSISTENT 2 /* one persistent kernel (a workgroup per CU), device-wide barrier between phases */

/* Who picks each level's action (top-down edge / top-down binned / bottom-up). */
#define GGPU_POLICY_HEURISTIC 0 /* the variant's rule: GAP alpha/beta for gpu_do */
#define GGPU_POLICY_LEARNED 1   /* the MLP set with ggpu_policy_set (heuristic if none) */
#define GGPU_POLICY_SCRIPTED 2  /* actions[level % n_actions] (tests, training data) */

#define GGPU_ACTION_TD_EDGE 0
#define GGPU_ACTION_TD_BINNED 1
#define GGPU_ACTION_BU 2

/* Vertex numbering the batched BFS traverses in (no reference counterpart;
 * depths always come back in the caller's ids). */
#define GGPU_BATCH_ORDER_GIVEN 0   /* the caller's numbering */
#define GGPU_BATCH_ORDER_DEGREE 1  /* renumbered by descending degree (reorder "degree") on the first run */
#define GGPU_BATCH_ORDER_AUTO 2    /* DEGREE when the maximum degree is >= 8x the mean, else GIVEN */

typedef struct ggpu_bfs_opts {
    uint32_t tg;         /* workgroup size of the level kernels: multiple of 64 in [64, 1024] */
    uint32_t small_max;  /* binned: degree <  small_max -> lane per vertex (wave-cooperative gather) */
    uint32_t large_min;  /* binned: degree >= large_min -> workgroup per vertex */
    uint32_t ept;        /* edge-balanced top-down: edges per thread (>= 1) */
    double alpha;        /* gpu_do: go bottom-up if m_f > m_u / alpha */
    double beta;         /* gpu_do: back to top-down when the frontier shrinks below n / beta */
    uint64_t hybrid_t;   /* hybrid: levels with work < hybrid_t run on the CPU */
    int cpu_threads;     /* hybrid: OpenMP threads of CPU levels (0 = all cores) */
    uint32_t loop;       /* GGPU_LOOP_* (HOST for hybrid) */
    uint32_t policy;     /* GGPU_POLICY_* (gpu_do / batched BFS only) */
    const uint8_t* actions; /* GGPU_POLICY_SCRIPTED: per-level actions, cycled */
    uint32_t n_actions;
    uint32_t graph_levels;  /* GGPU_LOOP_GRAPH: levels captured per graph launch */
    uint32_t time_levels;   /* HOST loop: 1 = record per-level GPU time with events */
    uint32_t batch_order;   /* batched BFS: GGPU_BATCH_ORDER_* (default AUTO) */
} ggpu_bfs_opts;
void ggpu_bfs_opts_default(ggpu_bfs_opts* o);
/* "gpu_edge" | "gpu_binned" | "gpu_do" | "hybrid" | "gpu_vertex" | "gpu_naive" -> GGPU_BFS_*
 * gpu_vertex and gpu_naive are top-down only and run with the host loop and
 * the heuristic policy (runVertex / runNaive, bfs_gpu.mm:115-175). */
int ggpu_bfs_parse_variant(const char* name, int* variant);

/* GpuBfs(ctx, g) (bfs_gpu.mm:45-74): copies the symmetric CSR to HBM. */
int ggpu_bfs_create(ggpu_ctx* ctx, uint64_t n, uint64_t m, const uint32_t* row_ptr, const uint32_t* col_idx,
                    ggpu_bfs** out);
void ggpu_bfs_destroy(ggpu_bfs* b);
/* runEdge / runBinned / runDO / runHybrid (bfs_gpu.hpp:48-66). The result
 * stays on the device (fetch with ggpu_bfs_fetch_dist). trace (optional)
 * receives up to trace_cap levels; *trace_len the level count. */
int ggpu_bfs_run(ggpu_bfs* b, int variant, uint32_t source, const ggpu_bfs_opts* opt, ggpu_level* trace,
                 size_t trace_cap, size_t* trace_len);
/* dist() (bfs_gpu.hpp:40): n uint32 distances, GGPU_INF if unreached. */
int ggpu_bfs_fetch_dist(ggpu_bfs* b, uint32_t* out);
/* lastGpuMs() / lastLevels() (bfs_gpu.hpp:68-72). */
double ggpu_bfs_last_gpu_ms(const ggpu_bfs* b);
uint32_t ggpu_bfs_last_levels(const ggpu_bfs* b);

/* ------------------------------------------------ BFS without allocation
 * No reference counterpart. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): one gpu_do traversal from `source` with the persistent device
 * loop, over a symmetric CSR that the caller already holds in device memory
 * (d_row_ptr: n + 1 offsets, d_col_idx: m vertex ids < n; not checked),
 * writing the distances into the caller's d_dist (n entries, GGPU_INF if
 * unreached); *levels (if not NULL) = the number of levels, as
 * ggpu_bfs_last_levels reports. The distances and the per-level decisions
 * equal ggpu_bfs_run(GGPU_BFS_DO, loop = GGPU_LOOP_PERSISTENT)'s. The call
 * allocates and frees nothing: every working buffer (visited and frontier
 * bitmaps, level counters, loop state, arrival counters) lives in one fixed
 * array of GGPU_NOALLOC_SCRATCH_BYTES of device memory inside the library,
 * which each GPU it runs on holds until the process ends; n above
 * GGPU_NOALLOC_MAX_N is rejected (GGPU_EINVAL). Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (two host threads, two
 * contexts, two streams) race and must not be made. The call runs on the
 * context's stream and returns when the traversal has finished. opt (NULL =
 * defaults): alpha, beta, small_max, large_min and the heuristic or learned
 * policy are used, `loop` is ignored, the scripted policy is rejected. */
#define GGPU_NOALLOC_MAX_N (1u << 24)
#define GGPU_NOALLOC_SCRATCH_BYTES 8406528u
int ggpu_bfs_run_noalloc(ggpu_ctx* ctx, uint64_t n, uint64_t m, const uint32_t* d_row_ptr,
                         const uint32_t* d_col_idx, uint32_t source, const ggpu_bfs_opts* opt, uint32_t* d_dist,
                         uint32_t* levels);

/* --------------------------------------------- batched multi-source BFS
 * No reference counterpart. k sources (1 <= k <= 256) traverse together as
 * per-vertex bit masks (one bit per source, ceil(k/64) u64 words per vertex);
 * each source's depths equal a single-source BFS from it. */
typedef struct ggpu_msbfs_result {
    uint32_t levels;          /* levels of the batch (max over sources + 1) */
    uint32_t pad;
    double gpu_ms;            /* device time of the whole batch (events around every launch of it) */
    uint64_t level_actions[3];/* levels run per action */
    double setup_ms;          /* device clock: launch start -> level 0 ready */
    double levels_ms;         /* device clock: the level loop (summed over a renumbered run's launches) */
    double finish_ms;         /* events: the depth pass kernels (depths, or depth records and their
                                 un-permute for a renumbered run) and state cleared, over all windows */
} ggpu_msbfs_result;
/* opt->policy picks the per-level action (HEURISTIC: GAP rule on batch counts;
 * LEARNED: the MLP; SCRIPTED: opt->actions). */
int ggpu_msbfs_run(ggpu_bfs* b, const uint32_t* sources, uint32_t k, const ggpu_bfs_opts* opt,
                   ggpu_msbfs_result* res, ggpu_level* trace, size_t trace_cap, size_t* trace_len);
/* k*n uint32, source-major: out[s*n + v] = depth of v from sources[s]. */
int ggpu_msbfs_fetch_dist(ggpu_bfs* b, uint32_t* out);

/* ------------------------------------------------------------ PageRank */
#define GGPU_PR_SCALAR 1   /* gpu_pull_scalar:   thread per vertex */
#define GGPU_PR_VECTOR 2   /* gpu_pull_vector:   wave per vertex */
#define GGPU_PR_ADAPTIVE 3 /* gpu_pull_adaptive: vertices binned by in-degree (lane / wave / workgroup) */
#define GGPU_PR_GS 4       /* gpu_pull_gs:       in-place sweeps in colour or natural order */
#define GGPU_PR_PUSH 5     /* gpu_push:          float atomics along out-arcs (the transpose of the in-CSR), Jacobi only */

typedef struct ggpu_pr_opts {
    float damping;         /* 0.85 */
    int iters;             /* fixed mode: sweeps; tol mode: cap */
    double tol;            /* > 0: stop when the FP64 L1 change of a sweep < tol */
    double tol_rel;        /* > 0 (tol mode): also require max |change| / max(r, 1/n) < tol_rel */
    uint32_t tg;           /* workgroup size: multiple of 64 in [64, 1024] */
    uint32_t small_max;    /* adaptive: in-degree <  small_max -> lane per vertex */
    uint32_t large_min;    /* adaptive: in-degree >= large_min -> workgroup per vertex */
    int in_place;          /* Jacobi variants: update contributions in place (Gauss-Seidel style) */
    uint32_t medium_lanes; /* adaptive bins: lanes per medium vertex (power of two <= 64) */
    uint32_t layout;       /* gpu_pull_adaptive work layout: GGPU_PR_LAYOUT_* */
    uint32_t order;        /* vertex numbering the sweeps run in: GGPU_PR_ORDER_* (ranks come back in caller ids) */
} ggpu_pr_opts;
/* GIVEN: the caller's ids. DEGREE: an internal copy relabeled by descending
 * out-degree, so the most-gathered contributions share cache lines. */
#define GGPU_PR_ORDER_GIVEN 0
#define GGPU_PR_ORDER_DEGREE 1
/* STREAM: consecutive short in-lists packed into ~2K-edge row blocks whose
 * gathers are staged in LDS, long lists split into chunks (default).
 * BINS: the reference's shape, one lane / lane group / workgroup per vertex by
 * in-degree (small_max, large_min, medium_lanes). */
#define GGPU_PR_LAYOUT_STREAM 0
#define GGPU_PR_LAYOUT_BINS 1
/* HOT: one 1024-lane workgroup per compute unit for the whole sweep, the
 * contributions of vertices [0, 30720) held in LDS (with ORDER_DEGREE those
 * are the most gathered ones), wave-sized row blocks of <= 512 entries.
 * Jacobi sweeps only: in-place sweeps with this layout use STREAM. */
#define GGPU_PR_LAYOUT_HOT 2
/* TILED: rows of in-degree >= large_min (1..1024) are summed per tile of 32768
 * consecutive sources whose contributions are held in LDS, one compensated
 * partial per (tile, row), merged per row in tile order by a second kernel;
 * the other rows as in HOT. Built for ORDER_DEGREE (hub rows and hot sources
 * first). Jacobi sweeps only: in-place sweeps with this layout use STREAM. */
#define GGPU_PR_LAYOUT_TILED 3
/* HOT_TAIL: HOT, except that rows of in-degree >= large_min leave their in-arcs
 * from tail sources (ids >= 2^19, past the LDS copy and an L2-sized band) to a
 * second kernel that sums them per tile of 2^19 sources in tile order (the tile
 * being read stays in L2), one compensated partial per (tile, row); a merge adds
 * each such row's head, tile and overflow partials in a fixed order. Built for
 * ORDER_DEGREE. Jacobi sweeps only: in-place sweeps with this layout use STREAM. */
#define GGPU_PR_LAYOUT_HOT_TAIL 4
void ggpu_pr_opts_default(ggpu_pr_opts* o);
/* "gpu_pull_scalar" | "gpu_pull_vector" | "gpu_pull_adaptive" | "gpu_pull_gs" | "gpu_push" -> GGPU_PR_* */
int ggpu_pr_parse_variant(const char* name, int* variant);

typedef struct ggpu_pr_result {
    int iters;            /* sweeps performed */
    int floor_stop;       /* tol mode: stopped because the L1 change stagnated (5 sweeps without a new best) */
    double last_delta;    /* FP64 L1 change of the last sweep (tol mode) */
    double last_max_rel;  /* max relative change of the last sweep (tol mode) */
    double gpu_ms;        /* device time of the run */
    double dangling;      /* FP64 dangling mass entering the last sweep */
    double sum;           /* FP64 sum of the ranks after the last sweep */
} ggpu_pr_result;

/* GpuPageRank(ctx, out, in, outdeg, false, true) (pagerank_gpu.mm:30-68):
 * in-CSR (transpose; the graph itself when undirected) + out-degrees. */
int ggpu_pr_create(ggpu_ctx* ctx, uint64_t n, uint64_t m, const uint32_t* in_row_ptr, const uint32_t* in_col_idx,
                   const uint32_t* outdeg, ggpu_pr** out);
void ggpu_pr_destroy(ggpu_pr* p);
/* prepareAdaptive(small, large) (pagerank_gpu.mm:70-86). */
int ggpu_pr_prepare_adaptive(ggpu_pr* p, uint32_t small_max, uint32_t large_min);
/* prepareGS(pairMaxOutdeg) (pagerank_gpu.mm:88-192); gsColored/gsSets/gsDeferred. */
int ggpu_pr_prepare_gs(ggpu_pr* p, uint32_t pair_max_outdeg);
int ggpu_pr_gs_info(const ggpu_pr* p, int* colored, uint64_t* sets, uint64_t* deferred);
/* run(variant, opts) (pagerank_gpu.mm:336-407). */
int ggpu_pr_run(ggpu_pr* p, int variant, const ggpu_pr_opts* opt, ggpu_pr_result* res);
/* ranks() (pagerank_gpu.mm:194-203): n floats in the caller's vertex ids. */
int ggpu_pr_fetch_ranks(ggpu_pr* p, float* out);
/* Autotuner (no reference counterpart): times candidate (small_max,
 * large_min, medium_lanes) on this graph for `iters` sweeps each and writes
 * the fastest into opt (ranks are unaffected by the choice). */
int ggpu_pr_autotune(ggpu_pr* p, int iters, ggpu_pr_opts* opt);

/* Personalized PageRank block (no reference counterpart): B <= 256 one-hot
 * restart vectors (seeds[b]) iterate together as an n x B FP32 block. Restart
 * and dangling mass of column b both return to seeds[b]. Work assignment from
 * opt: in-lists of >= large_min entries (>= 64) are cut into large_min-entry
 * chunks, a wave each; shorter ones get medium_lanes lanes per vertex
 * (clamped up to min(64, B rounded to a power of two)); tg workgroups. */
int ggpu_ppr_run(ggpu_pr* p, const uint32_t* seeds, uint32_t B, const ggpu_pr_opts* opt, ggpu_pr_result* res);
/* Autotuner of the block (no reference counterpart): times (large_min,
 * medium_lanes, tg) candidates for `iters` sweeps of this seed block and
 * writes the fastest into opt; the block holds the last candidate's ranks. */
int ggpu_ppr_autotune(ggpu_pr* p, const uint32_t* seeds, uint32_t B, int iters, ggpu_pr_opts* opt);
/* n*B floats, vertex-major: out[v*B + b]. */
int ggpu_ppr_fetch(ggpu_pr* p, float* out);

/* ------------------------------------------------------ learned policy
 * No reference counterpart. MLP 5 -> 32 -> 32 -> 3 (ReLU), FP32; 1,347
 * parameters laid out W1[32][5] b1[32] W2[32][32] b2[32] W3[3][32] b3[3].
 * Inputs are the per-level features of ggpu_policy_features; outputs are
 * predicted log-times (ln ms) of the actions TD_EDGE, TD_BINNED, BU; the
 * policy takes the argmin. A context holds one parameter set per slot:
 * single-source traversals (gpu_do) and batched traversals have different
 * costs per action. With no set for a slot, LEARNED falls back to the
 * heuristic. */
#define GGPU_MLP_PARAMS 1347
#define GGPU_POLICY_SINGLE 0   /* gpu_do */
#define GGPU_POLICY_BATCHED 1  /* ggpu_msbfs_run */
int ggpu_policy_set(ggpu_ctx* ctx, int slot, const float* params, uint32_t count);
int ggpu_policy_clear(ggpu_ctx* ctx, int slot);
int ggpu_policy_get(ggpu_ctx* ctx, int slot, float* params, uint32_t count, int* is_set);
/* features[5] of one level: log2(1+n_f)/32, log2(1+m_f)/32, log2(1+m_u)/32,
 * level/64, log2(1+mean degree)/16. */
void ggpu_policy_features(uint64_t n_f, uint64_t m_f, uint64_t m_u, uint32_t level, double mean_degree,
                          float* features);
/* GPU forward pass: y[count][3]. */
int ggpu_mlp_forward(ggpu_ctx* ctx, const float* params, const float* x, uint64_t count, float* y);
/* GPU forward + backward: loss and dLoss/dparams (MSE over entries with mask != 0). */
int ggpu_mlp_loss_grad(ggpu_ctx* ctx, const float* params, const float* x, const float* t, const uint8_t* mask,
                       uint64_t count, float* grad, float* loss);
/* GPU training: `epochs` full-batch Adam steps from params_inout. */
int ggpu_mlp_train(ggpu_ctx* ctx, float* params_inout, const float* x, const float* t, const uint8_t* mask,
                   uint64_t count, uint32_t epochs, float lr, float* final_loss);
/* GPU training: `epochs` full-batch SGD steps with momentum in [0, 1)
 * (v = momentum * v + grad, params -= lr * v; momentum 0 is plain SGD). */
int ggpu_mlp_train_sgd(ggpu_ctx* ctx, float* params_inout, const float* x, const float* t, const uint8_t* mask,
                       uint64_t count, uint32_t epochs, float lr, float momentum, float* final_loss);
/* Training data. The frontier and the counts of a level do not depend on
 * the actions taken before it, so the traversal is run once per action with
 * that action at every level: level l's three times are the three actions'
 * costs from the same state. Writes per level the 5 features (as the policy
 * sees them), the 3 times (ms, features[l*5..], times_ms[l*3..]) and the
 * trace record of the first run; *len = levels (at most cap are written).
 * Single source (gpu_do's level loop, device time per level from events): */
int ggpu_bfs_record_levels(ggpu_bfs* b, uint32_t source, const ggpu_bfs_opts* opt, float* features,
                           float* times_ms, ggpu_level* trace, size_t cap, size_t* len);
/* Batched (one ggpu_msbfs_run per action; times from the device clock): */
int ggpu_msbfs_record_levels(ggpu_bfs* b, const uint32_t* sources, uint32_t k, const ggpu_bfs_opts* opt,
                             float* features, float* times_ms, ggpu_level* trace, size_t cap, size_t* len);

#ifdef __cplusplus
}
#endif
#endif
