// BFS engine state shared by the host-driven loop (bfs.hip), the device-driven
// loops (bfs_loop.hip), the batched BFS (msbfs.hip) and the policy recorder.
#pragma once

#include <cmath>
#include <vector>

#include "bfs_kernels.hpp"
#include "graph.hpp"

namespace ggpu {

// Level-loop state: the scalars of runDirOpt (src/metal/bfs_gpu.mm:451-455)
// plus the frontier format, the action of the current level and exact m_f /
// m_u for the learned policy's features. Lives on the host for the host loop
// and in device memory for the device loops.
struct LoopState {
    uint64_t size;             // frontier vertices
    uint64_t total;            // frontier edges (sum of frontier degrees, exact)
    uint64_t scout;            // GAP m_f (reset to 1 after a bottom-up phase)
    uint64_t to_check;         // GAP m_u (edges not yet examined top-down)
    uint64_t unvisited_edges;  // exact sum of unvisited degrees
    uint64_t bins[3];          // binned frontier sizes
    uint32_t level;
    uint32_t cur;              // ping-pong index of queues / bins / bitmaps
    uint32_t fmt;              // Fmt of the current frontier
    uint32_t action;           // GGPU_ACTION_* of the current level
    uint32_t bottom_up;        // GAP direction flag
    uint32_t visited_fresh;    // device visited bitmap matches dist
    uint32_t done;
    uint32_t pad;
};

struct LevelOut {
    uint64_t found = 0, examined = 0, next_edges = 0;
    uint64_t bins[3] = {0, 0, 0};
};

// Where decide() finds the policy inputs (host or device pointers).
struct PolicyView {
    uint32_t mode;        // GGPU_POLICY_*
    uint32_t variant;     // GGPU_BFS_*
    double alpha, beta;
    const uint8_t* actions;
    uint32_t n_actions;
    const float* mlp;     // GGPU_MLP_PARAMS floats, or null (-> heuristic)
    double mean_degree;
};

__host__ __device__ inline uint32_t bin_of(uint32_t deg, uint32_t small_max, uint32_t large_min) {
    return deg < small_max ? 0u : (deg < large_min ? 1u : 2u);
}

__host__ __device__ inline void policy_features(uint64_t n_f, uint64_t m_f, uint64_t m_u, uint32_t level,
                                                double mean_degree, float* f) {
    f[0] = (float)(log2(1.0 + (double)n_f) / 32.0);
    f[1] = (float)(log2(1.0 + (double)m_f) / 32.0);
    f[2] = (float)(log2(1.0 + (double)m_u) / 32.0);
    f[3] = (float)level / 64.0f;
    f[4] = (float)(log2(1.0 + mean_degree) / 16.0);
}

// MLP forward 5 -> 32 -> 32 -> 3 (ReLU) for one sample, FP32.
__host__ __device__ inline void mlp_forward1(const float* p, const float* x, float* y) {
    const float* W1 = p;
    const float* b1 = W1 + 32 * 5;
    const float* W2 = b1 + 32;
    const float* b2 = W2 + 32 * 32;
    const float* W3 = b2 + 32;
    const float* b3 = W3 + 3 * 32;
    float h1[32], h2[32];
    for (int j = 0; j < 32; ++j) {
        float a = b1[j];
        for (int i = 0; i < 5; ++i) a += W1[j * 5 + i] * x[i];
        h1[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < 32; ++j) {
        float a = b2[j];
        for (int i = 0; i < 32; ++i) a += W2[j * 32 + i] * h1[i];
        h2[j] = a > 0.f ? a : 0.f;
    }
    for (int k = 0; k < 3; ++k) {
        float a = b3[k];
        for (int i = 0; i < 32; ++i) a += W3[k * 32 + i] * h2[i];
        y[k] = a;
    }
}

// The same forward pass on the 64 lanes of one wave: lane j computes hidden
// unit j of each layer and lane k output k, every sum in mlp_forward1's order
// (so the results are bit-identical); the other layer's values arrive by
// shuffles. All 64 lanes must call it; every lane returns y[0..2].
__device__ inline void mlp_forward_wave(const float* p, const float* x, float* y) {
    const float* W1 = p;
    const float* b1 = W1 + 32 * 5;
    const float* W2 = b1 + 32;
    const float* b2 = W2 + 32 * 32;
    const float* W3 = b2 + 32;
    const float* b3 = W3 + 3 * 32;
    const int lane = (int)__lane_id();
    const int j = lane & 31;
    float a = b1[j];
    for (int i = 0; i < 5; ++i) a += W1[j * 5 + i] * x[i];
    const float h1 = a > 0.f ? a : 0.f;
    a = b2[j];
    for (int i = 0; i < 32; ++i) a += W2[j * 32 + i] * __shfl(h1, i, 64);
    const float h2 = a > 0.f ? a : 0.f;
    const int k = lane < 3 ? lane : 0;
    a = b3[k];
    for (int i = 0; i < 32; ++i) a += W3[k * 32 + i] * __shfl(h2, i, 64);
    for (int q = 0; q < 3; ++q) y[q] = __shfl(a, q, 64);
}

// Picks the action of the level about to run (state after the previous level).
// y_pre: the MLP's outputs for this state when the caller already evaluated
// them (the device loops do, on one wave); otherwise computed here.
__host__ __device__ inline void decide_action(LoopState& st, const PolicyView& pv, uint64_t n,
                                              const float* y_pre = nullptr) {
    uint32_t act;
    if (pv.variant == GGPU_BFS_EDGE) {
        act = GGPU_ACTION_TD_EDGE;
    } else if (pv.variant == GGPU_BFS_BINNED) {
        act = GGPU_ACTION_TD_BINNED;
    } else if (pv.mode == GGPU_POLICY_SCRIPTED) {
        act = pv.actions[st.level % pv.n_actions];
        if (act > GGPU_ACTION_BU) act = GGPU_ACTION_BU;
        st.bottom_up = act == GGPU_ACTION_BU;
    } else if (pv.mode == GGPU_POLICY_LEARNED && pv.mlp) {
        float f[5], y[3];
        if (y_pre) {
            for (int q = 0; q < 3; ++q) y[q] = y_pre[q];
        } else {
            policy_features(st.size, st.total, st.unvisited_edges, st.level, pv.mean_degree, f);
            mlp_forward1(pv.mlp, f, y);
        }
        act = 0;
        for (uint32_t k = 1; k < 3; ++k)
            if (y[k] < y[act]) act = k;
        if (!(y[act] == y[act])) act = st.bottom_up ? GGPU_ACTION_BU : GGPU_ACTION_TD_EDGE;  // NaN guard
        st.bottom_up = act == GGPU_ACTION_BU;
    } else {
        // GAP rule (bfs_cpu.cpp:288-289 / bfs_gpu.mm:461-464)
        if (!st.bottom_up && (double)st.scout > (double)st.to_check / pv.alpha) st.bottom_up = 1;
        act = st.bottom_up ? GGPU_ACTION_BU : GGPU_ACTION_TD_EDGE;
    }
    if (act != GGPU_ACTION_BU) st.to_check -= st.to_check < st.scout ? st.to_check : st.scout;
    st.action = act;
}

// Folds a finished level's counters into the state (bfs_gpu.mm:518-532).
__host__ __device__ inline void fold_level(LoopState& st, const LevelOut& r, const PolicyView& pv, uint64_t n) {
    const uint64_t old = st.size;
    st.size = r.found;
    st.total = r.next_edges;
    st.unvisited_edges -= st.unvisited_edges < r.next_edges ? st.unvisited_edges : r.next_edges;
    for (int k = 0; k < 3; ++k) st.bins[k] = r.bins[k];
    const bool gap = pv.variant == GGPU_BFS_DO || pv.variant == GGPU_BFS_HYBRID;
    const bool heuristic = pv.mode == GGPU_POLICY_HEURISTIC || (pv.mode == GGPU_POLICY_LEARNED && !pv.mlp);
    if (st.action == GGPU_ACTION_BU) {
        st.fmt = FMT_BITMAP;
        st.visited_fresh = 1;
        if (gap && heuristic) {
            if (!(st.size >= old || (double)st.size > (double)n / pv.beta)) {
                st.bottom_up = 0;
                st.scout = 1;
            }
        } else {
            st.scout = st.total;
        }
    } else {
        st.fmt = st.action == GGPU_ACTION_TD_EDGE ? FMT_QUEUE : FMT_BINS;
        st.visited_fresh = 0;
        st.scout = st.total;
    }
    st.cur ^= 1;
    st.level += 1;
    if (st.size == 0) st.done = 1;
}

constexpr uint32_t kHeavyMin = 32;     // bottom-up: degree handled by a wave instead of a lane
constexpr uint32_t kHeavyChunk = 4096; // bottom-up: edges per heavy item

struct HybridMirror {
    std::vector<uint32_t> dist;
    std::vector<uint32_t> q[2], deg[2];
    std::vector<uint64_t> bm[2];
};

struct BfsEngine {
    ggpu_ctx* ctx;
    uint32_t n;
    uint64_t m;
    uint32_t nw;
    uint32_t grid_cap = 4096;
    Csr host;  // host copy of the graph: degrees, hybrid CPU levels, policy features
    DBuf<uint32_t> rp, ci, dist, qa, qb, dega, degb, tile, bins;
    DBuf<uint64_t> bms;
    DBuf<uint4> heavy;       // bottom-up heavy items
    uint32_t n_heavy = 0;
    DBuf<unsigned long long> stats;
    std::vector<hipEvent_t> events;
    double last_gpu_ms = 0;
    uint32_t last_levels = 0;
    HybridMirror hmirror;
    bool mirror_ready = false;
    void* h_mirror = nullptr;
    // device-driven loop resources (bfs_loop.hip), released through dl_free_fn
    void* dl_state = nullptr;
    void (*dl_free_fn)(void*) = nullptr;
    // batched BFS state (msbfs.hip), released through ms_free
    void* ms_state = nullptr;
    void (*ms_free)(void*) = nullptr;

    BfsEngine(ggpu_ctx* c, uint64_t n, uint64_t m, const uint32_t* rp, const uint32_t* ci);
    ~BfsEngine();
    BfsDev dev();
    void ensure_edge_buffers();
    void ensure_bins();
    void ensure_bitmaps();
    hipEvent_t event(size_t i);
    HybridMirror& mirror();

    void host_loop(int variant, uint32_t src, const ggpu_bfs_opts& o, std::vector<ggpu_level>* trace);
    // gpu_vertex / gpu_naive: top-down levels until nothing is discovered
    void simple_loop(int variant, uint32_t src, const ggpu_bfs_opts& o, std::vector<ggpu_level>* trace);
    void device_loop(int variant, uint32_t src, const ggpu_bfs_opts& o, std::vector<ggpu_level>* trace);
    void launch_level(BfsDev& d, LoopState& st, const ggpu_bfs_opts& o, unsigned long long* slot, uint32_t tg);
    void download_state(LoopState& st, HybridMirror& hm);
    void upload_state(LoopState& st, HybridMirror& hm);
    LevelOut cpu_level(LoopState& st, HybridMirror& hm, const ggpu_bfs_opts& o);

    LoopState loop_start(uint32_t src, int variant, const ggpu_bfs_opts& o) const {
        LoopState st{};
        const uint32_t dg = host.deg(src);
        st.size = 1;
        st.total = dg;
        st.scout = dg;
        st.to_check = m;
        st.unvisited_edges = m - dg;
        st.bins[bin_of(dg, o.small_max, o.large_min)] = 1;
        st.fmt = variant == GGPU_BFS_BINNED ? FMT_BINS : FMT_QUEUE;
        return st;
    }
    PolicyView policy_view(int variant, const ggpu_bfs_opts& o, const float* mlp) const {
        PolicyView pv{};
        pv.mode = o.policy;
        pv.variant = (uint32_t)variant;
        pv.alpha = o.alpha;
        pv.beta = o.beta;
        pv.actions = o.actions;
        pv.n_actions = o.n_actions;
        pv.mlp = mlp;
        pv.mean_degree = n ? (double)m / (double)n : 0.0;
        return pv;
    }
};

__host__ __device__ inline ggpu_level trace_record(const LoopState& st, const LevelOut& r, char device, double ms) {
    ggpu_level L{};
    L.level = st.level;
    L.action = st.action;
    L.frontier = st.size;
    L.edges = st.action == GGPU_ACTION_BU ? r.examined : st.total;
    L.frontier_edges = st.total;
    L.unvisited_edges = st.unvisited_edges;
    L.direction = st.action == GGPU_ACTION_BU ? 'B' : 'T';
    L.device = device;
    L.time_ms = ms;
    return L;
}

uint32_t blocks_for(uint64_t items, uint32_t per_block, uint32_t cap);
void check_tg(uint32_t tg);

}  // namespace ggpu
