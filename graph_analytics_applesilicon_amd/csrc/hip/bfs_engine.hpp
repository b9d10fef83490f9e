// BFS engine state shared by the host-driven loop (bfs.hip), the device-driven
// loops (bfs_loop.hip), the batched BFS (msbfs.hip) and the policy recorder.
#pragma once

#include <cmath>
#include <vector>

#include "bfs_kernels.hpp"
#include "graph.hpp"
#include "level_policy.hpp"

namespace ggpu {

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
        return loop_start_state(host.deg(src), m, (uint32_t)variant, o.small_max, o.large_min);
    }
    PolicyView policy_view(int variant, const ggpu_bfs_opts& o, const float* mlp) const;
};

// The decision inputs of a run on a graph of n vertices and m arcs.
inline PolicyView make_policy_view(int variant, const ggpu_bfs_opts& o, const float* mlp, uint64_t n, uint64_t m) {
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

inline PolicyView BfsEngine::policy_view(int variant, const ggpu_bfs_opts& o, const float* mlp) const {
    return make_policy_view(variant, o, mlp, n, m);
}

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
