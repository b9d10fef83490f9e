// BFS level-loop state and the direction decision, shared by every loop
// that decides per level: the host-driven GPU loop (bfs.hip), the device
// loops (bfs_loop.hip), the batched BFS's features (msbfs.hip), the policy
// trainer and the CPU direction-optimizing BFS (cpu_bfs.cpp). Plain C++ that
// hipcc also compiles for the device (GGPU_HD).
#pragma once

#include <cmath>
#include <cstdint>

#include "ggpu_hip.h"
#include "ggpu_types.h"

#if defined(__HIPCC__)
#define GGPU_HD __host__ __device__
#else
#define GGPU_HD
#endif

namespace ggpu {

// Frontier representation of the current level.
enum Fmt : uint32_t { FMT_QUEUE = 0, FMT_BINS = 1, FMT_BITMAP = 2 };

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

GGPU_HD inline uint32_t bin_of(uint32_t deg, uint32_t small_max, uint32_t large_min) {
    return deg < small_max ? 0u : (deg < large_min ? 1u : 2u);
}

// The state before level 0: the source alone in the frontier, `src_deg` its
// degree, m the graph's arcs.
GGPU_HD inline LoopState loop_start_state(uint32_t src_deg, uint64_t m, uint32_t variant, uint32_t small_max,
                                          uint32_t large_min) {
    LoopState st{};
    st.size = 1;
    st.total = src_deg;
    st.scout = src_deg;
    st.to_check = m;
    st.unvisited_edges = m - src_deg;
    st.bins[bin_of(src_deg, small_max, large_min)] = 1;
    st.fmt = variant == GGPU_BFS_BINNED ? FMT_BINS : FMT_QUEUE;
    return st;
}

GGPU_HD inline void policy_features(uint64_t n_f, uint64_t m_f, uint64_t m_u, uint32_t level,
                                                double mean_degree, float* f) {
    f[0] = (float)(log2(1.0 + (double)n_f) / 32.0);
    f[1] = (float)(log2(1.0 + (double)m_f) / 32.0);
    f[2] = (float)(log2(1.0 + (double)m_u) / 32.0);
    f[3] = (float)level / 64.0f;
    f[4] = (float)(log2(1.0 + mean_degree) / 16.0);
}

// MLP forward 5 -> 32 -> 32 -> 3 (ReLU) for one sample, FP32.
GGPU_HD inline void mlp_forward1(const float* p, const float* x, float* y) {
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

// Picks the action of the level about to run (state after the previous level).
// y_pre: the MLP's outputs for this state when the caller already evaluated
// them (the device loops do, on one wave); otherwise computed here.
GGPU_HD inline void decide_action(LoopState& st, const PolicyView& pv, uint64_t n,
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
GGPU_HD inline void fold_level(LoopState& st, const LevelOut& r, const PolicyView& pv, uint64_t n) {
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

}  // namespace ggpu
