// State shared between the PageRank engine (pagerank.hip) and the
// personalized-PageRank block (ppr.hip).
#pragma once

#include <vector>

#include "dev.hpp"

namespace ggpu {

struct PprView {
    ggpu_ctx* ctx;
    uint32_t n;
    const uint32_t* rp;
    const uint32_t* ci;
    const uint32_t* od;
    const uint32_t* od_host;
    const uint32_t* rp_host;
    const uint32_t* id_of;  // caller id -> id in this numbering; null: identity
    int space;              // numbering the arrays are in
};

struct PprState {
    DBuf<float> r, c;
    DBuf<double> part, dang, delta;
    DBuf<uint32_t> seeds;
    // in-lists of at least heavy_min entries cut into chunks of heavy_min:
    // chunks[i] = {first entry, end}; heavy[h] = {vertex, first chunk} plus a
    // sentinel {0, n_chunks}; cpart = per-chunk, per-column (hi, lo) sums
    DBuf<uint2> chunks, heavy;
    DBuf<float2> cpart;
    uint32_t heavy_min = 0, n_chunks = 0, n_heavy = 0;
    int heavy_space = -1;
    bool heavy_built = false;
    // device copy of the degree numbering's id map, for the fetch
    DBuf<uint32_t> id_of;
    bool id_of_ready = false;
    uint32_t order = 0;  // GGPU_PR_ORDER_* numbering of the block held
    int occ_tg = 0, occ_sweep = 0, occ_chunks = 0;  // resident workgroups per CU at occ_tg
    std::vector<hipEvent_t> ev;
    uint32_t B = 0;
    int slot = 0;
    bool have = false;
    ~PprState() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

struct PrEngine;
PprView ppr_view(PrEngine& e, uint32_t order);
PprState& ppr_state(PrEngine& e);
ggpu_pr_result ppr_run(PrEngine& e, const uint32_t* seeds, uint32_t B, const ggpu_pr_opts& o);
void ppr_autotune(PrEngine& e, const uint32_t* seeds, uint32_t B, int iters, ggpu_pr_opts& o);
void ppr_fetch(PrEngine& e, float* out);

}  // namespace ggpu
