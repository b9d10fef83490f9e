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
};

struct PprState {
    DBuf<float> r, c;
    DBuf<double> part, dang, delta;
    DBuf<uint32_t> seeds;
    std::vector<hipEvent_t> ev;
    uint32_t B = 0;
    int slot = 0;
    bool have = false;
    ~PprState() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

struct PrEngine;
PprView ppr_view(PrEngine& e);
PprState& ppr_state(PrEngine& e);
ggpu_pr_result ppr_run(PrEngine& e, const uint32_t* seeds, uint32_t B, const ggpu_pr_opts& o);
void ppr_fetch(PrEngine& e, float* out);

}  // namespace ggpu
