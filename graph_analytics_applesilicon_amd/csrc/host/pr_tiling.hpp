// Work plan of the TILED PageRank layout (csrc/hip/pr_tiled.hpp): which rows
// are tiled, which source tiles are dense, the tile-major arc arrays, the
// items, the tasks and their assignment to workgroups. Built on the host once
// per vertex numbering; the HIP backend uploads it as is. Item and task
// records are plain u32 words with the device structs' layout.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "graph.hpp"

namespace ggpu {

struct TiledParams {
    uint32_t min_indeg = 128;         // rows of in-degree >= this are tiled
    uint32_t tile = 32768;            // sources per tile
    uint32_t nnz_cap = 1024;          // arcs per item; longer segments are chunked
    uint32_t rows_cap = 64;           // segments (or untiled rows) per run
    uint32_t slots = 512;             // chunk partials per task
    uint32_t dense_min_arcs = 0;      // a tile with fewer arcs into the tiled rows joins the cold pseudo-tile
                                      // (0: max(tile / 16, tiled rows / 8))
    uint32_t task_arcs = 131072;      // arcs per task (a task never splits a long segment)
    uint32_t task_items = 2048;       // items per task at most
    uint32_t groups = 256;            // workgroups the tasks are assigned to
};

// Item kinds (the first word's low two bits; the partial array index above them).
enum : uint32_t { kTiRun = 0, kTiChunk = 1, kTiUntiled = 2 };
constexpr uint32_t kTiNoTile = 0xFFFFFFFFu;

struct TiledPlan {
    uint32_t R = 0;  // tiled rows
    uint32_t K = 0;  // dense tiles (partial array K is the cold pseudo-tile's)
    std::vector<uint32_t> trow;        // tiled row index -> vertex (ascending)
    std::vector<uint32_t> dense_tile;  // dense index -> source tile
    std::vector<uint16_t> tci;         // dense-tile arcs: source - tile * params.tile
    std::vector<uint32_t> cci;         // cold arcs: source ids
    std::vector<uint32_t> ent_row;     // short segments: tiled row index
    std::vector<uint16_t> ent_len;     // short segments: arcs
    std::vector<std::array<uint32_t, 4>> items;  // see TileItemKind
    std::vector<std::array<uint32_t, 8>> tasks;  // item range, LDS tile, k, long-segment range, pad
    std::vector<std::array<uint32_t, 4>> segs;   // long segments: {tiled row, first slot, chunks, dest}
    // pieces of segments longer than one task's slots: dest = o + 1 writes overflow
    // partial o (of tiled row ovf_row[o]); the merge adds row j's overflow
    // partials oidx[optr[j] .. optr[j + 1]) after the tile partials
    std::vector<uint32_t> ovf_row, optr, oidx;
    std::vector<std::array<uint32_t, 2>> wg;     // per workgroup: task range
    uint64_t dense_arcs = 0, cold_arcs = 0, untiled_arcs = 0;
};

TiledPlan plan_tiled(const Csr& in, const TiledParams& p);

// Executes a plan on the CPU in FP64: sums[v] = sum of contrib[u] over v's
// in-arcs, with every item, chunk and merge step the GPU takes (test hook;
// checks the plan's invariants and throws std::runtime_error on a violation).
void emulate_tiled(const Csr& in, const TiledParams& p, const TiledPlan& plan, const float* contrib, double* sums);

}  // namespace ggpu
