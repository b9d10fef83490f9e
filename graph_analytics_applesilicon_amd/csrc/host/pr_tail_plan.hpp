// Work plan of the HOT_TAIL PageRank layout (csrc/hip/pr_tail.hpp). In the
// degree-ordered numbering a Jacobi sweep gathers three kinds of sources:
// [0, 30720) from the LDS copy of the HOT sweep, a band of contributions
// above it, and the tail, whose gathers miss L2. The rows of in-degree >=
// min_indeg give most band and tail gathers; their in-lists are sorted, so the
// arcs of such a row from sources >= tail_begin are a suffix of its list. The
// HOT sweep sums the head [rp[v], split) of these "split rows" into partials;
// the suffixes are cut by source tile into pieces, each summed into one
// compensated partial per (tile, row) segment: band tiles (band_tile sources,
// [tail_begin, band_end)) by a kernel that copies the tile to LDS, the tail
// tiles beyond (tile sources, tile-major, so the tile being read stays in L2)
// by a kernel that gathers them; a merge adds a row's head partials, then its
// tile partials in tile order, then its overflow partials (pieces past the
// first of a segment longer than one item), and stores it. Without a band
// (band_tile 0) every tile is a tail tile of `tile` sources from tail_begin.
// No reference counterpart: the reference pulls every row from the whole
// contribution vector (src/shaders/pagerank.metal:306-381).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "graph.hpp"

namespace ggpu {

struct TailParams {
    uint32_t min_indeg = 64;         // rows of in-degree >= this have their tail split off
    uint32_t tail_begin = 1u << 19;  // first tail source id
    uint32_t tile = 1u << 19;        // tail sources per tile
    uint32_t nnz_cap = 512;          // arcs per piece and per item (a multiple of 64)
    uint32_t pieces_cap = 512;       // pieces per item
    uint32_t band_end = 0;           // band: sources [tail_begin, band_end) ...
    uint32_t band_tile = 0;          // ... in tiles of band_tile sources (0: no band)
};

// Band tiles of a plan over n vertices: tiles 0 .. band_tiles - 1.
inline uint32_t tail_band_tiles(const TailParams& p, uint64_t n) {
    if (!p.band_tile || n <= p.tail_begin) return 0;
    const uint64_t e = n < p.band_end ? n : p.band_end;
    return (uint32_t)((e - p.tail_begin + p.band_tile - 1) / p.band_tile);
}

// All tiles of a plan over n vertices: the band tiles, then the tail tiles.
inline uint64_t tail_tiles(const TailParams& p, uint64_t n) {
    if (n <= p.tail_begin) return 0;
    const uint64_t start = p.band_tile ? p.band_end : p.tail_begin;
    return tail_band_tiles(p, n) + (n > start ? (n - start + p.tile - 1) / p.tile : 0);
}

// The tile of source u >= tail_begin.
inline uint32_t tail_tile_of(const TailParams& p, uint32_t u) {
    if (p.band_tile && u < p.band_end) return (u - p.tail_begin) / p.band_tile;
    const uint32_t b0 = p.band_tile ? (p.band_end - p.tail_begin + p.band_tile - 1) / p.band_tile : 0u;
    const uint32_t start = p.band_tile ? p.band_end : p.tail_begin;
    return b0 + (u - start) / p.tile;
}

struct TailPlan {
    uint32_t R = 0;               // split rows
    uint32_t T = 0;               // tiles (band tiles first)
    std::vector<uint32_t> srow;   // split row index -> vertex (ascending)
    std::vector<uint32_t> split;  // split row index -> offset into ci of its first tail arc
    std::vector<uint32_t> tci;    // tail arcs of the split rows: tile-major, rows ascending within a tile
    std::vector<uint32_t> px;     // piece p: arcs tci[px[p], px[p + 1])
    std::vector<uint32_t> pdst;   // piece p: partial index, t * R + j (first piece of row j's tile-t segment)
                                  // or T * R + o (overflow slot o)
    std::vector<uint32_t> ovptr;  // split row j: overflow slots [ovptr[j], ovptr[j + 1]), in (tile, piece) order
    std::vector<std::array<uint32_t, 2>> items;  // {first piece, end piece}; one tile per item, tile-major
    uint64_t head_arcs = 0;       // arcs of the split rows' heads
};

TailPlan plan_tail(const Csr& in, const TailParams& p);

// Executes a plan on the CPU in FP64: sums[v] = sum of contrib[u] over v's
// in-arcs, split rows as head + tile partials + overflow partials in the
// merge's order (test hook; checks the plan's invariants and throws
// std::runtime_error on a violation).
void emulate_tail(const Csr& in, const TailParams& p, const TailPlan& plan, const float* contrib, double* sums);

}  // namespace ggpu
