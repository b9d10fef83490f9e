/* Test hooks of the HIP backend (no reference counterpart; not part of the
 * drop-in boundary): the level-loop primitives run on their own, as the
 * reference's tests run its Primitives (tests/test_main.cpp:355-460).
 * Callers: tests/ only. */
#ifndef GGPU_HIP_TESTING_H
#define GGPU_HIP_TESTING_H

#include "ggpu_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The in-place exclusive scan the edge-balanced top-down runs over frontier
 * degrees (two passes over tiles of 4096 values: tile sums, then each tile's
 * carry-in and its own scan), on n values copied in from `in` and out to
 * `out`; uint32 arithmetic wraps. max_blocks caps the grid (0: the engine's
 * cap; a small cap makes blocks loop over several tiles). Replaces the
 * reference test's Primitives::scanExclusive calls (tests/test_main.cpp:382-420). */
int ggpu_test_scan(ggpu_ctx* ctx, uint32_t n, const uint32_t* in, uint32_t* out, uint32_t max_blocks);

/* The host loop's frontier conversions on the graph of b, with workgroups of
 * tg threads: dist (n depths) is uploaded, the frontier bitmap of depth `level`
 * is built (bitmap: ceil(n/64) words; the device words start as all ones, so
 * every word must be written), then turned into a queue of vertex ids with
 * their degrees (queue, degrees: room for n entries; *count = its length, in
 * no particular order). Replaces queueToBitmap / bitmapToQueue in the
 * reference test (tests/test_main.cpp:422-460). */
int ggpu_test_frontier_convert(ggpu_bfs* b, const uint32_t* dist, uint32_t level, uint32_t tg, uint64_t* bitmap,
                               uint32_t* queue, uint32_t* degrees, uint32_t* count);

/* Batched BFS, routed top-down: a top-down level of a batch of at most 128
 * sources whose frontier has at most 4096 vertices and at least min_edges arcs
 * is routed through per-range bins in LDS (0: the default, 65536 arcs);
 * *routed_last (may be null) = routed levels of b's last batched run. */
int ggpu_test_msbfs_route(ggpu_bfs* b, uint64_t min_edges, uint32_t* routed_last);

/* Sets the HOT_TAIL geometry of p's later layouts: the first source past the
 * split rows' heads, the sources per tail tile, the end of the band and the
 * sources per band tile (at most 16384; 0: no band). tail_begin and tile both
 * 0: the defaults (30720, 2^19, 30720 + 30 * 16384, 16384), whatever the band
 * arguments. Small values make small graphs run bands and tails of many tiles.
 * GGPU_EINVAL for tile 0 with tail_begin > 0, band_tile > 16384, or a band that
 * does not end after tail_begin. A layout built with other values is rebuilt on
 * its next run. */
int ggpu_test_pr_tail_geometry(ggpu_pr* p, uint32_t tail_begin, uint32_t tile, uint32_t band_end,
                               uint32_t band_tile);

/* The HOT_TAIL layout last built for vertex order `order` (GGPU_PR_ORDER_*):
 * stats (8 words) = split rows, tiles, tail arcs (band included), items,
 * split rows merged by a wave, long-row chunks of the HOT items, band tiles,
 * band tasks; all zero if none. */
int ggpu_test_pr_tail_info(const ggpu_pr* p, uint32_t order, uint64_t* stats);

#ifdef __cplusplus
}
#endif
#endif
