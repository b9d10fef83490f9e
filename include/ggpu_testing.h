/* Test hooks of the host library (no reference counterpart; not part of the
 * drop-in boundary). Callers: tests/ only. */
#ifndef GGPU_TESTING_H
#define GGPU_TESTING_H

#include "ggpu_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Builds the TILED PageRank work plan for an in-CSR and executes it on the
 * CPU in FP64: sums[v] = sum of contrib[u] over v's in-arcs, taken item by
 * item as the GPU takes them. params (9 words): min in-degree of a tiled row,
 * sources per tile, arcs per item, segments per run, chunk slots per task,
 * min arcs of a dense tile, arcs per task, items per task, workgroups.
 * stats (8 words): tiled rows, dense tiles, dense arcs, cold arcs, untiled
 * arcs, items, tasks, long segments. Checks the plan's invariants: a
 * violation returns GGPU_EINVAL with the message in ggpu_host_last_error(). */
int ggpu_test_pr_tiled_emulate(uint64_t n, uint64_t m, const uint32_t* in_row_ptr, const uint32_t* in_col_idx,
                               const uint32_t* params, const float* contrib, double* sums, uint64_t* stats);

/* Builds the HOT_TAIL PageRank plan (the suffixes of the split rows from the
 * first tail source on, in tile-major pieces and items) and executes it on the
 * CPU in FP64: sums[v] = sum of contrib[u] over v's in-arcs, split rows as head
 * + tile partials + overflow partials in the merge's order. params (7 words):
 * min in-degree of a split row, first tail source, sources per tail tile, arcs
 * per piece and item, pieces per item, end of the band and sources per band
 * tile (0: no band). stats (8 words): split rows, tiles, tail arcs (band
 * included), head arcs, pieces, items, overflow partials, band tiles. Checks
 * the plan's invariants: a violation returns GGPU_EINVAL with the message in
 * ggpu_host_last_error(). */
int ggpu_test_pr_tail_emulate(uint64_t n, uint64_t m, const uint32_t* in_row_ptr, const uint32_t* in_col_idx,
                              const uint32_t* params, const float* contrib, double* sums, uint64_t* stats);

#ifdef __cplusplus
}
#endif
#endif
