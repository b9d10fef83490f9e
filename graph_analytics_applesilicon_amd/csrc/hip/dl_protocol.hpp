// Control protocol of the persistent single-source BFS loop (bfs_loop.hip),
// shared with the host-side model of that protocol that the tests run.
//
// One decider. Workgroup 0 is the only workgroup that folds level counters
// and decides the next action; it publishes each grid level as a command word
// that the other workgroups poll with coherent loads. No other workgroup
// derives control state of its own, so no two workgroups can take different
// branches: a workgroup runs exactly the grid levels workgroup 0 published.
//
// A grid level g (g counts grid levels, 0, 1, ...):
//   workgroup 0 publishes GRID(level, action) with a write-through store;
//   every workgroup runs its share, writes its counters to part slot g & 1
//     with write-through stores, waits for its own stores and atomics, then
//     arrives (a device-scope atomic on its group counter; the last arriver
//     of a group bumps the top counter);
//   workgroup 0 waits until the top counter reaches (g + 1) * groups, reads
//     every workgroup's counters with coherent loads, folds and decides.
// Levels that workgroup 0 runs alone (small top-down frontiers) are not grid
// levels: the others keep polling; workgroup 0 stores a heartbeat word (kind
// NONE, the current level) every 64 such levels so their bounded wait counts
// only polls without progress. The traversal ends with DONE; a workgroup that
// gives up waiting raises the abort flag and every waiter leaves on it.
// Part slot g & 1 is written only after command g, which workgroup 0 stores
// only after it has read slot (g - 1) & 1 completely, so a slot is never
// written while it is read (a single slot would do; two keep the writes of
// level g + 1 off the lines being read at level g).
// Every datum another workgroup reads during the launch is written with a
// device-scope atomic or a write-through store and read with a coherent load;
// nothing depends on cache writeback or invalidate fences. (The depths have
// no reader during the launch.)
#pragma once

#include "bfs_engine.hpp"

namespace ggpu {

// Levels run by workgroup 0 alone: decided top-down, at most kSbMax frontier
// vertices and kSbEdges frontier edges.
constexpr uint32_t kSbMax = 4096;
constexpr uint32_t kSbEdges = 1u << 14;

__host__ __device__ inline bool sb_ok(const LoopState& st) {
    return st.action != GGPU_ACTION_BU && st.size <= kSbMax && st.total <= kSbEdges;
}

enum DlCmdKind : uint32_t {
    kCmdNone = 0,   // heartbeat: workgroup 0 is running level `level` alone
    kCmdGrid = 1,   // run level `level` with `action` on the whole grid
    kCmdDone = 2,   // the traversal is over
    kCmdAbort = 3,  // workgroup 0 gave up waiting
};

// level + 1 in the low 32 bits (levels < n <= 2^32 - 1, so it fits; 0 = no
// command yet), action in bits 32..39, kind in bits 40..47.
__host__ __device__ inline uint64_t dl_cmd(uint32_t kind, uint32_t level, uint32_t action) {
    return (uint64_t)(level + 1u) | ((uint64_t)(action & 0xFFu) << 32) | ((uint64_t)(kind & 0xFFu) << 40);
}
__host__ __device__ inline uint32_t dl_cmd_kind(uint64_t w) { return (uint32_t)(w >> 40) & 0xFFu; }
__host__ __device__ inline uint32_t dl_cmd_action(uint64_t w) { return (uint32_t)(w >> 32) & 0xFFu; }
__host__ __device__ inline uint32_t dl_cmd_level1(uint64_t w) { return (uint32_t)w; }  // level + 1

// Part slot of grid level g.
__host__ __device__ inline uint32_t dl_part_slot(uint32_t g) { return g & 1u; }

// Barrier-group layout (the grid_sync.hpp words): workgroup b arrives on group
// b % groups; returns the size of group `g` and the number of groups.
__host__ __device__ inline uint32_t dl_group_size(uint32_t grid, uint32_t groups, uint32_t g) {
    return grid / groups + (g < grid % groups ? 1u : 0u);
}
__host__ __device__ inline uint32_t dl_groups(uint32_t grid, uint32_t groups) { return grid < groups ? grid : groups; }

}  // namespace ggpu
