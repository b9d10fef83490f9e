/* Plain-old-data records shared by the host graph library (ggpu_graph.h) and
 * the HIP backend (ggpu_hip.h). C99, no C++ or HIP types, so a cgo / ctypes /
 * JNI binding can mirror them field for field. */
#ifndef GGPU_TYPES_H
#define GGPU_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GGPU_INF 0xFFFFFFFFu   /* dist of an unreached vertex (cpu::kUnreached, src/cpu/bfs_cpu.hpp:14) */

/* Status codes: every entry point returns one (0 = success). */
#define GGPU_OK 0
#define GGPU_EINVAL 1      /* bad argument (size, option, variant name) */
#define GGPU_EDEVICE 2     /* HIP runtime error */
#define GGPU_EIO 3         /* file I/O or format error */
#define GGPU_ENOMEM 4      /* allocation failed */
#define GGPU_ESTATE 5      /* call out of order (e.g. fetch before run) */
#define GGPU_ETIMEOUT 6    /* a bounded device wait gave up (persistent kernel barrier) */

/* CSR flags, bit-compatible with the GGPUCSR1 header (src/graph/csr.hpp:14-18). */
#define GGPU_FLAG_DIRECTED 1u
#define GGPU_FLAG_SORTED 2u
#define GGPU_FLAG_DEDUP 4u

/* One BFS level, as util::BfsLevel (src/util/trace.hpp:7-14) plus the two
 * heuristic inputs and the action taken, which the learned policy trains on. */
typedef struct ggpu_level {
    uint32_t level;           /* distance of the input frontier */
    uint32_t action;          /* 0 edge-balanced top-down, 1 binned top-down, 2 bitmap bottom-up */
    uint64_t frontier;        /* vertices in the input frontier */
    uint64_t edges;           /* adjacency entries examined */
    uint64_t frontier_edges;  /* sum of the frontier's degrees (m_f, exact) */
    uint64_t unvisited_edges; /* sum of the unvisited vertices' degrees (m_u, exact) */
    char direction;           /* 'T' top-down, 'B' bottom-up */
    char device;              /* 'C' CPU, 'G' GPU */
    char pad[6];
    double time_ms;           /* wall time of the level (0 when the level ran inside one launch) */
} ggpu_level;

#ifdef __cplusplus
}
#endif
#endif
