// Device-wide barrier for persistent kernels on gfx950 (every workgroup resident).
//
// Each XCD has its own L2, and a plain store stays in the writer's L2 until
// it is written back, so a reader on another XCD can miss it. A fenced
// barrier (agent-scope release: L2 writeback by every block) costs about
// 24 us on 256 blocks. This barrier instead relies on a discipline:
//   * between two barriers, data that another block reads after the barrier
//     is written only with device-scope atomics or agent-scope (write-through)
//     stores -- both reach the device coherence point;
//   * on leaving, a block either
//       - kAcquireL2: invalidates its caches at agent scope (L1 and the
//         XCD's L2), after which plain loads see that data; or
//       - kAcquireL1: invalidates only its L1 and reads data other blocks
//         wrote with coherent (agent-scope) loads; the L2 then keeps the
//         read-only graph arrays across barriers. Plain loads of such data
//         can return stale L2 lines on MI355X (a kernel reading per-block
//         counters that way diverged between blocks).
// The agent-scope invalidate costs every block's L2 contents each barrier
// (tens of us per level when levels are small).
// Blocks arrive on kBarGroups counters (blockIdx % kBarGroups), the last of a
// group on a top counter, the last of those bumps a generation word: about
// 5.4 us on 256 blocks of 512 threads instead of 17 us for one counter.
//
// The wait is bounded: after `spin_limit` polls the block raises `abort` and
// leaves; every block returns false from its next barrier, so a launch whose
// grid is not co-resident (or any bug that loses an arrival) still ends.
// The fences above serve the batched BFS (kAcquireL2: plain loads after the
// barrier); the persistent single-source loop uses only grid_arrive and
// coherent accesses (dl_protocol.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ggpu {

constexpr uint32_t kBarGroups = 8;
constexpr uint32_t kBarStride = 16;                        // words between counters (separate lines)
constexpr uint32_t kBarTop = kBarGroups * kBarStride;      // top counter
constexpr uint32_t kBarGen = kBarTop + kBarStride;         // generation
constexpr uint32_t kBarWords = kBarGen + kBarStride;       // words to allocate and zero
constexpr uint64_t kBarSpinLimit = 1ull << 23;

template <class T, class V>
__device__ __forceinline__ void store_through(T* p, V v) {
    __hip_atomic_store(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ __forceinline__ T load_coherent(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

enum AcquireKind { kAcquireL2 = 0, kAcquireL1 = 1 };

// Called by every thread of every block. Returns false once any block has
// given up waiting; the caller then returns from the kernel.
template <AcquireKind kAcquire = kAcquireL2>
__device__ __forceinline__ bool grid_barrier(uint32_t* bar, uint32_t* abort, uint64_t spin_limit = kBarSpinLimit) {
    __builtin_amdgcn_s_waitcnt(0);  // this thread's stores and atomics have completed
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t gen = load_coherent(&bar[kBarGen]);
        const uint32_t G = gridDim.x;
        const uint32_t g = blockIdx.x % kBarGroups;
        const uint32_t gsize = G / kBarGroups + (g < G % kBarGroups ? 1u : 0u);
        const uint32_t ngroups = G < kBarGroups ? G : kBarGroups;
        // The counters only grow (zeroed per launch; 2^32 arrivals are out of
        // reach): the last arriver of a round is the one that completes a
        // multiple of the group size. Resetting them instead would race with
        // a fast block's arrival at the next barrier.
        bool released = false;
        if ((atomicAdd(&bar[g * kBarStride], 1u) + 1) % gsize == 0) {
            if ((atomicAdd(&bar[kBarTop], 1u) + 1) % ngroups == 0) {
                __hip_atomic_fetch_add(&bar[kBarGen], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                released = true;
            }
        }
        if (!released) {
            uint64_t spins = 0;
            while (load_coherent(&bar[kBarGen]) == gen) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > spin_limit) {
                    atomicExch(abort, 1u);
                    break;
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (kAcquire == kAcquireL2) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        else asm volatile("buffer_inv sc0" ::: "memory");  // the CU's L1 only
    }
    __syncthreads();
    return load_coherent(abort) == 0u;
}

// The arrival half of grid_barrier, for loops in which one workgroup alone
// observes completion (it polls bar[kBarTop] for rounds * groups arrivals)
// and releases the others by a command of its own. Called by every thread.
__device__ __forceinline__ void grid_arrive(uint32_t* bar) {
    __builtin_amdgcn_s_waitcnt(0);  // this thread's stores and atomics have completed
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t G = gridDim.x;
        const uint32_t g = blockIdx.x % kBarGroups;
        const uint32_t gsize = G / kBarGroups + (g < G % kBarGroups ? 1u : 0u);
        if ((atomicAdd(&bar[g * kBarStride], 1u) + 1) % gsize == 0) atomicAdd(&bar[kBarTop], 1u);
    }
}

}  // namespace ggpu
