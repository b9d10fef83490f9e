// Device-side building blocks for CDNA4 (gfx950, wave64) and the host-side
// context shared by every backend object.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "ggpu_hip.h"

namespace ggpu {

constexpr int kWave = 64;

// ------------------------------------------------------------ wave helpers

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Number of lanes below this one whose bit is set in `mask`.
__device__ __forceinline__ uint32_t rank_in(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// Inclusive prefix sum over the 64 lanes.
template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        T o = __shfl_up(v, off, kWave);
        if (lane >= (uint32_t)off) v += o;
    }
    return v;
}

// Reserve one slot per lane with `want` set; returns the lane's slot (valid
// only where want). One atomic per wave. Must be called by every lane of the
// wave (convergent).
__device__ __forceinline__ unsigned long long wave_append(bool want, unsigned long long* counter) {
    const uint64_t mask = __ballot(want);
    unsigned long long base = 0;
    if (mask) {
        const uint32_t leader = (uint32_t)__ffsll((long long)mask) - 1;
        if (lane_id() == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
        base = __shfl(base, (int)leader, kWave);
    }
    return base + rank_in(mask);
}

// Wave-aggregated counter add: lane 0 of the wave adds the wave's sum.
__device__ __forceinline__ void wave_add(unsigned long long* counter, unsigned long long v) {
    v = wave_sum(v);
    if (lane_id() == 0 && v) atomicAdd(counter, v);
}

// Block-wide sum of a per-thread value (blockDim.x a multiple of 64, <= 1024).
// `scratch` must hold blockDim.x/64 entries; returns the total in every thread.
template <class T>
__device__ __forceinline__ T block_sum(T v, T* scratch) {
    v = wave_sum(v);
    const uint32_t w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    __syncthreads();
    if (lane_id() == 0) scratch[w] = v;
    __syncthreads();
    T t = 0;
    for (uint32_t i = 0; i < nw; ++i) t += scratch[i];
    return t;
}

// ------------------------------------------------------------ host context

#define GGPU_HIP_TRY(expr)                                                                        \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw ::ggpu::DeviceError(std::string(#expr) + " (" + __FILE__ + ":" +                \
                                      std::to_string(__LINE__) + "): " + hipGetErrorString(e_), e_); \
    } while (0)

struct DeviceError {
    std::string what;
    hipError_t code;
    DeviceError(std::string w, hipError_t c) : what(std::move(w)), code(c) {}
};

struct ArgError {
    std::string what;
    int code;
    explicit ArgError(std::string w, int c = GGPU_EINVAL) : what(std::move(w)), code(c) {}
};

// Device buffer owned by a backend object.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t count = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void alloc(size_t c) {
        if (c <= count && p) return;
        release();
        const size_t bytes = (c ? c : 1) * sizeof(T);
        hipError_t e = hipMalloc((void**)&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            throw DeviceError("hipMalloc(" + std::to_string(bytes) + " bytes): " + hipGetErrorString(e), e);
        }
        count = c;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        count = 0;
    }
};

}  // namespace ggpu

struct ggpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 0;
    std::string name;
    std::string err;
    uint64_t hbm = 0;
    // learned policy parameters per slot (GGPU_POLICY_SINGLE / _BATCHED),
    // host and device copies of GGPU_MLP_PARAMS floats
    bool policy_set[2] = {false, false};
    std::vector<float> policy_host[2];
    float* policy_dev[2] = {nullptr, nullptr};
    // pinned host staging for per-level control readback
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
};

namespace ggpu {
// Pinned host buffer of at least `bytes` owned by the context.
void* ctx_pinned(ggpu_ctx* ctx, size_t bytes);
// With `device` current: finds the device address of the working array of
// ggpu_bfs_run_noalloc on that GPU (once per process and GPU).
void noalloc_setup(int device);
// Runs f, translating exceptions into a status and ctx->err.
template <class F>
int guarded(ggpu_ctx* ctx, F&& f);
int set_error(ggpu_ctx* ctx, int code, const std::string& msg);
}  // namespace ggpu

#include <exception>
#include <new>

namespace ggpu {
template <class F>
int guarded(ggpu_ctx* ctx, F&& f) {
    try {
        f();
        return GGPU_OK;
    } catch (const DeviceError& e) {
        return set_error(ctx, GGPU_EDEVICE, e.what);
    } catch (const ArgError& e) {
        return set_error(ctx, e.code, e.what);
    } catch (const std::bad_alloc&) {
        return set_error(ctx, GGPU_ENOMEM, "out of host memory");
    } catch (const std::exception& e) {
        return set_error(ctx, GGPU_EINVAL, e.what());
    }
}
}  // namespace ggpu
