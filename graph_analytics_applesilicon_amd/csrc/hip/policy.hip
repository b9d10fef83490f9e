// Learned direction policy (no reference counterpart): parameter slots, the
// GPU forward / backward / Adam kernels of the 5 -> 32 -> 32 -> 3 ReLU MLP,
// and the single-source level recorder that produces its training data.
//
// Backward is split for determinism: one thread per sample runs the forward
// pass and stores its activations and deltas; then one thread per parameter
// sums its gradient over a fixed chunk of samples in FP64, and a last kernel
// adds the chunks in order. The same inputs give the same bits on every run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "bfs_engine.hpp"
#include "dev.hpp"

namespace ggpu {

constexpr int kIn = 5, kHid = 32, kOut = 3;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kIn, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kOut * kHid, kParams = kOffB3 + kOut;
static_assert(kParams == GGPU_MLP_PARAMS, "MLP parameter layout");
// per-sample record of the backward pass: h1, h2 (post-ReLU), d1, d2 (dLoss/d pre-activation), dy
constexpr int kActH1 = 0, kActH2 = 32, kActD1 = 64, kActD2 = 96, kActDy = 128, kActWords = 131;
constexpr uint32_t kGradChunk = 1024;  // samples per partial gradient
constexpr uint32_t kMlpThreads = 256;

__global__ void __launch_bounds__(kMlpThreads) k_mlp_forward(const float* __restrict__ p, const float* __restrict__ x,
                                                            uint64_t count, float* __restrict__ y) {
    __shared__ float sp[kParams];
    for (uint32_t i = threadIdx.x; i < kParams; i += blockDim.x) sp[i] = p[i];
    __syncthreads();
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= count) return;
    float xi[kIn], yo[kOut];
    for (int i = 0; i < kIn; ++i) xi[i] = x[s * kIn + i];
    mlp_forward1(sp, xi, yo);
    for (int k = 0; k < kOut; ++k) y[s * kOut + k] = yo[k];
}

// Forward + per-sample deltas of the masked mean squared error.
__global__ void __launch_bounds__(kMlpThreads) k_mlp_samples(const float* __restrict__ p, const float* __restrict__ x,
                                                            const float* __restrict__ t, const uint8_t* __restrict__ mask,
                                                            uint64_t count, double norm, float* __restrict__ act,
                                                            double* __restrict__ loss_part) {
    __shared__ float sp[kParams];
    __shared__ double s_red[kMlpThreads / kWave];
    for (uint32_t i = threadIdx.x; i < kParams; i += blockDim.x) sp[i] = p[i];
    __syncthreads();
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double loss = 0;
    if (s < count) {
        const float* W1 = sp + kOffW1;
        const float* b1 = sp + kOffB1;
        const float* W2 = sp + kOffW2;
        const float* b2 = sp + kOffB2;
        const float* W3 = sp + kOffW3;
        const float* b3 = sp + kOffB3;
        float xi[kIn], a1[kHid], h1[kHid], a2[kHid], h2[kHid], dy[kOut];
        for (int i = 0; i < kIn; ++i) xi[i] = x[s * kIn + i];
        for (int j = 0; j < kHid; ++j) {
            float a = b1[j];
            for (int i = 0; i < kIn; ++i) a += W1[j * kIn + i] * xi[i];
            a1[j] = a;
            h1[j] = a > 0.f ? a : 0.f;
        }
        for (int j = 0; j < kHid; ++j) {
            float a = b2[j];
            for (int i = 0; i < kHid; ++i) a += W2[j * kHid + i] * h1[i];
            a2[j] = a;
            h2[j] = a > 0.f ? a : 0.f;
        }
        for (int k = 0; k < kOut; ++k) {
            float a = b3[k];
            for (int i = 0; i < kHid; ++i) a += W3[k * kHid + i] * h2[i];
            const bool on = mask ? mask[s * kOut + k] != 0 : true;
            const double e = (double)a - (double)t[s * kOut + k];
            dy[k] = on ? (float)(2.0 * e * norm) : 0.f;
            if (on) loss += e * e * norm;
        }
        float d2[kHid];
        for (int i = 0; i < kHid; ++i) {
            float g = 0.f;
            for (int k = 0; k < kOut; ++k) g += dy[k] * W3[k * kHid + i];
            d2[i] = a2[i] > 0.f ? g : 0.f;
        }
        float* r = act + s * kActWords;
        for (int i = 0; i < kHid; ++i) {
            float g = 0.f;
            for (int j = 0; j < kHid; ++j) g += d2[j] * W2[j * kHid + i];
            r[kActD1 + i] = a1[i] > 0.f ? g : 0.f;
            r[kActD2 + i] = d2[i];
            r[kActH1 + i] = h1[i];
            r[kActH2 + i] = h2[i];
        }
        for (int k = 0; k < kOut; ++k) r[kActDy + k] = dy[k];
    }
    const double tot = block_sum(loss, s_red);
    if (threadIdx.x == 0) loss_part[blockIdx.x] = tot;
}

// gpart[chunk][param] = sum over the chunk's samples of that parameter's gradient terms
__global__ void __launch_bounds__(kMlpThreads) k_mlp_grad_chunks(const float* __restrict__ act,
                                                                const float* __restrict__ x, uint64_t count,
                                                                double* __restrict__ gpart) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (uint32_t)kParams) return;
    const uint64_t s0 = (uint64_t)blockIdx.y * kGradChunk;
    const uint64_t s1 = s0 + kGradChunk < count ? s0 + kGradChunk : count;
    int da, db;  // delta word, input word (-1: bias, -2: x)
    if (q < (uint32_t)kOffB1) {
        da = kActD1 + (int)q / kIn;
        db = -2 - (int)(q % kIn);
    } else if (q < (uint32_t)kOffW2) {
        da = kActD1 + (int)(q - kOffB1);
        db = -1;
    } else if (q < (uint32_t)kOffB2) {
        da = kActD2 + (int)(q - kOffW2) / kHid;
        db = kActH1 + (int)(q - kOffW2) % kHid;
    } else if (q < (uint32_t)kOffW3) {
        da = kActD2 + (int)(q - kOffB2);
        db = -1;
    } else if (q < (uint32_t)kOffB3) {
        da = kActDy + (int)(q - kOffW3) / kHid;
        db = kActH2 + (int)(q - kOffW3) % kHid;
    } else {
        da = kActDy + (int)(q - kOffB3);
        db = -1;
    }
    double g = 0;
    for (uint64_t s = s0; s < s1; ++s) {
        const float* r = act + s * kActWords;
        const double d = r[da];
        if (db == -1) g += d;
        else if (db <= -2) g += d * (double)x[s * kIn + (-2 - db)];
        else g += d * (double)r[db];
    }
    gpart[(size_t)blockIdx.y * kParams + q] = g;
}

__global__ void __launch_bounds__(kMlpThreads) k_mlp_grad_final(const double* __restrict__ gpart, uint32_t chunks,
                                                               const double* __restrict__ loss_part,
                                                               uint32_t loss_blocks, float* __restrict__ grad,
                                                               double* __restrict__ loss) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < (uint32_t)kParams) {
        double g = 0;
        for (uint32_t c = 0; c < chunks; ++c) g += gpart[(size_t)c * kParams + q];
        grad[q] = (float)g;
    }
    if (q == 0) {
        double l = 0;
        for (uint32_t b = 0; b < loss_blocks; ++b) l += loss_part[b];
        *loss = l;
    }
}

__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                           float* __restrict__ m2, float lr, float c1, float c2) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (uint32_t)kParams) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gq = g[q];
    const float a = b1 * m1[q] + (1.f - b1) * gq;
    const float b = b2 * m2[q] + (1.f - b2) * gq * gq;
    m1[q] = a;
    m2[q] = b;
    p[q] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// Stochastic gradient descent with momentum (full batch): v = mu v + g, p -= lr v.
__global__ void k_mlp_sgd(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ v, float lr,
                          float mu) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (uint32_t)kParams) return;
    const float a = mu * v[q] + g[q];
    v[q] = a;
    p[q] -= lr * a;
}

// Device copies of one training set and the scratch of a backward pass.
struct MlpData {
    DBuf<float> p, x, t, act, grad, m1, m2;
    DBuf<uint8_t> mask;
    DBuf<double> gpart, loss_part, loss;
    uint64_t count = 0;
    double norm = 0;
    uint32_t chunks = 0, blocks = 0;
    bool has_mask = false;

    void load(hipStream_t s, const float* params, const float* hx, const float* ht, const uint8_t* hm, uint64_t n) {
        count = n;
        uint64_t used = 0;
        for (uint64_t i = 0; i < n * kOut; ++i) used += hm ? (hm[i] != 0) : 1;
        norm = used ? 1.0 / (double)used : 0.0;
        blocks = (uint32_t)((n + kMlpThreads - 1) / kMlpThreads);
        chunks = (uint32_t)((n + kGradChunk - 1) / kGradChunk);
        p.alloc(kParams);
        x.alloc(n * kIn);
        t.alloc(n * kOut);
        act.alloc(n * kActWords);
        grad.alloc(kParams);
        gpart.alloc((size_t)chunks * kParams);
        loss_part.alloc(blocks);
        loss.alloc(1);
        has_mask = hm != nullptr;
        GGPU_HIP_TRY(hipMemcpyAsync(p.p, params, kParams * 4, hipMemcpyHostToDevice, s));
        GGPU_HIP_TRY(hipMemcpyAsync(x.p, hx, n * kIn * 4, hipMemcpyHostToDevice, s));
        GGPU_HIP_TRY(hipMemcpyAsync(t.p, ht, n * kOut * 4, hipMemcpyHostToDevice, s));
        if (hm) {
            mask.alloc(n * kOut);
            GGPU_HIP_TRY(hipMemcpyAsync(mask.p, hm, n * kOut, hipMemcpyHostToDevice, s));
        }
    }
    void backward(hipStream_t s) {
        k_mlp_samples<<<blocks, kMlpThreads, 0, s>>>(p.p, x.p, t.p, has_mask ? mask.p : nullptr, count, norm, act.p,
                                                    loss_part.p);
        const uint32_t pb = (kParams + kMlpThreads - 1) / kMlpThreads;
        k_mlp_grad_chunks<<<dim3(pb, chunks), kMlpThreads, 0, s>>>(act.p, x.p, count, gpart.p);
        k_mlp_grad_final<<<pb, kMlpThreads, 0, s>>>(gpart.p, chunks, loss_part.p, blocks, grad.p, loss.p);
        GGPU_HIP_TRY(hipGetLastError());
    }
};

static void check_params(const float* params, uint32_t count) {
    if (!params) throw ArgError("params is null");
    if (count != GGPU_MLP_PARAMS) throw ArgError("policy parameters must have GGPU_MLP_PARAMS entries");
    for (uint32_t i = 0; i < count; ++i)
        if (!std::isfinite(params[i])) throw ArgError("policy parameters must be finite");
}

static void check_slot(int slot) {
    if (slot != GGPU_POLICY_SINGLE && slot != GGPU_POLICY_BATCHED) throw ArgError("unknown policy slot");
}

}  // namespace ggpu

using namespace ggpu;

ggpu::BfsEngine* ggpu_bfs_engine(ggpu_bfs* b);

extern "C" {

int ggpu_policy_set(ggpu_ctx* ctx, int slot, const float* params, uint32_t count) {
    if (!ctx) return GGPU_EINVAL;
    return guarded(ctx, [&] {
        check_slot(slot);
        check_params(params, count);
        GGPU_HIP_TRY(hipSetDevice(ctx->device));
        ctx->policy_host[slot].assign(params, params + count);
        GGPU_HIP_TRY(hipMemcpyAsync(ctx->policy_dev[slot], ctx->policy_host[slot].data(), count * sizeof(float),
                                    hipMemcpyHostToDevice, ctx->stream));
        GGPU_HIP_TRY(hipStreamSynchronize(ctx->stream));
        ctx->policy_set[slot] = true;
    });
}

int ggpu_policy_clear(ggpu_ctx* ctx, int slot) {
    if (!ctx) return GGPU_EINVAL;
    return guarded(ctx, [&] {
        check_slot(slot);
        ctx->policy_set[slot] = false;
        ctx->policy_host[slot].clear();
    });
}

int ggpu_policy_get(ggpu_ctx* ctx, int slot, float* params, uint32_t count, int* is_set) {
    if (!ctx) return GGPU_EINVAL;
    return guarded(ctx, [&] {
        check_slot(slot);
        if (is_set) *is_set = ctx->policy_set[slot] ? 1 : 0;
        if (params && ctx->policy_set[slot]) {
            if (count < GGPU_MLP_PARAMS) throw ArgError("params buffer too small");
            std::memcpy(params, ctx->policy_host[slot].data(), GGPU_MLP_PARAMS * sizeof(float));
        }
    });
}

void ggpu_policy_features(uint64_t n_f, uint64_t m_f, uint64_t m_u, uint32_t level, double mean_degree,
                          float* features) {
    if (features) policy_features(n_f, m_f, m_u, level, mean_degree, features);
}

int ggpu_mlp_forward(ggpu_ctx* ctx, const float* params, const float* x, uint64_t count, float* y) {
    if (!ctx) return GGPU_EINVAL;
    return guarded(ctx, [&] {
        if (!params || (count && (!x || !y))) throw ArgError("null argument");
        if (!count) return;
        GGPU_HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        DBuf<float> dp, dx, dy;
        dp.alloc(kParams);
        dx.alloc(count * kIn);
        dy.alloc(count * kOut);
        GGPU_HIP_TRY(hipMemcpyAsync(dp.p, params, kParams * 4, hipMemcpyHostToDevice, s));
        GGPU_HIP_TRY(hipMemcpyAsync(dx.p, x, count * kIn * 4, hipMemcpyHostToDevice, s));
        k_mlp_forward<<<(uint32_t)((count + kMlpThreads - 1) / kMlpThreads), kMlpThreads, 0, s>>>(dp.p, dx.p, count,
                                                                                              dy.p);
        GGPU_HIP_TRY(hipGetLastError());
        GGPU_HIP_TRY(hipMemcpyAsync(y, dy.p, count * kOut * 4, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
    });
}

int ggpu_mlp_loss_grad(ggpu_ctx* ctx, const float* params, const float* x, const float* t, const uint8_t* mask,
                       uint64_t count, float* grad, float* loss) {
    if (!ctx) return GGPU_EINVAL;
    return guarded(ctx, [&] {
        if (!params || !x || !t || !grad || !count) throw ArgError("null argument or empty training set");
        GGPU_HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        MlpData d;
        d.load(s, params, x, t, mask, count);
        d.backward(s);
        double l = 0;
        GGPU_HIP_TRY(hipMemcpyAsync(grad, d.grad.p, kParams * 4, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipMemcpyAsync(&l, d.loss.p, 8, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
        if (loss) *loss = (float)l;
    });
}

int ggpu_mlp_train(ggpu_ctx* ctx, float* params_inout, const float* x, const float* t, const uint8_t* mask,
                   uint64_t count, uint32_t epochs, float lr, float* final_loss) {
    if (!ctx) return GGPU_EINVAL;
    return guarded(ctx, [&] {
        if (!params_inout || !x || !t || !count) throw ArgError("null argument or empty training set");
        if (!(lr > 0) || !std::isfinite(lr)) throw ArgError("learning rate must be positive");
        check_params(params_inout, GGPU_MLP_PARAMS);
        GGPU_HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        MlpData d;
        d.load(s, params_inout, x, t, mask, count);
        d.m1.alloc(kParams);
        d.m2.alloc(kParams);
        GGPU_HIP_TRY(hipMemsetAsync(d.m1.p, 0, kParams * 4, s));
        GGPU_HIP_TRY(hipMemsetAsync(d.m2.p, 0, kParams * 4, s));
        const uint32_t pb = (kParams + kMlpThreads - 1) / kMlpThreads;
        double c1 = 1, c2 = 1;
        for (uint32_t e = 0; e < epochs; ++e) {
            d.backward(s);
            c1 *= 0.9;
            c2 *= 0.999;
            k_mlp_adam<<<pb, kMlpThreads, 0, s>>>(d.p.p, d.grad.p, d.m1.p, d.m2.p, lr, (float)(1 - c1),
                                                   (float)(1 - c2));
            GGPU_HIP_TRY(hipGetLastError());
        }
        d.backward(s);  // loss at the final parameters
        double l = 0;
        GGPU_HIP_TRY(hipMemcpyAsync(params_inout, d.p.p, kParams * 4, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipMemcpyAsync(&l, d.loss.p, 8, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
        if (final_loss) *final_loss = (float)l;
    });
}

int ggpu_mlp_train_sgd(ggpu_ctx* ctx, float* params_inout, const float* x, const float* t, const uint8_t* mask,
                       uint64_t count, uint32_t epochs, float lr, float momentum, float* final_loss) {
    if (!ctx) return GGPU_EINVAL;
    return guarded(ctx, [&] {
        if (!params_inout || !x || !t || !count) throw ArgError("null argument or empty training set");
        if (!(lr > 0) || !std::isfinite(lr)) throw ArgError("learning rate must be positive");
        if (!(momentum >= 0 && momentum < 1)) throw ArgError("momentum must be in [0, 1)");
        check_params(params_inout, GGPU_MLP_PARAMS);
        GGPU_HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        MlpData d;
        d.load(s, params_inout, x, t, mask, count);
        d.m1.alloc(kParams);  // the velocity
        GGPU_HIP_TRY(hipMemsetAsync(d.m1.p, 0, kParams * 4, s));
        const uint32_t pb = (kParams + kMlpThreads - 1) / kMlpThreads;
        for (uint32_t e = 0; e < epochs; ++e) {
            d.backward(s);
            k_mlp_sgd<<<pb, kMlpThreads, 0, s>>>(d.p.p, d.grad.p, d.m1.p, lr, momentum);
            GGPU_HIP_TRY(hipGetLastError());
        }
        d.backward(s);  // loss at the final parameters
        double l = 0;
        GGPU_HIP_TRY(hipMemcpyAsync(params_inout, d.p.p, kParams * 4, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipMemcpyAsync(&l, d.loss.p, 8, hipMemcpyDeviceToHost, s));
        GGPU_HIP_TRY(hipStreamSynchronize(s));
        if (final_loss) *final_loss = (float)l;
    });
}

int ggpu_bfs_record_levels(ggpu_bfs* b, uint32_t source, const ggpu_bfs_opts* opt, float* features,
                           float* times_ms, ggpu_level* trace, size_t cap, size_t* len) {
    BfsEngine* e = ggpu_bfs_engine(b);
    if (!e) return GGPU_EINVAL;
    ggpu_ctx* ctx = e->ctx;
    return guarded(ctx, [&] {
        ggpu_bfs_opts o;
        ggpu_bfs_opts_default(&o);
        if (opt) o = *opt;
        o.loop = GGPU_LOOP_HOST;
        o.policy = GGPU_POLICY_SCRIPTED;
        o.time_levels = 1;
        o.n_actions = 1;
        const size_t cap_all = (size_t)e->n + 2;
        std::vector<ggpu_level> runs[3];
        size_t levels = 0;
        for (uint8_t a = 0; a < 3; ++a) {
            o.actions = &a;
            runs[a].resize(cap_all);
            size_t ln = 0;
            const int rc = ggpu_bfs_run(b, GGPU_BFS_DO, source, &o, runs[a].data(), runs[a].size(), &ln);
            if (rc != GGPU_OK) throw ArgError(ctx->err, rc);
            if (a == 0) levels = ln;
            else if (ln != levels) throw ArgError("level counts differ between actions", GGPU_ESTATE);
        }
        const double mean_degree = e->n ? (double)e->m / (double)e->n : 0.0;
        const size_t cnt = std::min(levels, cap);
        for (size_t l = 0; l < cnt; ++l) {
            const ggpu_level& L = runs[0][l];
            if (features)
                policy_features(L.frontier, L.frontier_edges, L.unvisited_edges, L.level, mean_degree, features + l * 5);
            if (times_ms)
                for (int a = 0; a < 3; ++a) times_ms[l * 3 + a] = (float)runs[a][l].time_ms;
            if (trace) trace[l] = L;
        }
        if (len) *len = levels;
    });
}

}  // extern "C"
