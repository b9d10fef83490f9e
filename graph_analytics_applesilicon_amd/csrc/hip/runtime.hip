// Context: device selection, one stream, pinned staging memory and the
// error string (MetalContext, src/metal/context.hpp:21-59; the MI355X side has
// no runtime shader compilation: kernels are gfx950 code objects in this
// library).
#include <cstring>

#include "dev.hpp"

namespace ggpu {

int set_error(ggpu_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

void* ctx_pinned(ggpu_ctx* ctx, size_t bytes) {
    if (ctx->pinned_bytes >= bytes) return ctx->pinned;
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    ctx->pinned = nullptr;
    ctx->pinned_bytes = 0;
    GGPU_HIP_TRY(hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault));
    ctx->pinned_bytes = bytes;
    return ctx->pinned;
}

}  // namespace ggpu

using namespace ggpu;

extern "C" {

int ggpu_ctx_create(int device, ggpu_ctx** out) {
    if (!out) return GGPU_EINVAL;
    *out = nullptr;
    auto* c = new ggpu_ctx();
    c->device = device;
    int rc = guarded(c, [&] {
        int count = 0;
        GGPU_HIP_TRY(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw ArgError("device " + std::to_string(device) + " not present");
        GGPU_HIP_TRY(hipSetDevice(device));
        hipDeviceProp_t prop;
        GGPU_HIP_TRY(hipGetDeviceProperties(&prop, device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            throw ArgError(std::string("this backend is built for gfx950 only; device is ") + prop.gcnArchName);
        c->name = prop.name;
        if (c->name.empty()) c->name = std::string("AMD ") + prop.gcnArchName;  // marketing name not reported
        c->cus = prop.multiProcessorCount;
        c->hbm = prop.totalGlobalMem;
        GGPU_HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        noalloc_setup(device);
        for (auto& p : c->policy_dev) GGPU_HIP_TRY(hipMalloc((void**)&p, GGPU_MLP_PARAMS * sizeof(float)));
        ctx_pinned(c, 4096);
    });
    if (rc != GGPU_OK) {
        // keep the message reachable through a context-less path
        static thread_local std::string last;
        last = c->err;
        ggpu_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return GGPU_OK;
}

void ggpu_ctx_destroy(ggpu_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    for (auto p : ctx->policy_dev)
        if (p) (void)hipFree(p);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    delete ctx;
}

const char* ggpu_last_error(const ggpu_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int ggpu_device_info(ggpu_ctx* ctx, char* name, size_t name_len, int* cus, uint64_t* hbm_bytes) {
    if (!ctx) return GGPU_EINVAL;
    if (name && name_len) {
        std::strncpy(name, ctx->name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cus) *cus = ctx->cus;
    if (hbm_bytes) *hbm_bytes = ctx->hbm;
    return GGPU_OK;
}

int ggpu_ctx_sync(ggpu_ctx* ctx) {
    if (!ctx) return GGPU_EINVAL;
    return guarded(ctx, [&] { GGPU_HIP_TRY(hipStreamSynchronize(ctx->stream)); });
}

}  // extern "C"
