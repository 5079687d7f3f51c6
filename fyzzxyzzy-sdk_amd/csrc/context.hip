// Context lifetime, stream selection and error reporting of the C-ABI.
#include <memory>

#include "../../include/fyzz_hip.h"
#include "context.hpp"
#include "stages.hpp"

using namespace fyzz_hip;

extern "C" {

const char* fyzz_hip_version(void) { return "fyzz_hip gfx950 1"; }

int fyzz_hip_context_create(int device, fyzz_hip_context** out) {
    if (!out) return kInvalidArgument;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return kDeviceError;
    if (device < 0 || device >= count) return kInvalidArgument;
    auto ctx = std::make_unique<fyzz_hip_context>();
    ctx->device = device;
    const int rc = guarded(ctx.get(), [&] { sap::bind_fixed_workspace(device); });
    if (rc) return rc;
    *out = ctx.release();
    return kOk;
}

int fyzz_hip_context_destroy(fyzz_hip_context* ctx) {
    if (!ctx) return kInvalidArgument;
    delete ctx;
    return kOk;
}

const char* fyzz_hip_last_error(const fyzz_hip_context* ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

int fyzz_hip_set_stream(fyzz_hip_context* ctx, void* hip_stream) {
    return guarded(ctx, [&] { ctx->stream = static_cast<hipStream_t>(hip_stream); });
}

int fyzz_hip_set_cooperative(fyzz_hip_context* ctx, int cooperative) {
    return guarded(ctx, [&] { ctx->cooperative = cooperative != 0; });
}

int fyzz_hip_synchronize(fyzz_hip_context* ctx) {
    return guarded(ctx, [&] { FYZZ_HIP_CHECK(hipStreamSynchronize(ctx->stream)); });
}

int fyzz_hip_profile_enable(fyzz_hip_context* ctx, int enable) {
    return guarded(ctx, [&] { ctx->prof.enabled = enable != 0; });
}

int fyzz_hip_profile_read(fyzz_hip_context* ctx, const char* name, double* total_ms,
                          uint64_t* count) {
    return guarded(ctx, [&] {
        FYZZ_REQUIRE(name && total_ms && count, kInvalidArgument, "profile_read: null argument");
        ctx->prof.read(name, total_ms, count);
    });
}

}  // extern "C"
