// Context lifecycle, stream binding and launch configuration of the C ABI.
#include <cstring>

#include "dcpse_internal.hpp"

using namespace dcpse;

extern "C" {

int32_t dcpse_version(void) { return 100; }

dcpse_status dcpse_ctx_create(dcpse_ctx** out, int32_t device) {
    if (!out) return DCPSE_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DCPSE_ERR_HIP;
    if (device < 0 || device >= count) return DCPSE_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return DCPSE_ERR_HIP;
    dcpse_ctx* ctx = new dcpse_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->num_cus = prop.multiProcessorCount;
    // the device's default (null) stream until the caller binds another one
    ctx->stream = nullptr;
    if (noalloc_scratch_init(ctx) != DCPSE_OK) {
        delete ctx;
        return DCPSE_ERR_HIP;
    }
    *out = ctx;
    return DCPSE_OK;
}

dcpse_status dcpse_ctx_destroy(dcpse_ctx* ctx) {
    if (!ctx) return DCPSE_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->ws) (void)hipFree(ctx->ws);
    if (ctx->sup.buf) (void)hipFree(ctx->sup.buf);
    delete ctx;
    return DCPSE_OK;
}

dcpse_status dcpse_ctx_set_stream(dcpse_ctx* ctx, void* hip_stream) {
    if (!ctx) return DCPSE_ERR_INVALID;
    ctx->stream = static_cast<hipStream_t>(hip_stream);
    return DCPSE_OK;
}

void* dcpse_ctx_get_stream(dcpse_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

const char* dcpse_last_error(const dcpse_ctx* ctx) { return ctx ? ctx->err : "null context"; }

dcpse_status dcpse_sync(dcpse_ctx* ctx) {
    if (!ctx) return DCPSE_ERR_INVALID;
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    DCPSE_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return DCPSE_OK;
}

dcpse_status dcpse_set_apply_config(dcpse_ctx* ctx, int32_t group_lanes, int32_t elems_per_lane,
                                    int32_t vector_width, int32_t block_threads) {
    if (!ctx) return DCPSE_ERR_INVALID;
    const int g = group_lanes, e = elems_per_lane;
    bool ok = g == 0;
    if (vector_width == 1)
        ok = ok || (g == 4 && e == 16) || (g == 8 && (e == 8 || e == 16)) || (g == 16 && (e == 4 || e == 8)) ||
             (g == 32 && e == 2) || (g == 64 && e == 1);
    else if (vector_width == 4)
        ok = ok || (g == 4 && e == 8) || (g == 8 && (e == 4 || e == 8 || e == 12)) ||
             (g == 16 && (e == 4 || e == 8)) || (g == 32 && e == 4);
    else
        ok = false;
    if (!ok)
        return fail(ctx, DCPSE_ERR_INVALID,
                    "unsupported apply shape %d lanes x %d elements, vector width %d (compiled: width 1: 4x16, "
                    "8x8, 8x16, 16x4, 16x8, 32x2, 64x1; width 4: 4x8, 8x4, 8x8, 8x12, 16x4, 16x8, 32x4)",
                    g, e, vector_width);
    if (block_threads < 64 || block_threads > 512 || block_threads % 64)
        return fail(ctx, DCPSE_ERR_INVALID, "block_threads must be a multiple of 64 in [64, 512]");
    ctx->apply_group = g;
    ctx->apply_epl = e;
    ctx->apply_vec = vector_width;
    ctx->apply_block = block_threads;
    return DCPSE_OK;
}

dcpse_status dcpse_set_assembly_config(dcpse_ctx* ctx, int32_t waves_per_block, int32_t particles_per_wave) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (waves_per_block != 1)
        return fail(ctx, DCPSE_ERR_UNSUPPORTED, "the assembly kernel runs one wavefront per block");
    if (particles_per_wave < 0 || particles_per_wave > 4096)
        return fail(ctx, DCPSE_ERR_INVALID, "particles_per_wave must be in [0, 4096]");
    particles_per_wave = particles_per_wave > 64 ? 64 : particles_per_wave;
    ctx->asm_wpb = waves_per_block;
    ctx->asm_ppb = particles_per_wave > 0 ? particles_per_wave : 18;
    return DCPSE_OK;
}

dcpse_status dcpse_set_assembly_launch(dcpse_ctx* ctx, int32_t persistent, int32_t particles_per_batch,
                                       int32_t chunk, int32_t bucket_width) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (particles_per_batch < 0 || particles_per_batch > 64)
        return fail(ctx, DCPSE_ERR_INVALID, "particles_per_batch must be in [0, 64]");
    if (chunk != 32 && chunk != 64) return fail(ctx, DCPSE_ERR_INVALID, "chunk must be 32 or 64");
    if (bucket_width < 0 || bucket_width > DCPSE_MAX_ROW)
        return fail(ctx, DCPSE_ERR_INVALID, "bucket_width must be in [0, DCPSE_MAX_ROW]");
    ctx->asm_persistent = persistent ? 1 : 0;
    ctx->asm_ppb = particles_per_batch > 0 ? particles_per_batch : 18;
    ctx->asm_chunk = chunk;
    ctx->asm_bucket = bucket_width;
    return DCPSE_OK;
}

}  // extern "C"
