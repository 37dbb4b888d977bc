// Host-side internals shared by the translation units of libdcpse_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>

#include "dcpse_hip.h"

// Cell list and per-particle take masks of the last same-set, single-cut-off
// support count; the fill that follows with the same arguments gathers its
// rows from them instead of repeating the distance tests.
struct dcpse_support_cache {
    void* buf = nullptr;  // grow-only device buffer
    size_t cap = 0;
    bool valid = false;
    bool spill = false;   // some cells are left to the position kernel
    int dim = 0;
    const double* pos = nullptr;
    int64_t n = 0;
    double rcut = 0.0;
    double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};
    unsigned periodic = 0;
    int nws = 0;          // mask words per particle
};

struct dcpse_ctx {
    int device = 0;
    hipStream_t stream = nullptr;  // nullptr = the device's default stream
    int num_cus = 256;
    int apply_group = 0;   // lanes per row, 0 = automatic
    int apply_epl = 8;     // nonzeros per lane issued up front
    int apply_vec = 1;     // 1: 4-byte loads, 4: 16-byte loads (FP32 only)
    int apply_block = 256;
    int sell_groups = 2;   // SELL apply: groups of 4 slots per pipeline batch
    int apply_acc = 64;    // FP32 SELL forward applies: row sums in FP64 (64) or FP32 (32)
    int sell_depth = 2;    // SELL apply: pipeline stages (2)
    const void* grp_rows = nullptr;  // cache of the automatic apply shape
    int64_t grp_n = -1;
    int grp_g = 8;
    int grp_epl = 8;
    int grp_vec = 4;
    int asm_wpb = 1;       // waves per block in the assembly kernel
    int asm_ppb = 18;      // particles per batch (one wavefront walks a batch)
    int asm_chunk = 64;    // neighbours staged per LDS round (32 or 64)
    int asm_persistent = 0;  // 1: persistent grid taking batches from a counter
    int asm_bucket = 0;    // > 0: batches bucketed by row length, this bucket width
    void* ws = nullptr;    // grow-only device workspace
    size_t ws_size = 0;
    void* scratch = nullptr;  // this device's static no-alloc scratch (device address)
    dcpse_support_cache sup;
    char err[1024] = {0};
};

namespace dcpse {

inline dcpse_status fail(dcpse_ctx* ctx, dcpse_status code, const char* fmt, ...) {
    if (ctx) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define DCPSE_HIP_CHECK(ctx, expr)                                                        \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::dcpse::fail(ctx, DCPSE_ERR_HIP, "%s: %s (%s:%d)", #expr,            \
                                 hipGetErrorString(e_), __FILE__, __LINE__);              \
    } while (0)

#define DCPSE_LAUNCH_CHECK(ctx) DCPSE_HIP_CHECK(ctx, hipGetLastError())

// Bump allocator over the context workspace, valid for one entry-point call.
struct Workspace {
    char* base = nullptr;
    size_t used = 0, cap = 0;
    template <typename T>
    T* take(size_t n) {
        size_t off = (used + 255) & ~size_t(255);
        used = off + n * sizeof(T);
        return used <= cap ? reinterpret_cast<T*>(base + off) : nullptr;
    }
};

// Ensures ctx->ws holds at least `bytes`; waits for queued work before growing.
dcpse_status reserve_workspace(dcpse_ctx* ctx, size_t bytes, Workspace* w);

// Exclusive scan of n int32 counts into out[0..n] (int64), on ctx->stream.
dcpse_status scan_counts(dcpse_ctx* ctx, const int32_t* in, int64_t n, int64_t* out, Workspace* w);
size_t scan_workspace_bytes(int64_t n);

// Looks up this device's address of the static scratch array of the no-alloc
// entry points (once per device) and stores it in ctx->scratch; the device
// must be current.
dcpse_status noalloc_scratch_init(dcpse_ctx* ctx);

// Device fill helpers (kernels, so they can be captured in a graph).
void fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v);

}  // namespace dcpse
