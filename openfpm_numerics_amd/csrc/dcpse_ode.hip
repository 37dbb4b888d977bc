// Explicit time integration of a linear particle ODE df/dt = alpha * D f on
// the device (D a DC-PSE operator), FP32 or FP64 state.
//
// Replaces, for DC-PSE right-hand sides, the odeint adapter of
// src/OdeIntegrators/OdeIntegrators.hpp:63-135 and the per-stage vector
// kernels of src/OdeIntegrators/vector_algebra_ofp_gpu.hpp:13-362 (one launch
// per axpy of every stage), as exercised by
// src/OdeIntegrators/tests/OdeIntegratores_base_tests.cpp:478-577 (RK4 of a
// Laplacian right-hand side). Here a stage is ONE launch: the operator gather
// k = alpha * D u with the stage's two axpys fused into its epilogue,
//     out_a = y_a + ca * k,   out_b = y_b + cb * k,
// so classic RK4 (Butcher tableau 0, 1/2, 1/2, 1; weights 1/6, 1/3, 1/3, 1/6)
// is four launches per step and the state never leaves HBM:
//     stage 1: u1  = f + dt/2 k1      acc = f   + dt/6 k1      (u = f)
//     stage 2: u2  = f + dt/2 k2      acc = acc + dt/3 k2      (u = u1)
//     stage 3: u1  = f + dt   k3      acc = acc + dt/3 k3      (u = u2)
//     stage 4:                        f   = acc + dt/6 k4      (u = u1)
#include <algorithm>

#include "dcpse_common.hpp"
#include "dcpse_internal.hpp"

namespace dcpse {

template <typename T>
struct StageArgs {
    int64_t n_rows;
    const int64_t* row_ptr;
    const int32_t* col;
    const T* W;
    const T* u;
    T alpha;
    const T* y_a;
    const T* y_b;
    T* out_a;
    T* out_b;
    T ca, cb;
};

constexpr int kStageLanes = 8;  // lanes per CSR row

// G lanes per row stride the row, then reduce; one operator, difference form.
template <typename T, int G>
__global__ __launch_bounds__(256) void k_stage_csr(const StageArgs<T> a) {
    const int64_t row = (int64_t)blockIdx.x * (256 / G) + threadIdx.x / G;
    const int gl = threadIdx.x % G;
    if (row >= a.n_rows) return;  // whole G-lane groups leave together
    const int64_t rs = a.row_ptr[row], re = a.row_ptr[row + 1];
    const double up = a.u[row];
    double acc = 0.0;  // FP64 accumulation for both value types
    for (int64_t j = rs + gl; j < re; j += G) acc = fma((double)a.W[j], (double)a.u[a.col[j]] - up, acc);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kWave);
    if (gl == 0) {
        const double k = (double)a.alpha * acc;
        const double ya = a.out_a ? (double)a.y_a[row] : 0.0;
        const double yb = a.out_b ? (double)a.y_b[row] : 0.0;
        if (a.out_a) a.out_a[row] = (T)fma((double)a.ca, k, ya);
        if (a.out_b) a.out_b[row] = (T)fma((double)a.cb, k, yb);
    }
}

template <typename T>
static dcpse_status stage_csr(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                              const T* W, const T* u, T alpha, const T* y_a, T ca, T* out_a, const T* y_b, T cb,
                              T* out_b) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !col || !W || !u || n_rows < 0) return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    if ((out_a && !y_a) || (out_b && !y_b)) return fail(ctx, DCPSE_ERR_INVALID, "an output without its input");
    if (out_a == u || out_b == u) return fail(ctx, DCPSE_ERR_INVALID, "a stage output may not alias its input u");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0 || (!out_a && !out_b)) return DCPSE_OK;
    StageArgs<T> a{n_rows, row_ptr, col, W, u, alpha, y_a, y_b, out_a, out_b, ca, cb};
    const int64_t rows_per_block = 256 / kStageLanes;
    const int64_t blocks = (n_rows + rows_per_block - 1) / rows_per_block;
    k_stage_csr<T, kStageLanes><<<(unsigned)blocks, 256, 0, ctx->stream>>>(a);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

}  // namespace dcpse

using namespace dcpse;

extern "C" {

dcpse_status dcpse_stage_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                             const float* W, const float* u, float alpha, const float* y_a, float ca, float* out_a,
                             const float* y_b, float cb, float* out_b) {
    return stage_csr<float>(ctx, n_rows, row_ptr, col, W, u, alpha, y_a, ca, out_a, y_b, cb, out_b);
}

dcpse_status dcpse_stage_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                             const double* W, const double* u, double alpha, const double* y_a, double ca,
                             double* out_a, const double* y_b, double cb, double* out_b) {
    return stage_csr<double>(ctx, n_rows, row_ptr, col, W, u, alpha, y_a, ca, out_a, y_b, cb, out_b);
}

#define DCPSE_RK4_STEPS(STAGE, T, ...)                                                               \
    do {                                                                                             \
        T* u1 = work;                                                                                \
        T* u2 = work + n_rows;                                                                       \
        T* acc = work + 2 * n_rows;                                                                  \
        const T h = (T)dt, h2 = (T)(0.5 * dt), h3 = (T)(dt / 3.0), h6 = (T)(dt / 6.0);               \
        for (int64_t s = 0; s < n_steps; ++s) {                                                      \
            dcpse_status st;                                                                         \
            if ((st = STAGE(__VA_ARGS__, f, alpha, f, h2, u1, f, h6, acc))) return st;              \
            if ((st = STAGE(__VA_ARGS__, u1, alpha, f, h2, u2, acc, h3, acc))) return st;           \
            if ((st = STAGE(__VA_ARGS__, u2, alpha, f, h, u1, acc, h3, acc))) return st;            \
            if ((st = STAGE(__VA_ARGS__, u1, alpha, (const T*)nullptr, (T)0, (T*)nullptr, acc, h6, f))) \
                return st;                                                                           \
        }                                                                                            \
    } while (0)

dcpse_status dcpse_rk4_sell_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int64_t* slice_ptr,
                                const int32_t* sell_row, const int32_t* sell_col, const float* sell_W, float alpha,
                                double dt, int64_t n_steps, float* f, float* work) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!f || !work || n_steps < 0 || n_rows < 0) return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    DCPSE_RK4_STEPS(dcpse_stage_sell_f32, float, ctx, n_rows, row_ptr, slice_ptr, sell_row, sell_col, sell_W);
    return DCPSE_OK;
}

dcpse_status dcpse_rk4_bundle16_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                    const int64_t* slice_ptr, const int32_t* sell_row16, const int32_t* bundle,
                                    const int64_t* exc_ptr, const int32_t* exc_col, const float* exc_W,
                                    int64_t n_exc, float alpha, double dt, int64_t n_steps, float* f, float* work) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!f || !work || n_steps < 0 || n_rows < 0) return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    DCPSE_RK4_STEPS(dcpse_stage_bundle16_f32, float, ctx, n_rows, row_ptr, slice_ptr, sell_row16, bundle, exc_ptr,
                    exc_col, exc_W, n_exc);
    return DCPSE_OK;
}

dcpse_status dcpse_rk4_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                           const float* W, float alpha, double dt, int64_t n_steps, float* f, float* work) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!f || !work || n_steps < 0 || n_rows < 0) return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    DCPSE_RK4_STEPS(dcpse_stage_f32, float, ctx, n_rows, row_ptr, col, W);
    return DCPSE_OK;
}

dcpse_status dcpse_rk4_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                           const double* W, double alpha, double dt, int64_t n_steps, double* f, double* work) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!f || !work || n_steps < 0 || n_rows < 0) return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    DCPSE_RK4_STEPS(dcpse_stage_f64, double, ctx, n_rows, row_ptr, col, W);
    return DCPSE_OK;
}

}  // extern "C"
