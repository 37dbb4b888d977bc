// Export of an assembled DC-PSE operator as a sparse matrix with the
// reference's value_nz semantics (src/DCPSE/DCPSE_op/DCPSE_op.hpp:69-83):
// for every neighbour q of p the triplet (p, q, W_pq), and for p itself the
// sum of the (p, p, sign * W_pq) triplets, sign = getSign() (Dcpse.hpp:507-550),
// -1 for even nonzero |beta| and +1 otherwise. This is the matrix the
// reference's implicit solvers assemble row by row
// (src/DCPSE/DCPSE_op/DCPSE_Solver.hpp:978-1036); here it is written in one
// pass as a CSR with ascending columns, the diagonal in its sorted place.
#include <algorithm>

#include "dcpse_common.hpp"
#include "dcpse_internal.hpp"

namespace dcpse {

constexpr int kExportLanes = 8;

// G lanes per row. Row p of the output starts at row_ptr[p] + p; entry j of the
// input row moves to j + (col > p), the diagonal to (number of cols < p).
template <typename T, int G>
__global__ __launch_bounds__(256) void k_export_csr(int64_t n_rows, const int64_t* __restrict__ row_ptr,
                                                    const int32_t* __restrict__ col, const T* __restrict__ W,
                                                    double sign, int64_t* __restrict__ out_row_ptr,
                                                    int32_t* __restrict__ out_col, T* __restrict__ out_val) {
    const int64_t p = (int64_t)blockIdx.x * (256 / G) + threadIdx.x / G;
    const int gl = threadIdx.x % G;
    if (p > n_rows) return;  // whole G-lane groups leave together
    if (p == n_rows) {
        if (gl == 0) out_row_ptr[n_rows] = row_ptr[n_rows] + n_rows;
        return;
    }
    const int64_t rs = row_ptr[p], re = row_ptr[p + 1];
    const int64_t base = rs + p;
    double sum = 0.0;
    int below = 0;
    for (int64_t j = rs + gl; j < re; j += G) {
        const int32_t q = col[j];
        const T w = W[j];
        sum += (double)w;
        below += q < p;
        const int64_t dst = base + (j - rs) + (q > p ? 1 : 0);
        out_col[dst] = q;
        out_val[dst] = w;
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off, kWave);
        below += __shfl_xor(below, off, kWave);
    }
    if (gl == 0) {
        out_row_ptr[p] = base;
        out_col[base + below] = (int32_t)p;
        out_val[base + below] = (T)(sign * sum);
    }
}

template <typename T>
static dcpse_status export_csr(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                               const T* W, double sign, int64_t* out_row_ptr, int32_t* out_col, T* out_val) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !col || !W || !out_row_ptr || !out_col || !out_val || n_rows < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    if (sign != 1.0 && sign != -1.0) return fail(ctx, DCPSE_ERR_INVALID, "sign must be +1 or -1");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int64_t rows_per_block = 256 / kExportLanes;
    const int64_t blocks = (n_rows + 1 + rows_per_block - 1) / rows_per_block;
    k_export_csr<T, kExportLanes><<<(unsigned)blocks, 256, 0, ctx->stream>>>(n_rows, row_ptr, col, W, sign,
                                                                             out_row_ptr, out_col, out_val);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

// W_pq = s_p K_pq (G lanes per row): the pre-scaled FP64 weights of an
// operator from its kernels and the row factors s_p = eps_p^-|beta| the
// reference applies per row (Dcpse.hpp:467-499), one rounding per weight as
// in the assembly's epilogue.
template <int G>
__global__ __launch_bounds__(256) void k_scale_rows(int64_t n_rows, const int64_t* __restrict__ row_ptr,
                                                    const double* __restrict__ K, const double* __restrict__ s,
                                                    double* __restrict__ W) {
    const int64_t p = (int64_t)blockIdx.x * (256 / G) + threadIdx.x / G;
    if (p >= n_rows) return;
    const double sp = s[p];
    const int64_t re = row_ptr[p + 1];
    for (int64_t j = row_ptr[p] + threadIdx.x % G; j < re; j += G) W[j] = K[j] * sp;
}

}  // namespace dcpse

using namespace dcpse;

extern "C" {

dcpse_status dcpse_export_csr_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                                  const double* W, double sign, int64_t* out_row_ptr, int32_t* out_col,
                                  double* out_val) {
    return export_csr<double>(ctx, n_rows, row_ptr, col, W, sign, out_row_ptr, out_col, out_val);
}

dcpse_status dcpse_export_csr_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                                  const float* W, double sign, int64_t* out_row_ptr, int32_t* out_col,
                                  float* out_val) {
    return export_csr<float>(ctx, n_rows, row_ptr, col, W, sign, out_row_ptr, out_col, out_val);
}

dcpse_status dcpse_scale_rows_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const double* K,
                                  const double* row_scale, double* W) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (n_rows < 0 || (n_rows > 0 && (!row_ptr || !row_scale || !K || !W)))
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    if (n_rows == 0) return DCPSE_OK;
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    constexpr int G = 32;  // lanes per row: two rows per wavefront
    const int64_t rows_per_block = 256 / G;
    const int64_t blocks = (n_rows + rows_per_block - 1) / rows_per_block;
    k_scale_rows<G><<<(unsigned)blocks, 256, 0, ctx->stream>>>(n_rows, row_ptr, K, row_scale, W);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

}  // extern "C"
