/*
 * dcpse_hip.h - C ABI of the MI355X (gfx950) DC-PSE operator library.
 *
 * One shared library (libdcpse_hip.so) implements the DC-PSE hot path of
 * openfpm_numerics: neighbour-support construction, FP64 kernel assembly and
 * CSR operator application (forward and transposed). Host code in any
 * language binds these entry points; no C++ or torch types cross the ABI.
 *
 * Conventions
 *  - Positions are structure-of-arrays doubles: component d of particle i is
 *    pos[d * n + i] (n = that set's particle count).
 *  - Every pointer documented "device" is caller-allocated device memory; the
 *    library never frees caller memory. Pointers documented "host" are read
 *    before the call returns.
 *  - All device work is ordered on the context's stream and asynchronous
 *    unless stated otherwise; dcpse_sync() waits for it.
 *  - Errors: every entry point returns a dcpse_status; dcpse_last_error()
 *    gives the message. Per-particle assembly failures go to info[].
 *  - One context per host thread.
 *
 * Reference interfaces replaced (paths relative to the openfpm_numerics root):
 *  dcpse_support_count/fill  src/DCPSE/SupportBuilder.cuh:65-107
 *                            (SupportBuilderGPU::getSupport) and the
 *                            VL_NON_SYMMETRIC|VL_SKIP_REF_PART VerletList
 *                            used by every Dcpse test, e.g.
 *                            src/DCPSE/tests/Dcpse_unit_tests.cpp:97
 *  dcpse_assemble            src/DCPSE/Dcpse.hpp:690-809
 *                            (Dcpse::initializeStaticSize + computeKernel)
 *                            src/DCPSE/Dcpse.cuh:563-717 (Dcpse_gpu assembly)
 *  dcpse_apply_f32/f64       src/DCPSE/Dcpse.hpp:467-499 and :560-659
 *                            (Dcpse::computeDifferentialOperator) as driven by
 *                            src/DCPSE/DCPSE_op/DCPSE_op.hpp:64-67 (VECT_DCPSE)
 *                            and :389-401 (VECT_DCPSE_V_SUM, Laplacian)
 *  dcpse_transpose_*         the transpose of the operator whose rows
 *  dcpse_apply_transpose_*   src/DCPSE/DCPSE_op/DCPSE_op.hpp:69-83 (value_nz)
 *                            exports to src/DCPSE/DCPSE_op/DCPSE_Solver.hpp:978-1036
 *  dcpse_basis, dcpse_rhs    src/DCPSE/MonomialBasis.hpp:115-156,
 *                            src/DCPSE/DcpseRhs.hpp:27-60
 */
#ifndef DCPSE_HIP_H
#define DCPSE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dcpse_ctx dcpse_ctx;
typedef int32_t dcpse_status;

#define DCPSE_OK 0
#define DCPSE_ERR_INVALID 1     /* bad argument */
#define DCPSE_ERR_HIP 2         /* HIP runtime error */
#define DCPSE_ERR_UNSUPPORTED 3 /* configuration outside the compiled kernels */
#define DCPSE_ERR_RANGE 4       /* a size exceeds a documented limit */

/* per-particle assembly status written to info[] */
#define DCPSE_INFO_OK 0
#define DCPSE_INFO_TOO_FEW 1   /* fewer neighbours than monomials (Vandermonde.hpp:109-112) */
#define DCPSE_INFO_NOT_SPD 2   /* non-positive pivot in the FP64 Cholesky of A */
#define DCPSE_INFO_EPS_ZERO 3  /* all neighbours coincide with p (Vandermonde.hpp:80) */
#define DCPSE_INFO_NONFINITE 4 /* a kernel weight is not finite */

/* dcpse_assemble flags */
#define DCPSE_ASM_REFINE 1u /* one step of FP64 iterative refinement of the solve */

/* maximum neighbours per row handled by the support builder */
#define DCPSE_MAX_ROW 1024

/* ---- context ---------------------------------------------------------- */
int32_t dcpse_version(void);
dcpse_status dcpse_ctx_create(dcpse_ctx** ctx, int32_t device);
dcpse_status dcpse_ctx_destroy(dcpse_ctx* ctx);
/* Orders all further work of ctx on hip_stream, a caller-owned hipStream_t
 * (NULL = the device's default stream, which is also the initial binding). */
dcpse_status dcpse_ctx_set_stream(dcpse_ctx* ctx, void* hip_stream);
void* dcpse_ctx_get_stream(dcpse_ctx* ctx);
const char* dcpse_last_error(const dcpse_ctx* ctx);
dcpse_status dcpse_sync(dcpse_ctx* ctx);
/* Apply-kernel launch shape: lanes cooperating on one row, nonzeros each
 * lane loads before consuming, and load width in elements (1: 4-byte loads;
 * 4: 16-byte loads, FP32 only). Compiled shapes, width 1: 4x16, 8x8, 8x16,
 * 16x4, 16x8, 32x2, 64x1; width 4: 4x8, 8x4, 8x8, 8x12, 16x4, 16x8, 32x4.
 * group_lanes 0 =
 * automatic. FP64 applies ignore width 4 and use the width-1 shape with the
 * same lanes and elements when compiled (automatic otherwise).
 * block_threads: a multiple of 64 in 64..512. */
dcpse_status dcpse_set_apply_config(dcpse_ctx* ctx, int32_t group_lanes, int32_t elems_per_lane,
                                    int32_t vector_width, int32_t block_threads);
/* Assembly launch shape: wavefronts per block (must be 1) and particles each
 * wavefront assembles in sequence (0 = automatic, 18; values above 64 act as
 * 64, and on the static grid as 63). */
dcpse_status dcpse_set_assembly_config(dcpse_ctx* ctx, int32_t waves_per_block, int32_t particles_per_wave);
/* Assembly launch (kernels of src/DCPSE/Dcpse.hpp:690-809; replaces the
 * grid-stride assembleLocalMatrices_gpu<<<numSMs, 256>>> of
 * src/DCPSE/Dcpse.cuh:668-669). A wavefront assembles batches of
 * particles_per_batch particles (1..64, at most 63 on the static grid; 0 = 18),
 * one particle's ε and moments at a time over all 64 lanes, staging `chunk`
 * (32 or 64) neighbours per LDS round; two particles share the factorisation
 * when the basis fits half a wavefront, three when the shared system of a
 * nested pair fits 21 lanes. persistent != 0: one launch of as many wavefronts as the GPU
 * holds, each taking batches from a counter until none is left (0: one block
 * per batch). bucket_width > 0 (persistent grid only): batches ordered by row
 * length, longest first, in buckets of that width (particles keep their
 * index order within a bucket and tile of 2048); 0: index order. Results do
 * not depend on the shape. */
dcpse_status dcpse_set_assembly_launch(dcpse_ctx* ctx, int32_t persistent, int32_t particles_per_batch,
                                       int32_t chunk, int32_t bucket_width);

/* ---- host-only helpers (no device work) -------------------------------- */
/* Number of monomials m for signature beta[dim] and convergence order. */
int32_t dcpse_basis_size(int32_t dim, const uint32_t* beta, uint32_t order);
/* Writes m*dim exponents (host) in the reference enumeration order; returns
 * m, or -m if cap < m. */
int32_t dcpse_basis(int32_t dim, const uint32_t* beta, uint32_t order, uint32_t* exps, int32_t cap);
/* Writes the RHS b (host, m doubles); returns m, or -m if cap < m. */
int32_t dcpse_rhs(int32_t dim, const uint32_t* beta, uint32_t order, double* b, int32_t cap);

/* ---- neighbour support (Verlet CSR) ------------------------------------ */
/* Row p lists, in ascending index, every support particle q with
 * |x_p - x_q| < rcut_p (q != p when pos_dom == NULL, i.e. the domain set is
 * the support set; then n_dom must equal n_sup). A NULL pos_sup is accepted
 * when n_sup == 0 (an empty set has no rows, or rows without neighbours). rcut is a device array of 1 value (uniform) or n_dom
 * values (rcut_per_particle != 0). box_lo/box_hi: host, dim doubles; a
 * periodic dimension (bit d of periodic_mask) uses minimum-image
 * displacements and needs at least 3 cells of edge >= max rcut.
 * count: writes row_ptr[0..n_dom] (device int64) and *nnz (host; this call
 * synchronises the stream). fill: writes col[nnz] (device int32), for the
 * same arguments as the count and unchanged positions. With one set and one
 * cut-off, the context keeps the count's cell list and per-particle take
 * masks (about 64 bytes per particle) and a fill that directly follows a
 * count with the same arguments gathers its rows from them. */
dcpse_status dcpse_support_count(dcpse_ctx* ctx, int32_t dim,
                                 const double* pos_sup, int64_t n_sup,
                                 const double* pos_dom, int64_t n_dom,
                                 const double* rcut, int32_t rcut_per_particle,
                                 const double* box_lo, const double* box_hi, uint32_t periodic_mask,
                                 int64_t* row_ptr, int64_t* nnz);
dcpse_status dcpse_support_fill(dcpse_ctx* ctx, int32_t dim,
                                const double* pos_sup, int64_t n_sup,
                                const double* pos_dom, int64_t n_dom,
                                const double* rcut, int32_t rcut_per_particle,
                                const double* box_lo, const double* box_hi, uint32_t periodic_mask,
                                const int64_t* row_ptr, int32_t* col);
/* A domain that is a subset of the support set: domain particle p is
 * support particle dom_ids[p] (device int32[n_dom]) at pos_dom (device, its
 * positions gathered from pos_sup). Row p lists every support particle
 * q != dom_ids[p] with |x_p - x_q| < rcut_p, ascending: the rows of
 * vector_dist_subset::getVerlet<VL_NON_SYMMETRIC|VL_SKIP_REF_PART>, which the
 * reference's two-set operators run on
 * (src/DCPSE/DCPSE_op/tests/DCPSE_op_subset_test.cpp:113-141,
 * src/DCPSE/Dcpse.hpp:111-130). Same count/fill protocol as above. */
dcpse_status dcpse_support_count_subset(dcpse_ctx* ctx, int32_t dim,
                                        const double* pos_sup, int64_t n_sup,
                                        const double* pos_dom, const int32_t* dom_ids, int64_t n_dom,
                                        const double* rcut, int32_t rcut_per_particle,
                                        const double* box_lo, const double* box_hi, uint32_t periodic_mask,
                                        int64_t* row_ptr, int64_t* nnz);
dcpse_status dcpse_support_fill_subset(dcpse_ctx* ctx, int32_t dim,
                                       const double* pos_sup, int64_t n_sup,
                                       const double* pos_dom, const int32_t* dom_ids, int64_t n_dom,
                                       const double* rcut, int32_t rcut_per_particle,
                                       const double* box_lo, const double* box_hi, uint32_t periodic_mask,
                                       const int64_t* row_ptr, int32_t* col);

/* ---- kernel assembly (FP64) ------------------------------------------- */
/* Assembles sum_k coeffs[k] * D^{beta_k} for n_sig signatures that share one
 * monomial basis (same |beta|; e.g. the dim terms of the Laplacian), order
 * `order`, eps_p = mean_q |x_pq|_1 / h_over_eps (the reference uses 0.9).
 * betas/coeffs: host arrays (n_sig*dim, n_sig). Outputs (device, any may be
 * NULL): eps[n_dom], eps_inv_pow[n_dom] = eps^-|beta|, K64[nnz] raw kernel
 * (the reference's calcKernels), W32[nnz] and W64[nnz] = eps^-|beta| * K,
 * info[n_dom] (W64 left NULL costs nothing and is derived on demand by
 * dcpse_scale_rows_f64). pos_dom == NULL means the domain set is the support
 * set. */
dcpse_status dcpse_assemble(dcpse_ctx* ctx, int32_t dim,
                            const double* pos_sup, int64_t n_sup,
                            const double* pos_dom, int64_t n_dom,
                            const double* box_lo, const double* box_hi, uint32_t periodic_mask,
                            const int64_t* row_ptr, const int32_t* col,
                            int32_t n_sig, const uint32_t* betas, const double* coeffs,
                            uint32_t order, double h_over_eps, uint32_t flags,
                            double* eps, double* eps_inv_pow, double* K64,
                            float* W32, double* W64, int32_t* info);
/* n_ops operators on one support in one call: operator o has n_sig[o]
 * signatures (betas/coeffs concatenated over operators) and its own output
 * arrays (eps[o], ... info[o]; any array pointer or entry may be NULL).
 * Operators whose bases pair up at order 2 (a second derivative or the
 * Laplacian with a first derivative, two operators of one basis) are
 * assembled two per kernel pass, sharing the neighbour gathers, epsilon and
 * the moments, and a pair whose bases nest (the Laplacian and a first
 * derivative, two operators of one basis) one Cholesky factorisation - the
 * reference builds each Dcpse separately (DCPSE_op.hpp:649-786, 1350-1487).
 * Results equal separate dcpse_assemble calls to rounding (eps and the
 * status flags exactly). */
dcpse_status dcpse_assemble_multi(dcpse_ctx* ctx, int32_t dim,
                                  const double* pos_sup, int64_t n_sup,
                                  const double* pos_dom, int64_t n_dom,
                                  const double* box_lo, const double* box_hi, uint32_t periodic_mask,
                                  const int64_t* row_ptr, const int32_t* col, int32_t n_ops,
                                  const int32_t* n_sig, const uint32_t* betas, const double* coeffs,
                                  uint32_t order, double h_over_eps, uint32_t flags,
                                  double* const* eps, double* const* eps_inv_pow, double* const* K64,
                                  float* const* W32, double* const* W64, int32_t* const* info);

/* ---- operator application --------------------------------------------- */
/* For n_ops operators sharing one support, each with pre-scaled weights
 * W[o] (host array of device pointers):
 *   out[o][b*ld_out + p] = alpha * sum_j W[o][j] * (f[b*ld_f + col_j] - f_self[b*ld_self + p])
 *                          + beta * out[o][b*ld_out + p]
 * for rows p < n_rows and fields b < n_fields. f_self == NULL means f (the
 * domain set is the support set). beta == 0 never reads out. FP32 batches of
 * 4 or more fields of one operator on one set are copied particle-major into
 * the context's workspace (16 x n_rows floats, grown on first use) and
 * applied 16, 8 or 4 fields per pass; each field's result equals its
 * single-field apply bit for bit. The same holds for
 * dcpse_apply_transpose_f32. */
dcpse_status dcpse_apply_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                             int32_t n_ops, const float* const* W,
                             const float* f, int64_t ld_f, int32_t n_fields,
                             const float* f_self, int64_t ld_self,
                             float* const* out, int64_t ld_out, float alpha, float beta);
dcpse_status dcpse_apply_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                             int32_t n_ops, const double* const* W,
                             const double* f, int64_t ld_f, int32_t n_fields,
                             const double* f_self, int64_t ld_self,
                             double* const* out, int64_t ld_out, double alpha, double beta);

/* ---- transposed operator (adjoint), gather only, no float atomics ------ */
/* The forward operator is A = W - diag(rowsum W) (difference form). Its
 * transpose is stored as a CSR over the n_cols support particles:
 * rowT_ptr[n_cols+1], colT[nnz] (ascending), WT[nnz], plus diag[n_rows] =
 * -rowsum(W) (accumulated in FP64). Needs n_rows == n_cols (same set).
 * count writes rowT_ptr (device); fill writes colT, WT, diag. */
dcpse_status dcpse_transpose_count(dcpse_ctx* ctx, int64_t n_rows, int64_t n_cols,
                                   const int64_t* row_ptr, const int32_t* col, int64_t* rowT_ptr);
dcpse_status dcpse_transpose_fill_f32(dcpse_ctx* ctx, int64_t n_rows, int64_t n_cols,
                                      const int64_t* row_ptr, const int32_t* col, const float* W,
                                      const int64_t* rowT_ptr, int32_t* colT, float* WT, float* diag);
dcpse_status dcpse_transpose_fill_f64(dcpse_ctx* ctx, int64_t n_rows, int64_t n_cols,
                                      const int64_t* row_ptr, const int32_t* col, const double* W,
                                      const int64_t* rowT_ptr, int32_t* colT, double* WT, double* diag);
/* out[b*ld_out + q] = alpha * (sum_j WT[j] g[b*ld_g + colT_j] + diag[q] g[b*ld_g + q])
 *                     + beta * out[b*ld_out + q] */
dcpse_status dcpse_apply_transpose_f32(dcpse_ctx* ctx, int64_t n_cols, const int64_t* rowT_ptr,
                                       const int32_t* colT, const float* WT, const float* diag,
                                       const float* g, int64_t ld_g, int32_t n_fields,
                                       float* out, int64_t ld_out, float alpha, float beta);
dcpse_status dcpse_apply_transpose_f64(dcpse_ctx* ctx, int64_t n_cols, const int64_t* rowT_ptr,
                                       const int32_t* colT, const double* WT, const double* diag,
                                       const double* g, int64_t ld_g, int32_t n_fields,
                                       double* out, int64_t ld_out, double alpha, double beta);

/* ---- diagnostics -------------------------------------------------------- */
/* Per-particle diagnostics of an assembled operator (any argument of
 * dcpse_assemble, plus its outputs eps and K64), the reference's DCPSE_VERBOSE
 * statistics (src/DCPSE/Dcpse.hpp:786-795, minSpacing of
 * src/DCPSE/Vandermonde.hpp:63-81), checkMomenta's moment sums (Dcpse.hpp:381-457)
 * and the conditioning its conditionNumber measures (Dcpse.hpp:812-824):
 *   residual[p]    = max_alpha |sum_q K_pq m_alpha(x_pq/eps_p) - b_alpha| / max|b|
 *   cond1[p]       = 1-norm condition number estimate of the solved system
 *                    A = (E V)^T (E V) (exact ||A||_1, Hager-Higham estimate of
 *                    ||A^-1||_1 as LAPACK dpocon); +inf where A is not
 *                    positive definite or the particle has fewer than m neighbours
 *   min_spacing[p] = min_q |x_pq|_2
 * (device arrays of n_dom; any may be NULL) and summary (host, 10 doubles;
 * synchronises): neighbours min, mean, max; <h> = h_over_eps * mean eps;
 * min spacing mean, min, max; residual max; cond1 max over finite values;
 * the number of particles with an infinite cond1. A diagnostic pass, a few
 * times the assembly's cost; bases of up to 64 monomials. */
dcpse_status dcpse_operator_diagnostics(dcpse_ctx* ctx, int32_t dim,
                                        const double* pos_sup, int64_t n_sup,
                                        const double* pos_dom, int64_t n_dom,
                                        const double* box_lo, const double* box_hi, uint32_t periodic_mask,
                                        const int64_t* row_ptr, const int32_t* col,
                                        int32_t n_sig, const uint32_t* betas, const double* coeffs,
                                        uint32_t order, double h_over_eps, const double* eps, const double* K64,
                                        double* residual, double* cond1, double* min_spacing, double* summary);

/* ---- elementwise helper of the expression evaluation ------------------ */
/* y[i] = alpha * x[i] + beta * y[i] for i < n (beta == 0 never reads y); the
 * device counterpart of the per-particle SUM/MUL nodes the reference's host
 * evaluation loop applies (src/Operators/Vector/vector_dist_operators.hpp,
 * evaluated by src/Operators/Vector/cuda/vector_dist_operators_cuda.cuh:445-458). */
dcpse_status dcpse_axpby_f32(dcpse_ctx* ctx, int64_t n, float alpha, const float* x, float beta, float* y);
dcpse_status dcpse_axpby_f64(dcpse_ctx* ctx, int64_t n, double alpha, const double* x, double beta, double* y);

/* ---- sliced-ELL execution layout (FP32 apply) ------------------------- */
/* The same operator laid out one wavefront per slice of 64 consecutive rows:
 * slice s holds rows [64s, 64s+64) with width w_s = longest row rounded up to
 * 4. Within a slice the rows are ordered by decreasing length (ties by
 * index): lane l serves row sell_row[64s+l] (-1: no row), and slot j of lane
 * l sits at slice_ptr[s] + (j/4)*256 + 4l + (j%4). Padded slots hold column
 * 0 and weight 0. sell_col/sell_W must be 16-byte aligned. This is an
 * execution copy: the CSR arrays stay the operator's canonical storage
 * (calcKernels/kerOffsets).
 * count: writes slice_ptr[n_slices+1] and sell_row[64*n_slices] (device,
 * n_slices = ceil(n_rows/64)) and *n_slots (host, synchronises).
 * fill_cols/fill_weights_f32 write sell_col[n_slots] / sell_W[n_slots]
 * (device) from the CSR. */
dcpse_status dcpse_sell_count(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, int64_t* slice_ptr,
                              int32_t* sell_row, int64_t* n_slots);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments and results as dcpse_sell_count, but its
 * scratch (per-slice widths, overflow flag, scan tile sums) is a 4 MiB
 * static device array that each GPU holds until the process ends, so it
 * calls no hipMalloc/hipFree. n_rows above DCPSE_NOALLOC_MAX_ROWS is rejected
 * with DCPSE_ERR_RANGE. Every call on one GPU shares that GPU's array, so
 * calls that overlap on one GPU (from several contexts or host threads) race. */
#define DCPSE_NOALLOC_MAX_ROWS 33554432
dcpse_status dcpse_sell_count_noalloc(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, int64_t* slice_ptr,
                                      int32_t* sell_row, int64_t* n_slots);
dcpse_status dcpse_sell_fill_cols(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                                  const int64_t* slice_ptr, const int32_t* sell_row, int32_t* sell_col);
dcpse_status dcpse_sell_fill_weights_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const float* W,
                                         const int64_t* slice_ptr, const int32_t* sell_row, float* sell_W);
/* Same result as dcpse_apply_f32 with one field, from the SELL layout:
 * out[o][p] = alpha * sum_j sell_W[o][.] (f[col] - f_self[p]) + beta * out[o][p]. */
dcpse_status dcpse_apply_sell_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int64_t* slice_ptr,
                                  const int32_t* sell_row, const int32_t* sell_col, int32_t n_ops,
                                  const float* const* sell_W, const float* f, const float* f_self,
                                  float* const* out, float alpha, float beta);
/* FP64 weights in the same slot order (sell_W 32-byte aligned), and the FP64
 * apply from them and the columns of dcpse_sell_fill_cols: same result as
 * dcpse_apply_f64 with one field (FP64 sums in another order). */
dcpse_status dcpse_sell_fill_weights_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const double* W,
                                         const int64_t* slice_ptr, const int32_t* sell_row, double* sell_W);
dcpse_status dcpse_apply_sell_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int64_t* slice_ptr,
                                  const int32_t* sell_row, const int32_t* sell_col, int32_t n_ops,
                                  const double* const* sell_W, const double* f, const double* f_self,
                                  double* const* out, double alpha, double beta);
/* Interleaved form for n_ops (1..4) operators sharing the support: slot
 * group g of slice s is one block of (1 + n_ops) * 256 words - the 64 lanes'
 * columns, then each operator's weights (float bits) - starting at
 * (1 + n_ops) * slice_ptr[s]; bundle holds (1 + n_ops) * n_slots words
 * (16-byte aligned). The apply reads one contiguous span per slot group. */
dcpse_status dcpse_sell_fill_bundle_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                                        int32_t n_ops, const float* const* W, const int64_t* slice_ptr,
                                        const int32_t* sell_row, int32_t* bundle);
dcpse_status dcpse_apply_sell_bundle_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                         const int64_t* slice_ptr, const int32_t* sell_row, int32_t n_ops,
                                         const int32_t* bundle, const float* f, const float* f_self,
                                         float* const* out, float alpha, float beta);
/* Bundle with 16-bit columns (same set of n < 2^30 particles on both sides):
 * the column q of row p is stored as d = (q - p) taken modulo n into
 * [-n/2, n/2); slot group g of slice s is one block of 128 + 256 * n_ops
 * words (the 64 lanes' four 16-bit offsets, then each operator's weights)
 * starting at (128 + 256 * n_ops) * slice_ptr[s] / 256; bundle holds
 * (0.5 + n_ops) * n_slots words. Neighbours with |d| > 32767 are stored as
 * d = 0 and listed per lane instead: count writes exc_ptr[64*n_slices+1]
 * (device) and *n_exc (host, synchronises); fill writes the bundle,
 * sell_row16 (sell_row with bit 30 set on lanes that have such neighbours),
 * exc_col[n_exc] and exc_W[n_ops*n_exc] (weights of operator o at
 * o*n_exc + k). The apply adds the listed neighbours after the slot loop. */
dcpse_status dcpse_sell_bundle16_count(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                                       const int32_t* sell_row, int64_t* exc_ptr, int64_t* n_exc);
dcpse_status dcpse_sell_fill_bundle16_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                          const int32_t* col, int32_t n_ops, const float* const* W,
                                          const int64_t* slice_ptr, const int32_t* sell_row,
                                          const int64_t* exc_ptr, int64_t n_exc, int32_t* bundle,
                                          int32_t* sell_row16, int32_t* exc_col, float* exc_W);
dcpse_status dcpse_apply_sell_bundle16_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                           const int64_t* slice_ptr, const int32_t* sell_row16, int32_t n_ops,
                                           const int32_t* bundle, const int64_t* exc_ptr, const int32_t* exc_col,
                                           const float* exc_W, int64_t n_exc, const float* f, const float* f_self,
                                           float* const* out, float alpha, float beta);
/* The FP64 16-bit bundle: slot group g of slice s is one block of
 * 128 + 512 * n_ops words (the 64 lanes' four 16-bit offsets, then each
 * operator's four FP64 weights per lane) starting at
 * (128 + 512 * n_ops) * slice_ptr[s] / 256; bundle holds
 * (0.5 + 2 * n_ops) * n_slots words. Exceptions as for the FP32 bundle
 * (dcpse_sell_bundle16_count; exc_W holds FP64 weights). 18 bytes per
 * neighbour for one operator instead of 20 with 32-bit columns; the apply
 * (out = alpha * D f + beta * out, the FP64 kernel of src/DCPSE/Dcpse.hpp:560-659)
 * equals dcpse_apply_sell_f64 bit for bit when no row has far neighbours. */
dcpse_status dcpse_sell_fill_bundle16_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                          const int32_t* col, int32_t n_ops, const double* const* W,
                                          const int64_t* slice_ptr, const int32_t* sell_row,
                                          const int64_t* exc_ptr, int64_t n_exc, int32_t* bundle,
                                          int32_t* sell_row16, int32_t* exc_col, double* exc_W);
dcpse_status dcpse_apply_sell_bundle16_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                           const int64_t* slice_ptr, const int32_t* sell_row16, int32_t n_ops,
                                           const int32_t* bundle, const int64_t* exc_ptr, const int32_t* exc_col,
                                           const double* exc_W, int64_t n_exc, const double* f,
                                           const double* f_self, double* const* out, double alpha, double beta);
/* Same result as dcpse_apply_transpose_f32 with one field, from the SELL
 * layout of the transposed CSR (built with the three calls above). */
dcpse_status dcpse_apply_transpose_sell_f32(dcpse_ctx* ctx, int64_t n_cols, const int64_t* rowT_ptr,
                                            const int64_t* slice_ptr, const int32_t* sell_rowT,
                                            const int32_t* sell_colT, const float* sell_WT, const float* diag,
                                            const float* g, float* out, float alpha, float beta);
/* Same result as dcpse_apply_transpose_f32 with one field, from a 16-bit
 * bundle (dcpse_sell_bundle16_count / dcpse_sell_fill_bundle16_f32) of the
 * transposed weights WT on the operator's own row_ptr / col and slices: for
 * a structurally symmetric support (one set, one cut-off) the transposed CSR
 * has the same row_ptr and col. out = alpha (WT g + diag .* g) + beta out. */
dcpse_status dcpse_apply_transpose_bundle16_f32(dcpse_ctx* ctx, int64_t n, const int64_t* row_ptr,
                                                const int64_t* slice_ptr, const int32_t* sell_row16,
                                                const int32_t* bundle, const int64_t* exc_ptr,
                                                const int32_t* exc_col, const float* exc_W, int64_t n_exc,
                                                const float* diag, const float* g, float* out, float alpha,
                                                float beta);
/* Software pipeline of the SELL apply: groups of 4 slots per batch (2 or 4)
 * and stages (2: the stream loads of one batch overlap the gathers and FMAs
 * of the other). Default (2, 2). */
dcpse_status dcpse_set_sell_pipeline(dcpse_ctx* ctx, int32_t groups, int32_t depth);
/* Accumulation of the FP32 sliced-ELL forward applies (dcpse_apply_sell_f32,
 * dcpse_apply_sell_bundle_f32, dcpse_apply_sell_bundle16_f32): 64 (default)
 * sums the FP32 products (f_q - f_p) W_pq in FP64 and rounds each row once;
 * 32 does every operation in FP32, differences included, as the reference's
 * apply does for float properties (src/DCPSE/Dcpse.hpp:560-659). Transposed
 * applies, stages and the CSR kernels always sum in FP64. */
dcpse_status dcpse_set_apply_accumulation(dcpse_ctx* ctx, int32_t bits);

/* ---- sparse-matrix export -------------------------------------------- */
/* The assembled operator as a CSR matrix in the reference's value_nz
 * semantics (src/DCPSE/DCPSE_op/DCPSE_op.hpp:69-83, consumed by
 * src/DCPSE/DCPSE_op/DCPSE_Solver.hpp:978-1036): row p holds W_pq at every
 * neighbour q and sign * sum_q W_pq at column p (sign = getSign(): -1 for
 * even nonzero |beta|, +1 otherwise), columns ascending. Same-set operators
 * (n_rows == n_cols). Writes out_row_ptr[n_rows+1] (= row_ptr[p] + p),
 * out_col / out_val[nnz + n_rows] (device). The diagonal is summed in FP64. */
dcpse_status dcpse_export_csr_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                                  const double* W, double sign, int64_t* out_row_ptr, int32_t* out_col,
                                  double* out_val);
dcpse_status dcpse_export_csr_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                                  const float* W, double sign, int64_t* out_row_ptr, int32_t* out_col,
                                  float* out_val);

/* FP64 pre-scaled weights W[j] = row_scale[p] * K[j] for every neighbour j of
 * row p, one rounding per weight: the W64 output of dcpse_assemble (bit for
 * bit, with row_scale = its eps_inv_pow) for an operator assembled with
 * W64 = NULL. Replaces the per-row eps^-|beta| factor the reference applies
 * to each row sum (src/DCPSE/Dcpse.hpp:467-499). Device pointers; K and W of
 * row_ptr[n_rows] entries (non-null when n_rows > 0). */
dcpse_status dcpse_scale_rows_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const double* K,
                                  const double* row_scale, double* W);

/* ---- explicit time integration (df/dt = alpha * D f) ----------------- */
/* One stage of an explicit Runge-Kutta step as ONE launch: the gather
 * k = alpha * D u (difference form, one operator) with two axpys fused into
 * its epilogue: out_a = y_a + ca * k and out_b = y_b + cb * k (an output is
 * skipped when null; outputs may alias y_a/y_b but not u). Replaces the
 * per-stage vector kernels of src/OdeIntegrators/vector_algebra_ofp_gpu.hpp:13-362
 * that odeint's runge_kutta4 drives through src/OdeIntegrators/OdeIntegrators.hpp:63-135.
 * CSR (f32/f64) and sliced-ELL (f32) layouts as for the apply. */
dcpse_status dcpse_stage_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                             const float* W, const float* u, float alpha, const float* y_a, float ca, float* out_a,
                             const float* y_b, float cb, float* out_b);
dcpse_status dcpse_stage_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                             const double* W, const double* u, double alpha, const double* y_a, double ca,
                             double* out_a, const double* y_b, double cb, double* out_b);
dcpse_status dcpse_stage_sell_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int64_t* slice_ptr,
                                  const int32_t* sell_row, const int32_t* sell_col, const float* sell_W,
                                  const float* u, float alpha, const float* y_a, float ca, float* out_a,
                                  const float* y_b, float cb, float* out_b);
/* The stage on the 16-bit bundle of ONE operator (dcpse_sell_fill_bundle16_f32
 * with n_ops = 1; support and domain one particle set). */
dcpse_status dcpse_stage_bundle16_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                      const int64_t* slice_ptr, const int32_t* sell_row16, const int32_t* bundle,
                                      const int64_t* exc_ptr, const int32_t* exc_col, const float* exc_W,
                                      int64_t n_exc, const float* u, float alpha, const float* y_a, float ca,
                                      float* out_a, const float* y_b, float cb, float* out_b);
/* n_steps classic RK4 steps of df/dt = alpha * D f (the integration of
 * src/OdeIntegrators/tests/OdeIntegratores_base_tests.cpp:478-577 with
 * runge_kutta4), state f updated in place; 4 stage launches per step.
 * work: 3 * n_rows device elements of scratch. */
dcpse_status dcpse_rk4_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                           const float* W, float alpha, double dt, int64_t n_steps, float* f, float* work);
dcpse_status dcpse_rk4_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                           const double* W, double alpha, double dt, int64_t n_steps, double* f, double* work);
dcpse_status dcpse_rk4_sell_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int64_t* slice_ptr,
                                const int32_t* sell_row, const int32_t* sell_col, const float* sell_W, float alpha,
                                double dt, int64_t n_steps, float* f, float* work);
dcpse_status dcpse_rk4_bundle16_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                    const int64_t* slice_ptr, const int32_t* sell_row16, const int32_t* bundle,
                                    const int64_t* exc_ptr, const int32_t* exc_col, const float* exc_W,
                                    int64_t n_exc, float alpha, double dt, int64_t n_steps, float* f, float* work);

#ifdef __cplusplus
}
#endif

#endif /* DCPSE_HIP_H */
