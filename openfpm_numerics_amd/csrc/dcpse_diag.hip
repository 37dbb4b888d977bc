// Per-particle diagnostics of an assembled DC-PSE operator.
//
// The reference prints, under DCPSE_VERBOSE, the global average spacing
// <h> = c * mean(eps), and the mean and range of each particle's smallest
// neighbour distance (src/DCPSE/Dcpse.hpp:786-795, minSpacing from
// src/DCPSE/Vandermonde.hpp:63-81); checkMomenta (Dcpse.hpp:381-457) reports
// the moments sum_q K_pq m_alpha(x_pq / eps) that every particle should
// reproduce as b_alpha; conditionNumber (Dcpse.hpp:812-824) measures the
// conditioning of a Vandermonde matrix. Here one wavefront per particle
// recomputes, from the positions, eps and the assembled kernels:
//   * the moment residual max_alpha |sum_q K_pq m_alpha(z_q) - b_alpha| / max|b|;
//   * the smallest neighbour distance |x_pq|_2;
//   * an estimate of the 1-norm condition number of the system the assembly
//     solved, A = (E V)^T (E V) = sum_q w_q v_q v_q^T: A is formed and
//     Cholesky-factored in LDS, ||A||_1 is exact and ||A^-1||_1 comes from the
//     Hager-Higham estimator (the method of LAPACK's dpocon/dlacon), so the
//     estimate is a lower bound that is usually exact.
// A second kernel reduces them to the run's summary. This is a diagnostic
// pass (a few times the assembly's cost); the assembly kernels do not call it.
#include <algorithm>
#include <cmath>
#include <vector>

#include "dcpse_common.hpp"
#include "dcpse_internal.hpp"

namespace dcpse {

struct DiagArgs {
    const double* pos_sup;
    int64_t n_sup;
    const double* pos_dom;
    int64_t n_dom;
    const int64_t* row_ptr;
    const int32_t* col;
    double len[3], ilen[3];
    unsigned periodic;
    const double* eps;
    const double* K64;
    const uint32_t* exps;  // M * dim monomial exponents (device)
    const double* b;       // M right-hand side entries (device)
    int M;
    double bmax;
    double* residual;
    double* cond1;
    double* min_spacing;
};

__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}

__device__ __forceinline__ double ipow_d(double x, unsigned e) {
    double r = 1.0;
    for (unsigned k = 0; k < e; ++k) r *= x;
    return r;
}

// x = A^-1 r for the Cholesky factor L (lower triangle of the M x M row-major
// LDS matrix, diagonal L_kk): lane i holds r_i and returns x_i.
__device__ double chol_solve_lds(const double* L, int M, double r, int lane) {
    for (int k = 0; k < M; ++k) {
        if (lane == k) r /= L[k * M + k];
        const double yk = readlane_d(r, k);
        if (lane > k && lane < M) r = fma(-L[lane * M + k], yk, r);
    }
    for (int k = M - 1; k >= 0; --k) {
        if (lane == k) r /= L[k * M + k];
        const double xk = readlane_d(r, k);
        if (lane < k) r = fma(-L[k * M + lane], xk, r);
    }
    return lane < M ? r : 0.0;
}

template <int DIM, int MB>
__global__ __launch_bounds__(64) void k_diag(const DiagArgs a) {
    extern __shared__ double sh[];
    double* A = sh;            // MB x MB (M x M used, row-major, stride M)
    double* vq = sh + MB * MB;  // monomial values of the current neighbour
    const int lane = threadIdx.x;
    const int M = a.M;
    const bool act = lane < M;
    unsigned e[3] = {0u, 0u, 0u};
#pragma unroll
    for (int d = 0; d < DIM; ++d) e[d] = act ? a.exps[lane * DIM + d] : 0u;
    const double bi = act ? a.b[lane] : 0.0;
    for (int64_t p = blockIdx.x; p < a.n_dom; p += gridDim.x) {
        const int64_t rs = a.row_ptr[p];
        const int n = (int)(a.row_ptr[p + 1] - rs);
        const double eps = a.eps[p];
        double xp[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < DIM; ++d) xp[d] = a.pos_dom[d * a.n_dom + p];
        double arow[MB];
#pragma unroll
        for (int j = 0; j < MB; ++j) arow[j] = 0.0;
        double r = 0.0, dmin2 = INFINITY;
        const double ieps = eps > 0.0 ? 1.0 / eps : 0.0;
        for (int q = 0; q < n; ++q) {
            const int c = a.col[rs + q];
            double z[3] = {0.0, 0.0, 0.0}, z2 = 0.0, d2 = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const double x = min_image_r(xp[d] - a.pos_sup[d * a.n_sup + c], a.len[d], a.ilen[d],
                                             (a.periodic >> d) & 1u);
                d2 = fma(x, x, d2);
                z[d] = x * ieps;
                z2 = fma(z[d], z[d], z2);
            }
            dmin2 = fmin(dmin2, d2);
            const double w = exp(-z2);
            double v = 1.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) v *= ipow_d(z[d], e[d]);
            v = act ? v : 0.0;
            r = fma(a.K64[rs + q], v, r);
            vq[lane] = v;
            wave_lds_sync();
            const double wv = w * v;
#pragma unroll
            for (int j = 0; j < MB; ++j)
                if (j < M) arow[j] = fma(wv, vq[j], arow[j]);
            wave_lds_sync();
        }
        const double res = wave_max_d(act ? fabs(r - bi) : 0.0) / a.bmax;
        double rowsum = 0.0;
#pragma unroll
        for (int j = 0; j < MB; ++j)
            if (j < M) rowsum += fabs(arow[j]);
        const double anorm = wave_max_d(act ? rowsum : 0.0);  // ||A||_1 = ||A||_inf (symmetric)
        if (act) {
#pragma unroll
            for (int j = 0; j < MB; ++j)
                if (j < M) A[lane * M + j] = arow[j];
        }
        wave_lds_sync();
        // Cholesky A = L L^T in place (lower triangle)
        bool ok = n >= M && eps > 0.0;
        for (int k = 0; k < M && ok; ++k) {
            const double dkk = A[k * M + k];
            if (!(dkk > 0.0) || !isfinite(dkk)) {
                ok = false;
                break;
            }
            const double lkk = sqrt(dkk);
            wave_lds_sync();
            if (lane > k && lane < M) A[lane * M + k] /= lkk;
            if (lane == k) A[k * M + k] = lkk;
            wave_lds_sync();
            if (lane > k && lane < M) {
                const double lik = A[lane * M + k];
                for (int j = k + 1; j <= lane; ++j) A[lane * M + j] = fma(-lik, A[j * M + k], A[lane * M + j]);
            }
            wave_lds_sync();
        }
        double cond = INFINITY;
        if (ok) {
            // Hager-Higham estimate of ||A^-1||_1 (A symmetric: A^-T = A^-1)
            double x = act ? 1.0 / M : 0.0, est = 0.0;
            for (int it = 0; it < 5; ++it) {
                const double y = chol_solve_lds(A, M, x, lane);
                const double est_new = wave_sum(fabs(y));
                if (it > 0 && est_new <= est) break;
                est = est_new;
                const double xi = act ? (y >= 0.0 ? 1.0 : -1.0) : 0.0;
                const double zz = chol_solve_lds(A, M, xi, lane);
                const double zmax = wave_max_d(fabs(zz));
                const double ztx = wave_sum(zz * x);
                if (it > 0 && zmax <= ztx) break;
                const unsigned long long at = __ballot(act && fabs(zz) == zmax);
                const int j = at ? __builtin_ctzll(at) : 0;
                x = lane == j ? 1.0 : 0.0;
            }
            // the alternative estimate of dlacon: x_i = (-1)^i (1 + i / (M - 1))
            const double xa = act ? ((lane & 1) ? -1.0 : 1.0) * (1.0 + (M > 1 ? (double)lane / (M - 1) : 0.0)) : 0.0;
            const double ya = chol_solve_lds(A, M, xa, lane);
            est = fmax(est, 2.0 * wave_sum(fabs(ya)) / (3.0 * M));
            cond = anorm * est;
        }
        wave_lds_sync();
        if (lane == 0) {
            a.residual[p] = res;
            a.cond1[p] = cond;
            a.min_spacing[p] = n > 0 ? sqrt(dmin2) : 0.0;
        }
    }
}

// Summary over particles (one block): n min / mean / max, <h> = c mean(eps)
// over assembled particles, min spacing mean / min / max, residual max, cond1
// max, the count of particles whose system could not be factored.
constexpr int kSumThreads = 256;

__global__ __launch_bounds__(kSumThreads) void k_diag_summary(int64_t n_dom, const int64_t* __restrict__ row_ptr,
                                                       const double* __restrict__ eps, double h_over_eps,
                                                       const double* __restrict__ residual,
                                                       const double* __restrict__ cond1,
                                                       const double* __restrict__ min_spacing,
                                                       double* __restrict__ out) {
    constexpr int NV = 10;
    __shared__ double red[NV][kSumThreads];
    double nmin = INFINITY, nsum = 0.0, nmax = 0.0, esum = 0.0, ecnt = 0.0, ssum = 0.0, smin = INFINITY,
           smax = 0.0, rmax = 0.0, cmax = 0.0;
    double bad = 0.0;
    for (int64_t p = threadIdx.x; p < n_dom; p += blockDim.x) {
        const double n = (double)(row_ptr[p + 1] - row_ptr[p]);
        nmin = fmin(nmin, n);
        nmax = fmax(nmax, n);
        nsum += n;
        if (eps[p] > 0.0) {
            esum += eps[p];
            ecnt += 1.0;
        }
        ssum += min_spacing[p];
        smin = fmin(smin, min_spacing[p]);
        smax = fmax(smax, min_spacing[p]);
        if (residual[p] == residual[p]) rmax = fmax(rmax, residual[p]);
        if (isfinite(cond1[p])) cmax = fmax(cmax, cond1[p]);
        else bad += 1.0;
    }
    const int t = threadIdx.x;
    red[0][t] = nmin; red[1][t] = nsum; red[2][t] = nmax; red[3][t] = esum; red[4][t] = ecnt;
    red[5][t] = ssum; red[6][t] = smin; red[7][t] = smax; red[8][t] = rmax; red[9][t] = cmax;
    __syncthreads();
    for (int off = kSumThreads / 2; off > 0; off >>= 1) {
        if (t < off) {
            red[0][t] = fmin(red[0][t], red[0][t + off]);
            red[1][t] += red[1][t + off];
            red[2][t] = fmax(red[2][t], red[2][t + off]);
            red[3][t] += red[3][t + off];
            red[4][t] += red[4][t + off];
            red[5][t] += red[5][t + off];
            red[6][t] = fmin(red[6][t], red[6][t + off]);
            red[7][t] = fmax(red[7][t], red[7][t + off]);
            red[8][t] = fmax(red[8][t], red[8][t + off]);
            red[9][t] = fmax(red[9][t], red[9][t + off]);
        }
        __syncthreads();
    }
    __shared__ double badsum[kSumThreads];
    badsum[t] = bad;
    __syncthreads();
    for (int off = kSumThreads / 2; off > 0; off >>= 1) {
        if (t < off) badsum[t] += badsum[t + off];
        __syncthreads();
    }
    if (t == 0) {
        const double nd = n_dom > 0 ? (double)n_dom : 1.0;
        out[0] = n_dom > 0 ? red[0][0] : 0.0;
        out[1] = red[1][0] / nd;
        out[2] = red[2][0];
        out[3] = red[4][0] > 0.0 ? h_over_eps * red[3][0] / red[4][0] : 0.0;
        out[4] = red[5][0] / nd;
        out[5] = n_dom > 0 ? red[6][0] : 0.0;
        out[6] = red[7][0];
        out[7] = red[8][0];
        out[8] = red[9][0];
        out[9] = badsum[0];
    }
}

template <int DIM, int MB>
static void launch_diag(const DiagArgs& a, int64_t blocks, hipStream_t s) {
    k_diag<DIM, MB><<<(unsigned)blocks, kWave, (size_t)(MB * MB + kWave) * sizeof(double), s>>>(a);
}

}  // namespace dcpse

using namespace dcpse;

extern "C" {

dcpse_status dcpse_operator_diagnostics(dcpse_ctx* ctx, int32_t dim, const double* pos_sup, int64_t n_sup,
                                        const double* pos_dom, int64_t n_dom, const double* box_lo,
                                        const double* box_hi, uint32_t periodic_mask, const int64_t* row_ptr,
                                        const int32_t* col, int32_t n_sig, const uint32_t* betas,
                                        const double* coeffs, uint32_t order, double h_over_eps, const double* eps,
                                        const double* K64, double* residual, double* cond1, double* min_spacing,
                                        double* summary) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (dim != 2 && dim != 3) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "dim must be 2 or 3");
    if (!pos_sup || !row_ptr || !eps || !K64 || !betas || n_sig < 1 || n_sup < 0 || n_dom < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    if (periodic_mask && (!box_lo || !box_hi))
        return fail(ctx, DCPSE_ERR_INVALID, "periodic boundaries need box_lo/box_hi");
    // a NULL domain is the support set itself: its row count is n_sup (a
    // smaller n_dom would size the outputs for fewer rows than are written)
    if (pos_dom == nullptr && n_dom != n_sup)
        return fail(ctx, DCPSE_ERR_INVALID, "pos_dom NULL (the support set as domain) needs n_dom == n_sup");
    if (pos_dom == nullptr) pos_dom = pos_sup;
    const int M = dcpse_basis_size(dim, betas, order);
    if (M <= 0 || M > 64) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "diagnostics need a basis of 1..64 monomials");
    std::vector<uint32_t> ex((size_t)M * dim);
    dcpse_basis(dim, betas, order, ex.data(), M);
    std::vector<double> b((size_t)M, 0.0), bk((size_t)M);
    for (int k = 0; k < n_sig; ++k) {
        if (dcpse_basis_size(dim, betas + (size_t)k * dim, order) != M)
            return fail(ctx, DCPSE_ERR_INVALID, "combined signatures must share one basis");
        dcpse_rhs(dim, betas + (size_t)k * dim, order, bk.data(), M);
        const double c = coeffs ? coeffs[k] : 1.0;
        for (int i = 0; i < M; ++i) b[(size_t)i] += c * bk[(size_t)i];
    }
    double bmax = 0.0;
    for (double v : b) bmax = std::max(bmax, std::fabs(v));
    if (bmax == 0.0) bmax = 1.0;
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    Workspace w;
    const size_t need = (size_t)M * dim * 4 + (size_t)M * 8 + 3 * (size_t)std::max<int64_t>(n_dom, 1) * 8 + 16 * 8 +
                        6 * 256;
    dcpse_status st = reserve_workspace(ctx, need, &w);
    if (st) return st;
    uint32_t* d_ex = w.take<uint32_t>((size_t)M * dim);
    double* d_b = w.take<double>((size_t)M);
    double* r_ws = w.take<double>((size_t)std::max<int64_t>(n_dom, 1));
    double* c_ws = w.take<double>((size_t)std::max<int64_t>(n_dom, 1));
    double* s_ws = w.take<double>((size_t)std::max<int64_t>(n_dom, 1));
    double* d_sum = w.take<double>(16);
    if (!d_sum) return fail(ctx, DCPSE_ERR_INVALID, "diagnostics workspace sizing error");
    hipStream_t s = ctx->stream;
    DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(d_ex, ex.data(), ex.size() * 4, hipMemcpyHostToDevice, s));
    DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(d_b, b.data(), b.size() * 8, hipMemcpyHostToDevice, s));
    DiagArgs a;
    a.pos_sup = pos_sup;
    a.n_sup = n_sup;
    a.pos_dom = pos_dom;
    a.n_dom = n_dom;
    a.row_ptr = row_ptr;
    a.col = col;
    for (int d = 0; d < 3; ++d) {
        a.len[d] = (d < dim && box_lo && box_hi) ? box_hi[d] - box_lo[d] : 1.0;
        a.ilen[d] = 1.0 / a.len[d];
    }
    a.periodic = periodic_mask;
    a.eps = eps;
    a.K64 = K64;
    a.exps = d_ex;
    a.b = d_b;
    a.M = M;
    a.bmax = bmax;
    a.residual = residual ? residual : r_ws;
    a.cond1 = cond1 ? cond1 : c_ws;
    a.min_spacing = min_spacing ? min_spacing : s_ws;
    if (n_dom > 0) {
        const int64_t blocks = std::min<int64_t>(n_dom, (int64_t)ctx->num_cus * 16);
        if (dim == 2) {
            if (M <= 16) launch_diag<2, 16>(a, blocks, s);
            else if (M <= 32) launch_diag<2, 32>(a, blocks, s);
            else launch_diag<2, 64>(a, blocks, s);
        } else {
            if (M <= 16) launch_diag<3, 16>(a, blocks, s);
            else if (M <= 32) launch_diag<3, 32>(a, blocks, s);
            else launch_diag<3, 64>(a, blocks, s);
        }
        DCPSE_LAUNCH_CHECK(ctx);
    }
    k_diag_summary<<<1, kSumThreads, 0, s>>>(n_dom, row_ptr, eps, h_over_eps, a.residual, a.cond1, a.min_spacing, d_sum);
    DCPSE_LAUNCH_CHECK(ctx);
    if (summary) {
        DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(summary, d_sum, 10 * sizeof(double), hipMemcpyDeviceToHost, s));
        DCPSE_HIP_CHECK(ctx, hipStreamSynchronize(s));
    }
    return DCPSE_OK;
}

}  // extern "C"
