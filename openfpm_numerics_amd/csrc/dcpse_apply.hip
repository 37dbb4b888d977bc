// DC-PSE operator application for gfx950: CSR neighbour gather.
//
// Restates the bulk apply of src/DCPSE/Dcpse.hpp:467-499 (and the per-key
// form :560-659 that the expression nodes of
// src/DCPSE/DCPSE_op/DCPSE_op.hpp:64-67 and :389-401 call) as
//     Df_p = sum_q W_pq (f_q - f_p),   W = eps_p^-|beta| K_pq (pre-scaled),
// the difference form: for odd |beta| the constant monomial is in the basis
// and its moment condition gives sum_q K_pq = 0, so the reference's
// (f_q + f_p) equals (f_q - f_p) up to rounding; for even |beta| it is the
// reference's own sign.
//
// The kernel is HBM-bound (about 8 bytes per nonzero per operator). G lanes
// cooperate on one row so the W/col streams are read as contiguous segments;
// several operators sharing one support (e.g. Laplacian and Dx) and several
// fields share the col stream and the f gather in one pass. Workgroups are
// remapped so that each XCD walks one contiguous slice of rows, keeping the
// f halo of a slice in that XCD's L2.
#include <algorithm>
#include <type_traits>

#include "dcpse_common.hpp"
#include "dcpse_internal.hpp"

namespace dcpse {

// Loads of the col/W streams: ordinary cached loads (rows are not aligned to
// cache lines, so a line split between two load instructions is found in L2 by
// the second).
#define DCPSE_STREAM_LOAD(p) (*(p))

constexpr int kMaxOps = 4;

template <typename T>
struct ApplyArgs {
    int64_t n_rows;
    const int64_t* row_ptr;
    const int32_t* col;
    const T* W[kMaxOps];
    const T* f;
    int64_t ld_f;
    const T* f_self;
    int64_t ld_self;
    const T* diag;  // transpose mode only
    T* out[kMaxOps];
    int64_t ld_out;
    T alpha, beta;
    int64_t blocks_needed;  // logical blocks
    int64_t per_xcd;        // logical blocks per XCD slice
};

enum ApplyMode { kDiff = 0, kExplicitDiag = 1 };

constexpr int kApplyMaxBlock = 512;  // launch bound of k_apply; larger blocks are rejected

// G lanes cooperate on one row. The first G*EPL nonzeros of the row are
// handled with every load issued before any is consumed: EPL col and
// NOPS*EPL weight loads per lane (indices clamped into the row, so no branch
// separates them), then EPL gathers of f, then the FMAs. Longer rows finish
// in a strided loop. Invalid (clamped) slots are masked out by a select.
template <typename T, int G, int EPL, int NOPS, int NF, int MODE>
__global__ __launch_bounds__(kApplyMaxBlock) void k_apply(const ApplyArgs<T> a) {
    // XCD-aware remap: hardware block b runs on XCD b % 8; give each XCD a
    // contiguous range of logical blocks.
    const int64_t hb = blockIdx.x;
    const int64_t lb = (hb & 7) * a.per_xcd + (hb >> 3);
    if (lb >= a.blocks_needed) return;
    const int rows_per_block = blockDim.x / G;
    const int gl = threadIdx.x % G;
    const int64_t row = lb * rows_per_block + threadIdx.x / G;
    if (row >= a.n_rows) return;  // whole G-lane groups leave together
    const int64_t rs = a.row_ptr[row];
    const int64_t re = a.row_ptr[row + 1];
    T fp[NF];
#pragma unroll
    for (int b = 0; b < NF; ++b) fp[b] = a.f_self[b * a.ld_self + row];
    double acc[NOPS][NF];  // FP64 accumulation for both value types
#pragma unroll
    for (int o = 0; o < NOPS; ++o)
#pragma unroll
        for (int b = 0; b < NF; ++b) acc[o][b] = 0.0;
    if (re > rs) {
        int32_t q[EPL];
        T w[NOPS][EPL];
        bool v[EPL];
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            const int64_t j = rs + gl + k * G;
            v[k] = j < re;
            const int64_t jj = v[k] ? j : re - 1;
            q[k] = DCPSE_STREAM_LOAD(a.col + jj);
#pragma unroll
            for (int o = 0; o < NOPS; ++o) w[o][k] = DCPSE_STREAM_LOAD(a.W[o] + jj);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int b = 0; b < NF; ++b) {
            T fq[EPL];
#pragma unroll
            for (int k = 0; k < EPL; ++k) fq[k] = a.f[b * a.ld_f + q[k]];
#pragma unroll
            for (int k = 0; k < EPL; ++k) {
                const double d = MODE == kDiff ? (double)fq[k] - (double)fp[b] : (double)fq[k];
#pragma unroll
                for (int o = 0; o < NOPS; ++o) acc[o][b] = v[k] ? fma((double)w[o][k], d, acc[o][b]) : acc[o][b];
            }
        }
        for (int64_t j = rs + gl + G * EPL; j < re; j += G) {
            const int32_t qq = DCPSE_STREAM_LOAD(a.col + j);
            T ww[NOPS];
#pragma unroll
            for (int o = 0; o < NOPS; ++o) ww[o] = DCPSE_STREAM_LOAD(a.W[o] + j);
#pragma unroll
            for (int b = 0; b < NF; ++b) {
                const T fq = a.f[b * a.ld_f + qq];
                const double d = MODE == kDiff ? (double)fq - (double)fp[b] : (double)fq;
#pragma unroll
                for (int o = 0; o < NOPS; ++o) acc[o][b] = fma((double)ww[o], d, acc[o][b]);
            }
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
        for (int o = 0; o < NOPS; ++o)
#pragma unroll
            for (int b = 0; b < NF; ++b) acc[o][b] += __shfl_xor(acc[o][b], off, kWave);
    if (gl == 0) {
#pragma unroll
        for (int o = 0; o < NOPS; ++o)
#pragma unroll
            for (int b = 0; b < NF; ++b) {
                double val = acc[o][b];
                if (MODE == kExplicitDiag) val = fma((double)a.diag[row], (double)fp[b], val);
                T* dst = a.out[o] + b * a.ld_out + row;
                val *= (double)a.alpha;
                if (a.beta != T(0)) val = fma((double)a.beta, (double)*dst, val);
                *dst = (T)val;
            }
    }
}

// 4-byte-aligned 4-vectors: one 16-byte load per lane at any element offset.
typedef int32_t vi4u __attribute__((ext_vector_type(4), aligned(4)));
typedef float vf4u __attribute__((ext_vector_type(4), aligned(4)));

// FP32 variant: each lane loads 4 consecutive nonzeros of its row with one
// 16-byte load per stream (col and every weight array), so a wavefront
// instruction moves a contiguous 64*G-byte segment per row. NCH such chunks per
// lane are issued before the gathers; longer rows finish in a strided loop.
// The last rows of the CSR, whose 16-byte loads could pass the end of the
// arrays, take the scalar path.
template <int G, int NCH, int NOPS, int NF, int MODE>
__global__ __launch_bounds__(kApplyMaxBlock) void k_apply_v4(const ApplyArgs<float> a) {
    const int64_t hb = blockIdx.x;
    const int64_t lb = (hb & 7) * a.per_xcd + (hb >> 3);
    if (lb >= a.blocks_needed) return;
    const int rows_per_block = blockDim.x / G;
    const int gl = threadIdx.x % G;
    const int64_t row = lb * rows_per_block + threadIdx.x / G;
    if (row >= a.n_rows) return;
    const int64_t rs = a.row_ptr[row];
    const int64_t re = a.row_ptr[row + 1];
    const int64_t nnz_total = a.row_ptr[a.n_rows];
    float fp[NF];
#pragma unroll
    for (int b = 0; b < NF; ++b) fp[b] = a.f_self[b * a.ld_self + row];
    double acc[NOPS][NF];  // FP64 accumulation of the FP32 products
#pragma unroll
    for (int o = 0; o < NOPS; ++o)
#pragma unroll
        for (int b = 0; b < NF; ++b) acc[o][b] = 0.0;
    int64_t jtail = rs + gl;  // first element of the strided scalar loop
    // 16-byte loads read elements e..e+3 with e <= re-1, so re + 3 <= nnz_total
    // keeps them inside the arrays; empty rows skip the vector section.
    if (re > rs && re + 3 <= nnz_total) {
        vi4u q[NCH];
        vf4u w[NOPS][NCH];
        int64_t e0[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            e0[c] = rs + 4 * (gl + c * G);
            const int64_t ee = e0[c] < re ? e0[c] : re - 1;  // clamped chunks are masked below
            q[c] = DCPSE_STREAM_LOAD((const vi4u*)(a.col + ee));
#pragma unroll
            for (int o = 0; o < NOPS; ++o) w[o][c] = DCPSE_STREAM_LOAD((const vf4u*)(a.W[o] + ee));
        }
        // keep every stream load ahead of the gathers that depend on col
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int b = 0; b < NF; ++b) {
            float fq[NCH][4];
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                fq[c][0] = a.f[b * a.ld_f + q[c].x];
                fq[c][1] = a.f[b * a.ld_f + q[c].y];
                fq[c][2] = a.f[b * a.ld_f + q[c].z];
                fq[c][3] = a.f[b * a.ld_f + q[c].w];
            }
#pragma unroll
            for (int c = 0; c < NCH; ++c)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const bool v = e0[c] + i < re;
                    const double d = MODE == kDiff ? (double)fq[c][i] - (double)fp[b] : (double)fq[c][i];
#pragma unroll
                    for (int o = 0; o < NOPS; ++o) acc[o][b] = v ? fma((double)w[o][c][i], d, acc[o][b]) : acc[o][b];
                }
        }
        jtail = rs + 4 * G * NCH + gl;
    }
    for (int64_t j = jtail; j < re; j += G) {
        const int32_t qq = DCPSE_STREAM_LOAD(a.col + j);
        float ww[NOPS];
#pragma unroll
        for (int o = 0; o < NOPS; ++o) ww[o] = DCPSE_STREAM_LOAD(a.W[o] + j);
#pragma unroll
        for (int b = 0; b < NF; ++b) {
            const float fq = a.f[b * a.ld_f + qq];
            const double d = MODE == kDiff ? (double)fq - (double)fp[b] : (double)fq;
#pragma unroll
            for (int o = 0; o < NOPS; ++o) acc[o][b] = fma((double)ww[o], d, acc[o][b]);
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
        for (int o = 0; o < NOPS; ++o)
#pragma unroll
            for (int b = 0; b < NF; ++b) acc[o][b] += __shfl_xor(acc[o][b], off, kWave);
    if (gl == 0) {
#pragma unroll
        for (int o = 0; o < NOPS; ++o)
#pragma unroll
            for (int b = 0; b < NF; ++b) {
                double val = acc[o][b];
                if (MODE == kExplicitDiag) val = fma((double)a.diag[row], (double)fp[b], val);
                float* dst = a.out[o] + b * a.ld_out + row;
                val *= (double)a.alpha;
                if (a.beta != 0.0f) val = fma((double)a.beta, (double)*dst, val);
                *dst = (float)val;
            }
    }
}

// Batches of FP32 fields on one set: the NF fields are first interleaved
// particle-major (fil[q * NF + b]), so a neighbour's NF values are one
// contiguous 4 NF-byte read instead of NF separate gathers, and the col/W
// streams are read once for all of them. Each field's sum runs over the
// nonzeros in the order of k_apply_v4 (same lanes, same chunks, same
// reduction), so the results equal the field-major kernel's bit for bit.
template <int NF>
__global__ __launch_bounds__(256) void k_interleave(int64_t n, const float* __restrict__ f, int64_t ld,
                                                    float* __restrict__ fil) {
    __shared__ float t[NF][257];
    const int64_t q0 = (int64_t)blockIdx.x * 256;
    const int64_t q = q0 + threadIdx.x;
#pragma unroll
    for (int b = 0; b < NF; ++b) t[b][threadIdx.x] = q < n ? f[b * ld + q] : 0.0f;
    __syncthreads();
    for (int k = threadIdx.x; k < 256 * NF; k += 256) {
        const int qq = k / NF, b = k % NF;
        if (q0 + qq < n) fil[(q0 + qq) * NF + b] = t[b][qq];
    }
}

typedef float vf4a __attribute__((ext_vector_type(4)));

template <int G, int NCH, int NF, int MODE>
__global__ __launch_bounds__(kApplyMaxBlock) void k_apply_il(const ApplyArgs<float> a) {
    static_assert(NF % 4 == 0, "fields in groups of 4");
    const int64_t hb = blockIdx.x;
    const int64_t lb = (hb & 7) * a.per_xcd + (hb >> 3);
    if (lb >= a.blocks_needed) return;
    const int rows_per_block = blockDim.x / G;
    const int gl = threadIdx.x % G;
    const int64_t row = lb * rows_per_block + threadIdx.x / G;
    if (row >= a.n_rows) return;
    const int64_t rs = a.row_ptr[row];
    const int64_t re = a.row_ptr[row + 1];
    const int64_t nnz_total = a.row_ptr[a.n_rows];
    auto gather = [&](int64_t q, float* v) {
        const vf4a* p = reinterpret_cast<const vf4a*>(a.f + q * NF);
#pragma unroll
        for (int c = 0; c < NF / 4; ++c) {
            const vf4a x = p[c];
            v[4 * c] = x.x;
            v[4 * c + 1] = x.y;
            v[4 * c + 2] = x.z;
            v[4 * c + 3] = x.w;
        }
    };
    float fp[NF];
    gather(row, fp);
    double acc[NF];
#pragma unroll
    for (int b = 0; b < NF; ++b) acc[b] = 0.0;
    int64_t jtail = rs + gl;
    if (re > rs && re + 3 <= nnz_total) {
        vi4u q[NCH];
        vf4u w[NCH];
        int64_t e0[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            e0[c] = rs + 4 * (gl + c * G);
            const int64_t ee = e0[c] < re ? e0[c] : re - 1;
            q[c] = DCPSE_STREAM_LOAD((const vi4u*)(a.col + ee));
            w[c] = DCPSE_STREAM_LOAD((const vf4u*)(a.W[0] + ee));
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < NCH; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const bool v = e0[c] + i < re;
                float fq[NF];
                gather(q[c][i], fq);
#pragma unroll
                for (int b = 0; b < NF; ++b) {
                    const double d = MODE == kDiff ? (double)fq[b] - (double)fp[b] : (double)fq[b];
                    acc[b] = v ? fma((double)w[c][i], d, acc[b]) : acc[b];
                }
            }
        jtail = rs + 4 * G * NCH + gl;
    }
    for (int64_t j = jtail; j < re; j += G) {
        const int32_t qq = DCPSE_STREAM_LOAD(a.col + j);
        const float ww = DCPSE_STREAM_LOAD(a.W[0] + j);
        float fq[NF];
        gather(qq, fq);
#pragma unroll
        for (int b = 0; b < NF; ++b) {
            const double d = MODE == kDiff ? (double)fq[b] - (double)fp[b] : (double)fq[b];
            acc[b] = fma((double)ww, d, acc[b]);
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
        for (int b = 0; b < NF; ++b) acc[b] += __shfl_xor(acc[b], off, kWave);
    if (gl == 0) {
#pragma unroll
        for (int b = 0; b < NF; ++b) {
            double val = acc[b];
            if (MODE == kExplicitDiag) val = fma((double)a.diag[row], (double)fp[b], val);
            float* dst = a.out[0] + b * a.ld_out + row;
            val *= (double)a.alpha;
            if (a.beta != 0.0f) val = fma((double)a.beta, (double)*dst, val);
            *dst = (float)val;
        }
    }
}

template <int G, int NCH, int NF, int MODE>
static void launch_apply_il(ApplyArgs<float> a, int block, hipStream_t s) {
    const int rows_per_block = block / G;
    a.blocks_needed = (a.n_rows + rows_per_block - 1) / rows_per_block;
    a.per_xcd = (a.blocks_needed + 7) / 8;
    k_apply_il<G, NCH, NF, MODE><<<(unsigned)(a.per_xcd * 8), block, 0, s>>>(a);
}

// The interleaved batch kernel for the shapes of the 16-byte-load kernel
// (il_shape: whether one exists).
static bool il_shape(int G, int epl) {
    return (G == 4 && epl == 8) || (G == 8 && (epl == 4 || epl == 8 || epl == 12)) ||
           (G == 16 && (epl == 4 || epl == 8)) || (G == 32 && epl == 4);
}
template <int MODE, int NF>
static bool dispatch_il(const ApplyArgs<float>& a, int G, int epl, int block, hipStream_t s) {
    if (G == 4 && epl == 8) return launch_apply_il<4, 2, NF, MODE>(a, block, s), true;
    if (G == 8 && epl == 4) return launch_apply_il<8, 1, NF, MODE>(a, block, s), true;
    if (G == 8 && epl == 8) return launch_apply_il<8, 2, NF, MODE>(a, block, s), true;
    if (G == 8 && epl == 12) return launch_apply_il<8, 3, NF, MODE>(a, block, s), true;
    if (G == 16 && epl == 4) return launch_apply_il<16, 1, NF, MODE>(a, block, s), true;
    if (G == 16 && epl == 8) return launch_apply_il<16, 2, NF, MODE>(a, block, s), true;
    if (G == 32 && epl == 4) return launch_apply_il<32, 1, NF, MODE>(a, block, s), true;
    return false;
}

template <typename T, int G, int EPL, int NOPS, int NF, int MODE>
static void launch_apply(ApplyArgs<T> a, int block, hipStream_t s) {
    const int rows_per_block = block / G;
    a.blocks_needed = (a.n_rows + rows_per_block - 1) / rows_per_block;
    a.per_xcd = (a.blocks_needed + 7) / 8;
    const int64_t grid = a.per_xcd * 8;
    if constexpr (EPL >= 100) {
        if constexpr (std::is_same<T, float>::value)
            k_apply_v4<G, EPL - 100, NOPS, NF, MODE><<<(unsigned)grid, block, 0, s>>>(a);
    } else {
        k_apply<T, G, EPL, NOPS, NF, MODE><<<(unsigned)grid, block, 0, s>>>(a);
    }
}

template <typename T, int G, int EPL, int MODE>
static bool dispatch_nops_nf(const ApplyArgs<T>& a, int nops, int nf, int block, hipStream_t s) {
    if (nf == 1) {
        switch (nops) {
            case 1: launch_apply<T, G, EPL, 1, 1, MODE>(a, block, s); return true;
            case 2: launch_apply<T, G, EPL, 2, 1, MODE>(a, block, s); return true;
            case 3: launch_apply<T, G, EPL, 3, 1, MODE>(a, block, s); return true;
            case 4: launch_apply<T, G, EPL, 4, 1, MODE>(a, block, s); return true;
        }
        return false;
    }
    if (nops != 1) return false;
    // fields per pass are capped at 32 / EPL (see max_fields) to bound registers
    switch (nf) {
        case 2: if constexpr (2 * (EPL >= 100 ? 4 * (EPL - 100) : EPL) <= 32) { launch_apply<T, G, EPL, 1, 2, MODE>(a, block, s); return true; }
                break;
        case 4: if constexpr (4 * (EPL >= 100 ? 4 * (EPL - 100) : EPL) <= 32) { launch_apply<T, G, EPL, 1, 4, MODE>(a, block, s); return true; }
                break;
        case 8: if constexpr (8 * (EPL >= 100 ? 4 * (EPL - 100) : EPL) <= 32) { launch_apply<T, G, EPL, 1, 8, MODE>(a, block, s); return true; }
                break;
        case 16: if constexpr (16 * (EPL >= 100 ? 4 * (EPL - 100) : EPL) <= 32) { launch_apply<T, G, EPL, 1, 16, MODE>(a, block, s); return true; }
                break;
    }
    return false;
}

// (lanes per row, elements per lane issued up front): every shape covers 64
// nonzeros per pass except the last two, which cover 128. Template EPL >= 100
// selects the 16-byte-load kernel with EPL - 100 chunks of 4 per lane.
template <typename T, int MODE>
static bool dispatch_group(const ApplyArgs<T>& a, int G, int epl, int nops, int nf, int block, hipStream_t s) {
    if (G == 4 && epl == 16) return dispatch_nops_nf<T, 4, 16, MODE>(a, nops, nf, block, s);
    if (G == 8 && epl == 8) return dispatch_nops_nf<T, 8, 8, MODE>(a, nops, nf, block, s);
    if (G == 16 && epl == 4) return dispatch_nops_nf<T, 16, 4, MODE>(a, nops, nf, block, s);
    if (G == 32 && epl == 2) return dispatch_nops_nf<T, 32, 2, MODE>(a, nops, nf, block, s);
    if (G == 64 && epl == 1) return dispatch_nops_nf<T, 64, 1, MODE>(a, nops, nf, block, s);
    if (G == 8 && epl == 16) return dispatch_nops_nf<T, 8, 16, MODE>(a, nops, nf, block, s);
    if (G == 16 && epl == 8) return dispatch_nops_nf<T, 16, 8, MODE>(a, nops, nf, block, s);
    return false;
}

// 16-byte-load shapes (FP32 only): epl = 4 * chunks per lane.
template <int MODE>
static bool dispatch_group_v4(const ApplyArgs<float>& a, int G, int epl, int nops, int nf, int block,
                              hipStream_t s) {
    if (G == 4 && epl == 8) return dispatch_nops_nf<float, 4, 102, MODE>(a, nops, nf, block, s);
    if (G == 8 && epl == 4) return dispatch_nops_nf<float, 8, 101, MODE>(a, nops, nf, block, s);
    if (G == 8 && epl == 8) return dispatch_nops_nf<float, 8, 102, MODE>(a, nops, nf, block, s);
    if (G == 8 && epl == 12) return dispatch_nops_nf<float, 8, 103, MODE>(a, nops, nf, block, s);
    if (G == 16 && epl == 4) return dispatch_nops_nf<float, 16, 101, MODE>(a, nops, nf, block, s);
    if (G == 16 && epl == 8) return dispatch_nops_nf<float, 16, 102, MODE>(a, nops, nf, block, s);
    if (G == 32 && epl == 4) return dispatch_nops_nf<float, 32, 101, MODE>(a, nops, nf, block, s);
    return false;
}

struct ApplyShape {
    int G, epl, vec;
};

// Launch shape from the mean row length (16-byte loads, 8 lanes per row): up
// to 24 nonzeros 4 per lane, otherwise 8 per lane (longer rows finish in the
// strided tail; covering them with a third chunk measured slower). The mean is read once per
// (row_ptr, n_rows) and cached in the context, so repeated applies of one
// operator never synchronise the stream. dcpse_set_apply_config overrides it.
static ApplyShape auto_shape(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr) {
    if (ctx->apply_group > 0) return {ctx->apply_group, ctx->apply_epl, ctx->apply_vec};
    if (ctx->grp_rows == row_ptr && ctx->grp_n == n_rows) return {ctx->grp_g, ctx->grp_epl, ctx->grp_vec};
    int64_t nnz = 0;
    ApplyShape sh{8, 8, 4};
    if (n_rows > 0 &&
        hipMemcpyAsync(&nnz, row_ptr + n_rows, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
        hipStreamSynchronize(ctx->stream) == hipSuccess) {
        double mean = (double)nnz / (double)n_rows;
        if (mean <= 24) sh = {8, 4, 4};
    }
    ctx->grp_rows = row_ptr;
    ctx->grp_n = n_rows;
    ctx->grp_g = sh.G;
    ctx->grp_epl = sh.epl;
    ctx->grp_vec = sh.vec;
    return sh;
}

// Fields per pass: the largest compiled NF <= the remaining count and
// <= 32 / elems_per_lane (every field keeps EPL gathered values live).
static int field_chunk(int remaining, int nops, int epl) {
    if (nops != 1) return 1;
    const int cap = 32 / epl;
    for (int nf = 16; nf >= 2; nf >>= 1)
        if (remaining >= nf && nf <= cap) return nf;
    return 1;
}

template <typename T, int MODE>
static dcpse_status apply_impl(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                               int nops, const T* const* W, const T* f, int64_t ld_f, int nfields,
                               const T* f_self, int64_t ld_self, const T* diag, T* const* out, int64_t ld_out,
                               T alpha, T beta) {
    if (!row_ptr || !col || !W || !f || !out || n_rows < 0 || nfields < 1 || nops < 1)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    if (nops > kMaxOps) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "at most %d operators per pass", kMaxOps);
    if (nops > 1 && nfields > 1)
        return fail(ctx, DCPSE_ERR_UNSUPPORTED, "several operators and several fields in one pass");
    if (nfields > 1 && (ld_f < 1 || ld_out < n_rows || (f_self && ld_self < n_rows)))
        return fail(ctx, DCPSE_ERR_INVALID, "leading dimensions must cover the rows");
    for (int o = 0; o < nops; ++o)
        if (!W[o] || !out[o]) return fail(ctx, DCPSE_ERR_INVALID, "null operator pointer");
    if (n_rows == 0) return DCPSE_OK;
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const ApplyShape sh = auto_shape(ctx, n_rows, row_ptr);
    const int block = ctx->apply_block;
    if (block % sh.G != 0 || block < 64 || block > kApplyMaxBlock)
        return fail(ctx, DCPSE_ERR_INVALID, "block %d incompatible with group %d", block, sh.G);
    ApplyArgs<T> a;
    a.n_rows = n_rows;
    a.row_ptr = row_ptr;
    a.col = col;
    a.diag = diag;
    a.ld_f = ld_f;
    a.ld_self = f_self ? ld_self : ld_f;
    a.ld_out = ld_out;
    a.alpha = alpha;
    a.beta = beta;
    for (int o = 0; o < kMaxOps; ++o) {
        a.W[o] = o < nops ? W[o] : nullptr;
        a.out[o] = nullptr;
    }
    int b0 = 0;
    if constexpr (std::is_same<T, float>::value) {
        // batches of 4 or more fields on one set (columns and rows index the
        // same n_rows particles): interleaved 16 (8, 4) at a time through the
        // workspace (the col/W streams read once per group)
        if (nops == 1 && nfields >= 4 && sh.vec == 4 && il_shape(sh.G, sh.epl) &&
            (!f_self || (f_self == f && ld_self == ld_f))) {
            Workspace ws;
            dcpse_status st = reserve_workspace(ctx, (size_t)n_rows * 16 * sizeof(float) + 256, &ws);
            if (st) return st;
            float* fil = ws.take<float>((size_t)n_rows * 16);
            if (!fil) return fail(ctx, DCPSE_ERR_INVALID, "apply workspace sizing error");
            while (nfields - b0 >= 4) {
                const int nf = nfields - b0 >= 16 ? 16 : (nfields - b0 >= 8 ? 8 : 4);
                const unsigned ib = (unsigned)((n_rows + 255) / 256);
                const float* fb = f + (int64_t)b0 * ld_f;
                if (nf == 16) k_interleave<16><<<ib, 256, 0, ctx->stream>>>(n_rows, fb, ld_f, fil);
                else if (nf == 8) k_interleave<8><<<ib, 256, 0, ctx->stream>>>(n_rows, fb, ld_f, fil);
                else k_interleave<4><<<ib, 256, 0, ctx->stream>>>(n_rows, fb, ld_f, fil);
                DCPSE_LAUNCH_CHECK(ctx);
                a.f = fil;
                a.f_self = fil;
                a.out[0] = out[0] + (int64_t)b0 * ld_out;
                const bool ok = nf == 16 ? dispatch_il<MODE, 16>(a, sh.G, sh.epl, block, ctx->stream)
                                : nf == 8 ? dispatch_il<MODE, 8>(a, sh.G, sh.epl, block, ctx->stream)
                                          : dispatch_il<MODE, 4>(a, sh.G, sh.epl, block, ctx->stream);
                if (!ok) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "no batched apply kernel for group=%d x %d", sh.G,
                                     sh.epl);
                DCPSE_LAUNCH_CHECK(ctx);
                b0 += nf;
            }
        }
    }
    while (b0 < nfields) {
        const int nf = field_chunk(nfields - b0, nops, sh.epl);
        a.f = f + (int64_t)b0 * ld_f;
        a.f_self = (f_self ? f_self + (int64_t)b0 * ld_self : a.f);
        for (int o = 0; o < nops; ++o) a.out[o] = out[o] + (int64_t)b0 * ld_out;
        bool ok = false;
        if constexpr (std::is_same<T, float>::value)
            ok = sh.vec == 4 ? dispatch_group_v4<MODE>(a, sh.G, sh.epl, nops, nf, block, ctx->stream)
                             : dispatch_group<T, MODE>(a, sh.G, sh.epl, nops, nf, block, ctx->stream);
        else
            ok = dispatch_group<T, MODE>(a, sh.G, sh.epl, nops, nf, block, ctx->stream) ||
                 (sh.vec == 4 && dispatch_group<T, MODE>(a, 16, 4, nops, nf, block, ctx->stream));
        if (!ok)
            return fail(ctx, DCPSE_ERR_UNSUPPORTED, "no apply kernel for group=%d x %d ops=%d fields=%d", sh.G,
                        sh.epl, nops, nf);
        DCPSE_LAUNCH_CHECK(ctx);
        b0 += nf;
    }
    return DCPSE_OK;
}

// y = alpha * x + beta * y (beta == 0 never reads y)
template <typename T>
__global__ void k_axpby(int64_t n, T alpha, const T* __restrict__ x, T beta, T* __restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = beta != T(0) ? fma(alpha, x[i], beta * y[i]) : alpha * x[i];
}

template <typename T>
static dcpse_status axpby(dcpse_ctx* ctx, int64_t n, T alpha, const T* x, T beta, T* y) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (n < 0 || (n > 0 && (!x || !y))) return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n == 0) return DCPSE_OK;
    const int64_t b = std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cus * 8);
    k_axpby<T><<<(unsigned)b, 256, 0, ctx->stream>>>(n, alpha, x, beta, y);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

}  // namespace dcpse

using namespace dcpse;

extern "C" {

dcpse_status dcpse_apply_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                             int32_t n_ops, const float* const* W, const float* f, int64_t ld_f, int32_t n_fields,
                             const float* f_self, int64_t ld_self, float* const* out, int64_t ld_out, float alpha,
                             float beta) {
    if (!ctx) return DCPSE_ERR_INVALID;
    return apply_impl<float, kDiff>(ctx, n_rows, row_ptr, col, n_ops, W, f, ld_f, n_fields, f_self, ld_self,
                                    nullptr, out, ld_out, alpha, beta);
}

dcpse_status dcpse_apply_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                             int32_t n_ops, const double* const* W, const double* f, int64_t ld_f,
                             int32_t n_fields, const double* f_self, int64_t ld_self, double* const* out,
                             int64_t ld_out, double alpha, double beta) {
    if (!ctx) return DCPSE_ERR_INVALID;
    return apply_impl<double, kDiff>(ctx, n_rows, row_ptr, col, n_ops, W, f, ld_f, n_fields, f_self, ld_self,
                                     nullptr, out, ld_out, alpha, beta);
}

dcpse_status dcpse_apply_transpose_f32(dcpse_ctx* ctx, int64_t n_cols, const int64_t* rowT_ptr,
                                       const int32_t* colT, const float* WT, const float* diag, const float* g,
                                       int64_t ld_g, int32_t n_fields, float* out, int64_t ld_out, float alpha,
                                       float beta) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!diag) return fail(ctx, DCPSE_ERR_INVALID, "diag is null");
    const float* Wp[1] = {WT};
    float* Op[1] = {out};
    return apply_impl<float, kExplicitDiag>(ctx, n_cols, rowT_ptr, colT, 1, Wp, g, ld_g, n_fields, nullptr, ld_g,
                                            diag, Op, ld_out, alpha, beta);
}

dcpse_status dcpse_apply_transpose_f64(dcpse_ctx* ctx, int64_t n_cols, const int64_t* rowT_ptr,
                                       const int32_t* colT, const double* WT, const double* diag,
                                       const double* g, int64_t ld_g, int32_t n_fields, double* out,
                                       int64_t ld_out, double alpha, double beta) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!diag) return fail(ctx, DCPSE_ERR_INVALID, "diag is null");
    const double* Wp[1] = {WT};
    double* Op[1] = {out};
    return apply_impl<double, kExplicitDiag>(ctx, n_cols, rowT_ptr, colT, 1, Wp, g, ld_g, n_fields, nullptr, ld_g,
                                             diag, Op, ld_out, alpha, beta);
}

dcpse_status dcpse_axpby_f32(dcpse_ctx* ctx, int64_t n, float alpha, const float* x, float beta, float* y) {
    return axpby<float>(ctx, n, alpha, x, beta, y);
}

dcpse_status dcpse_axpby_f64(dcpse_ctx* ctx, int64_t n, double alpha, const double* x, double beta, double* y) {
    return axpby<double>(ctx, n, alpha, x, beta, y);
}

}  // extern "C"
