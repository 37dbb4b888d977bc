// DC-PSE kernel assembly in FP64 for gfx950: one wavefront per particle.
//
// Restates src/DCPSE/Dcpse.hpp:690-809 (Dcpse::initializeStaticSize +
// computeKernel) with an MI355X-first structure:
//   * eps_p = mean_q |x_pq|_1 / h_over_eps       (Vandermonde.hpp:63-81)
//   * the moment matrix A = B^T B with B = E V (Dcpse.hpp:755-757) is built
//     from its distinct entries, the weighted moments
//        M(gamma) = sum_q w_q z_q^gamma,  z = x_pq / eps,  w = exp(-|z|^2) = E_qq^2,
//     so A_ij = M(alpha_i + alpha_j). Neighbours are staged 64 (or 32) at a
//     time as per-dimension power tables in LDS; each lane owns a segment of
//     one moment column.
//   * A is factored by an in-register FP64 Cholesky (lane i holds row i,
//     columns broadcast through LDS; two particles' systems at once for
//     bases of up to 32 monomials) instead of the reference's
//     colPivHouseholderQr (Dcpse.hpp:766); two operators whose bases nest
//     (the Laplacian and Dx, two operators of one basis) share one
//     factorisation; an optional FP64 iterative refinement step recomputes
//     the residual against the moments.
//   * epilogue: lanes own neighbours again and write, for both operators of
//     a pass,
//        K_pq = w_q * sum_i a_i z_q^alpha_i              (Dcpse.hpp:798-809)
//     (Horner form) coalesced, together with the eps^-|beta| pre-scaled
//     FP64/FP32 weights.
// There is no dense n x n E matrix and no per-particle heap traffic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dcpse_common.hpp"
#include "dcpse_internal.hpp"

namespace dcpse {

// One operator of a pass: its right-hand side (summed over its signatures)
// and outputs.
struct AsmOp {
    double b[64];
    unsigned border;  // |beta|
    double* eps;
    double* eps_inv_pow;
    double* K64;
    float* W32;
    double* W64;
    int32_t* info;
};

struct AsmArgs {
    const double* pos_sup;
    int64_t n_sup;
    const double* pos_dom;
    int64_t n_dom;
    const int64_t* row_ptr;
    const int32_t* col;
    double len[3];
    double ilen[3];  // 1 / len
    unsigned periodic;
    double h_over_eps;
    int refine;
    int ppb;              // particles per batch (one wavefront walks a batch)
    const int32_t* order; // batch b holds particles order[b * ppb ...]; nullptr: b * ppb ...
    unsigned* queue;      // persistent launch: counter handing out batches; nullptr: batch = block
    int64_t n_batches;
    AsmOp op[2];  // op[1] used by the two-operator kernels
};

template <int DIM, int L, int AMIN>
__constant__ BasisTables<DIM, L, AMIN> kTab = BasisTables<DIM, L, AMIN>();

// The output fields of operator k, read again from the kernel-argument
// segment where they are needed (a scalar load through a pointer the
// compiler cannot match with the kernel's first read): kept from the start,
// they would hold scalar registers across the whole particle loop. The
// kernel's one argument, the AsmArgs, starts the segment.
struct AsmOut {
    unsigned border;
    double* eps;
    double* eps_inv_pow;
    double* K64;
    float* W32;
    double* W64;
    int32_t* info;
};
__device__ __forceinline__ AsmOut kernarg_out(int k) {
    typedef __attribute__((address_space(4))) const AsmArgs KArgs;
    KArgs* kp = (KArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(kp));
    AsmOut o;
    o.border = kp->op[k].border;
    o.eps = kp->op[k].eps;
    o.eps_inv_pow = kp->op[k].eps_inv_pow;
    o.K64 = kp->op[k].K64;
    o.W32 = kp->op[k].W32;
    o.W64 = kp->op[k].W64;
    o.info = kp->op[k].info;
    return o;
}

// The (y, z) exponent pairs (b, c) with b + c <= P (c = 0 in 2D): the moment
// pass accumulates M(a, b, c) for every a <= P - b - c of each.
template <int DIM, int P>
struct PairTable {
    static constexpr int N = DIM == 3 ? (P + 1) * (P + 2) / 2 : P + 1;
    int bc[N][2];
    constexpr PairTable() : bc{} {
        int c = 0;
        for (int cc = 0; cc <= (DIM == 3 ? P : 0); ++cc)
            for (int b = 0; b + cc <= P; ++b) {
                bc[c][0] = b;
                bc[c][1] = cc;
                ++c;
            }
    }
};

// Lanes per neighbour in the moment pass: the smallest power of two covering
// the pairs.
template <int DIM, int P>
constexpr int moment_lpn() {
    constexpr int n = PairTable<DIM, P>::N;
    return n <= 8 ? 8 : (n <= 16 ? 16 : (n <= 32 ? 32 : 64));
}

// The moment pass's lane segments: pair (b, c) needs the x powers
// a = 0 .. P - b - c; chains longer than T are split in two, [0, T) and
// [T, P - b - c], so that no lane runs more than T steps. T is the smallest
// length for which the segments still fit the lanes of a neighbour (and at
// most two segments per chain). A second segment starts from x^T y^b, staged
// in the power table after the y powers (b <= P - T), so it costs no extra
// arithmetic in the loop. Per segment: the y-factor and wz-factor slots of
// the table row, its exponents and a0, len.
template <int DIM, int P>
struct SegTable {
    static constexpr int LPN = moment_lpn<DIM, P>();
    static constexpr int find_t() {
        int t = P + 1;
        for (int tt = P + 1; tt >= 1; --tt) {
            if (2 * tt < P + 1) break;
            int n = 0;
            for (int c = 0; c <= (DIM == 3 ? P : 0); ++c)
                for (int b = 0; b + c <= P; ++b) n += (P - b - c + 1 > tt) ? 2 : 1;
            if (n <= LPN) t = tt;
        }
        return t;
    }
    static constexpr int T = find_t();
    static constexpr int NX = T < P + 1 ? P - T + 1 : 0;  // x^T y^b slots, b = 0 .. P - T
    static constexpr int count() {
        int n = 0;
        for (int c = 0; c <= (DIM == 3 ? P : 0); ++c)
            for (int b = 0; b + c <= P; ++b) n += (P - b - c + 1 > T) ? 2 : 1;
        return n;
    }
    static constexpr int N = count();
    int yslot[N], zslot[N], b[N], c[N], a0[N], len[N];
    constexpr SegTable() : yslot{}, zslot{}, b{}, c{}, a0{}, len{} {
        int k = 0;
        for (int cc = 0; cc <= (DIM == 3 ? P : 0); ++cc)
            for (int bb = 0; bb + cc <= P; ++bb) {
                const int full = P - bb - cc + 1;
                // row slots: 0 = x, 1 + e = y^e, 1 + (P + 1) + e = x^T y^e,
                // 1 + (P + 1) + NX + e = w z^e
                yslot[k] = 1 + bb;
                zslot[k] = 1 + (P + 1) + NX + cc;
                b[k] = bb;
                c[k] = cc;
                a0[k] = 0;
                len[k] = full > T ? T : full;
                ++k;
                if (full > T) {
                    yslot[k] = 1 + (P + 1) + bb;
                    zslot[k] = 1 + (P + 1) + NX + cc;
                    b[k] = bb;
                    c[k] = cc;
                    a0[k] = T;
                    len[k] = full - T;
                    ++k;
                }
            }
    }
};

template <int DIM, int P>
__constant__ SegTable<DIM, P> kSeg = SegTable<DIM, P>();

__device__ __forceinline__ double ipow_rt(double x, unsigned e) {
    double r = 1.0;
    while (e) {
        if (e & 1u) r *= x;
        x *= x;
        e >>= 1u;
    }
    return r;
}

// The moment grid of the moments M(a, b, c), a + b + c <= P (GW = P + 1):
// M(a, b, c) sits at a + GW b + S2 c with S2 = P GW + 1. The index is linear
// in the exponents, so A_ik = M(alpha_i + alpha_k) is at index(alpha_i) +
// index(alpha_k), and injective on the simplex a + b + c <= P (slice c spans
// at most (P - c) GW < S2): P S2 + 1 entries (259 for P = 6) instead of GW^3.
__host__ __device__ constexpr int grid_s2(int GW) { return (GW - 1) * GW + 1; }
__host__ __device__ constexpr int grid_size(int DIM, int GW) {
    return DIM == 3 ? (GW - 1) * grid_s2(GW) + 1 : (GW - 1) * GW + 1;
}
template <int GW>
__device__ __forceinline__ constexpr int grid_index(int a, int b, int c) { return a + GW * b + grid_s2(GW) * c; }

// Lane i's offset of A row i in the moment grid: the moment index of
// monomial alpha_i (lanes past the basis take row 0).
template <int DIM, int L, int AMIN, int GW>
__device__ __forceinline__ int mom_row_base(int lane) {
    constexpr int M = BasisShape<DIM, L, AMIN>::M;
    const int li = lane < M ? lane : 0;
    const int e2 = DIM == 3 ? kTab<DIM, L, AMIN>.e[li][2] : 0;
    return grid_index<GW>(kTab<DIM, L, AMIN>.e[li][0], kTab<DIM, L, AMIN>.e[li][1], e2);
}

template <int LL, int AA>
struct BasisTag {
    static constexpr int L = LL, A = AA;
};

// Writes the "no kernel" result of one operator for particle p.
template <class Op>
__device__ __forceinline__ void asm_fail(const Op& op, int64_t p, int64_t rs, int n, int status, int lane) {
    for (int j = lane; j < n; j += kWave) {
        if (op.K64) op.K64[rs + j] = 0.0;
        if (op.W64) op.W64[rs + j] = 0.0;
        if (op.W32) op.W32[rs + j] = 0.0f;
    }
    if (lane == 0) {
        if (op.eps) op.eps[p] = 0.0;
        if (op.eps_inv_pow) op.eps_inv_pow[p] = 0.0;
        if (op.info) op.info[p] = status;
    }
}

// 1/sqrt(d) of a Cholesky pivot: the hardware estimate r (~23 bits) and one
// third-order step r (1 + e/2 + 3e^2/8), e = 1 - d r^2, whose error O(e^3)
// is below the final rounding; one estimate and four dependent FP64
// operations instead of a square root and a division. A pivot that is not
// positive and finite gives NaN (negative or NaN: the estimate is NaN; zero
// or +inf: d * r is 0 * inf).
__device__ __forceinline__ double pivot_rsqrt(double d) {
    const double r = __builtin_amdgcn_rsq(d);
    const double e = fma(-d * r, r, 1.0);
    return fma(r * e, fma(e, 0.375, 0.5), r);
}

// LDS doubles between the L factors of two slots: rows x
// even row stride, padded so that slot s starts 8 s banks on (the slots' lanes
// read the same row and column at once).
__host__ __device__ constexpr int lm_slot_stride(int rows) {
    int s = rows * (rows + (rows & 1));
    while ((2 * s) % 32 != 8) ++s;
    return s;
}

// Moment-system factorisation and solve of one operator for the particles of
// SLOTS lane groups at once (LS = 64 / SLOTS lanes each, one particle per
// group): lane slot * LS + i holds row i of its particle's A = (M(alpha_i +
// alpha_k))_k, read from the moment grid (grid_index) at abase + offset(k).
// A = L L^T by a right-looking Cholesky; forward and backward substitution;
// optionally one step of FP64 iterative refinement against the moments. The
// solution entry x_i is left in sx[lane]; okl is false when a pivot of the
// lane's system is not positive and finite. The groups share every
// instruction and read only their own LDS entries, so a failed or absent
// particle in one group cannot reach another.
template <int DIM, int L, int AMIN, int GW, int SLOTS>
__device__ __forceinline__ void asm_factor_solve(const double* mom, int abase, double b_lane, int refine,
                                                 double* lmat, double2* lbuf, double* sx, int lane, bool& okl) {
    constexpr int M = BasisShape<DIM, L, AMIN>::M;
    constexpr int LS = kWave / SLOTS;
    static_assert(M <= LS, "basis larger than a lane group");
    constexpr BasisTables<DIM, L, AMIN> E{};
    const int slot = min(lane / LS, SLOTS - 1);  // three groups: lane 63 is row 21 of the last
    const int base = slot * LS;
    // row index, opaque so the lane == k tests are not hoisted out of the
    // particle loop as spilled masks
    int me = lane - base;
    asm volatile("" : "+v"(me));
    const int li = me < M ? me : 0;

    double ar[M];
#pragma unroll
    for (int k = 0; k < M; ++k) ar[k] = mom[abase + grid_index<GW>(E.e[k][0], E.e[k][1], E.e[k][2])];

    // ---- A = L L^T, right-looking; lane i holds row i ---------------------
    // One group: step k takes the pivot from lane k (v_readlane), scales
    // column k and publishes it in LDS; every lane reads the trailing entries
    // l_jk as broadcasts and updates its row. Two groups: step k publishes
    // the unscaled column (with the lane's right-hand side entry), each lane
    // reads its group's pivot and entries from it, and updates with
    // (a_ik / d_kk) a_jk = l_ik l_jk. Entries right of
    // the diagonal are updated too and never read as part of L. The pivots
    // are checked once after the loop, from each lane's own 1 / L_kk (no
    // per-step masks); a failed factorisation writes zero weights whatever
    // it computed.
    double dinv = 0.0;  // lane k: 1 / L_kk
    // two groups: the right-hand side rides along as column M of A, so the
    // factorisation also performs the forward substitution L y = b (lane k
    // keeps y_k)
    double rb = b_lane, yb = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        // alternate buffers between steps
        double* col = reinterpret_cast<double*>(lbuf) + (k & 1) * kWave;
        double2* col2 = lbuf + (k & 1) * kWave;
        double dkk, rk = 0.0;
        double cv[M];  // two groups: the unscaled column k of the lane's group
        if constexpr (SLOTS == 1) {
            dkk = readlane_d(ar[k], k);
        } else {
            col2[lane] = make_double2(ar[k], rb);
            wave_lds_sync();
            const double2 pk = col2[base + k];
            dkk = pk.x;
            rk = pk.y;
#pragma unroll
            for (int j = k + 1; j < M; ++j) cv[j] = col2[base + j].x;
        }
        const double inv = pivot_rsqrt(dkk);
        dinv = (me == k) ? inv : dinv;
        // lane k's own entry becomes dkk * inv = L_kk
        ar[k] *= inv;
        if constexpr (SLOTS == 1) {
            const double lik = ar[k];
            if (k + 1 < M) {
                col[lane] = lik;
                wave_lds_sync();
#pragma unroll
                for (int j = k + 1; j < M; ++j) ar[j] = fma(-lik, col[j], ar[j]);
            }
        } else {
            const double t = ar[k] * inv;
            const double yk = rk * inv;
            yb = (me == k) ? yk : yb;
            rb = fma(-ar[k], yk, rb);
#pragma unroll
            for (int j = k + 1; j < M; ++j) ar[j] = fma(-t, cv[j], ar[j]);
        }
        // the selects stay in their step (not sunk to the end of the loop
        // with every step's values live)
        asm volatile("" : "+v"(dinv), "+v"(yb));
        __builtin_amdgcn_sched_barrier(0);
    }
    // lane k holds 1 / sqrt(pivot k), finite exactly when the pivot is
    // positive and finite; fma(v, 0, 0) is 0 exactly when v is finite (an
    // explicit fma: v - v may be contracted with the product that forms v)
    {
        const uint64_t bad = __ballot(me < M && fma(dinv, 0.0, 0.0) != 0.0);
        const uint64_t grp = LS == kWave ? ~0ull : ((1ull << LS) - 1ull) << base;
        okl = (bad & grp) == 0ull;
    }
    // keep the strictly lower part of L (the diagonal is in dinv): a lane's
    // partial sums then stop changing once its unknown is determined, so the
    // substitutions need no per-lane capture of their results
#pragma unroll
    for (int k = 0; k < M; ++k) ar[k] = k < me ? ar[k] : 0.0;
    // L rows to LDS (row stride MS, one block per group): the backward solve
    // reads column `me`
    constexpr int MS = M + (M & 1);
    double* lm = lmat + slot * lm_slot_stride(M);
    if (me < M) {
#pragma unroll
        for (int k = 0; k < M; ++k) lm[me * MS + k] = ar[k];
    }
    wave_lds_sync();

    // unknown k of each group to all lanes of the group (v_readlane per group)
    auto bcast = [&](double v, int k) -> double {
        if constexpr (SLOTS == 1) {
            return readlane_d(v, k);
        } else if constexpr (SLOTS == 2) {
            const double y0 = readlane_d(v, k), y1 = readlane_d(v, LS + k);
            return slot ? y1 : y0;
        } else {
            const double y0 = readlane_d(v, k), y1 = readlane_d(v, LS + k), y2 = readlane_d(v, 2 * LS + k);
            return slot == 2 ? y2 : (slot ? y1 : y0);
        }
    };
    // forward (L y = r) and backward (L^T x = y) substitution: step k
    // broadcasts unknown k and every lane subtracts its multiple; lane k's own
    // sum is final from step k on, so it ends holding y_k, then x_k
    auto back_solve = [&](double t) {
#pragma unroll
        for (int j = M - 1; j >= 0; --j) t = fma(-lm[j * MS + li], bcast(t * dinv, j), t);
        return t * dinv;
    };
    auto chol_solve = [&](double rhs) {
        double r = rhs;
#pragma unroll
        for (int k = 0; k < M; ++k) r = fma(-ar[k], bcast(r * dinv, k), r);
        return back_solve(r * dinv);
    };
    double xv = SLOTS == 1 ? chol_solve(b_lane) : back_solve(yb);
    if (refine) {
        // one step of iterative refinement against the moments: x += A^-1 (b - A x)
        sx[lane] = xv;
        wave_lds_sync();
        double rr = b_lane;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const double aik = mom[abase + grid_index<GW>(E.e[k][0], E.e[k][1], E.e[k][2])];
            rr = fma(-aik, sx[base + k], rr);
        }
        xv += chol_solve(rr);
    }
    sx[lane] = xv;
    wave_lds_sync();
}

// Two operators on one support whose bases nest: operator 1's basis is
// operator 0's monomials that it shares (S of them) plus at most the constant
// monomial (E = 1), e.g. Dx (degrees 0-2) beside the Laplacian (degrees 1-3),
// or two operators of one basis (S = M0, E = 0). Operator 1's moment matrix is
// then, in the row order [shared, constant], the leading S x S block of
// operator 0's (rows ordered [shared, the rest]) bordered by the constant's
// row, and its Cholesky factor is the first S columns of operator 0's factor
// plus that row eliminated by the same S pivots: one factorisation of
// operator 0's system with the constant's row appended as row M0 serves both.
// p1[i], q2[i]: the operator-0 and operator-1 basis index of solve row i (-1
// when the row is not in that basis); ex[i]: its exponents.
template <int DIM, int L1, int A1, int L2, int A2>
struct FusePlan {
    static constexpr int M1 = BasisShape<DIM, L1, A1>::M;
    static constexpr int M2 = BasisShape<DIM, (L2 > 0 ? L2 : 2), A2>::M;
    bool ok;
    int S, E;
    int p1[M1 + 1], q2[M1 + 1], ex[M1 + 1][3];
    static constexpr int find(const int (*e)[3], int m, const int* x) {
        for (int k = 0; k < m; ++k)
            if (e[k][0] == x[0] && e[k][1] == x[1] && e[k][2] == x[2]) return k;
        return -1;
    }
    constexpr FusePlan() : ok(false), S(0), E(0), p1{}, q2{}, ex{} {
        const BasisTables<DIM, L1, A1> B1{};
        const BasisTables<DIM, (L2 > 0 ? L2 : 2), A2> B2{};
        int s = 0;
        for (int pass = 0; pass < 2; ++pass)
            for (int i = 0; i < M1; ++i)
                if ((find(B2.e, M2, B1.e[i]) >= 0) == (pass == 0)) p1[s++] = i;
        for (int i = 0; i < M1; ++i) {
            if (find(B2.e, M2, B1.e[i]) >= 0) ++S;
            for (int d = 0; d < 3; ++d) ex[i][d] = B1.e[p1[i]][d];
            q2[i] = find(B2.e, M2, ex[i]);
        }
        int extra = 0, idx = -1;
        bool konst = true;
        for (int k = 0; k < M2; ++k)
            if (find(B1.e, M1, B2.e[k]) < 0) {
                ++extra;
                idx = k;
                konst = konst && B2.e[k][0] == 0 && B2.e[k][1] == 0 && B2.e[k][2] == 0;
            }
        E = extra;
        ok = L2 > 0 && S > 0 && (extra == 0 || (extra == 1 && konst));
        p1[M1] = -1;
        q2[M1] = extra == 1 ? idx : -1;
        ex[M1][0] = ex[M1][1] = ex[M1][2] = 0;
    }
};
template <int DIM, int L1, int A1, int L2, int A2>
__constant__ FusePlan<DIM, L1, A1, L2, A2> kFuse = FusePlan<DIM, L1, A1, L2, A2>();

// The factorisation and both solves of a nested pair (FusePlan) for NS
// particles at once, one per lane group of LS = 64 / NS lanes (the last group
// also takes the lanes past NS * LS, as rows beyond the system): lane
// slot * LS + i holds solve row i
// (M0 + E rows; abase: its moment row in its slot's grid; b1v, b2v: its
// right-hand side entries of both operators). Right-looking Cholesky over the
// M0 pivots of operator 0 as in asm_factor_solve, with both right-hand sides
// riding along (operator 1's over the S shared pivots); the constant row's
// Schur complement after the S shared pivots is operator 1's last pivot. The
// backward substitutions share the L rows of the shared pivots; the optional
// refinement step (x += A^-1 (b - A x) against the moments) does the same
// for both operators. Solutions go to sx (operator 0) and sx2 (operator 1)
// in basis order; ok1 / ok2 are false when a pivot of that operator's system
// is not positive and finite.
template <int DIM, int L1, int A1, int L2, int A2, int GW, int NS>
__device__ __forceinline__ void asm_factor_solve_fused(const double* mom, int abase, double b1v, double b2v,
                                                       int p1, int q2, int refine, double* lmat, double2* lbuf,
                                                       double* sx, double* sx2, int lane, bool& ok1, bool& ok2) {
    using FP = FusePlan<DIM, L1, A1, L2, A2>;
    constexpr FP F{};
    constexpr int M1 = FP::M1, S = F.S, E = F.E, MA = M1 + E;
    constexpr int LS = kWave / NS;
    static_assert(NS == 2 || NS == 3, "two or three particles per wave");
    static_assert(MA <= LS, "nested pair larger than a lane group");
    const int slot = min(lane / LS, NS - 1);
    const int base = slot * LS;
    int me = lane - base;
    asm volatile("" : "+v"(me));
    const int li = me < MA ? me : 0;

    double ar[MA];
#pragma unroll
    for (int k = 0; k < MA; ++k) ar[k] = mom[abase + grid_index<GW>(F.ex[k][0], F.ex[k][1], F.ex[k][2])];
    double dinv = 0.0, dc = 0.0;
    double rb = b1v, yb = 0.0, rb2 = b2v, yb2 = 0.0;
#pragma unroll
    for (int k = 0; k < M1; ++k) {
        // alternate buffers between steps; operator 1's right-hand side
        // column goes through the solution buffers, free until the end
        double2* col2 = lbuf + (k & 1) * kWave;
        double* colb = (k & 1) ? sx : sx2;
        col2[lane] = make_double2(ar[k], rb);
        if (k < S) colb[lane] = rb2;
        wave_lds_sync();
        const double2 pk = col2[base + k];
        const double dkk = pk.x, rk = pk.y;
        const double rk2 = k < S ? colb[base + k] : 0.0;
        double cv[MA];
#pragma unroll
        for (int j = k + 1; j < MA; ++j) cv[j] = col2[base + j].x;
        const double inv = pivot_rsqrt(dkk);
        dinv = (me == k) ? inv : dinv;
        ar[k] *= inv;
        const double t = ar[k] * inv;
        const double yk = rk * inv;
        yb = (me == k) ? yk : yb;
        rb = fma(-ar[k], yk, rb);
        if (k < S) {
            const double yk2 = rk2 * inv;
            yb2 = (me == k) ? yk2 : yb2;
            rb2 = fma(-ar[k], yk2, rb2);
        }
        // the selects stay in their step: sunk to the end of the loop they
        // would keep every step's inverse and unknowns live (245 instead of
        // 200 VGPRs)
        asm volatile("" : "+v"(dinv), "+v"(yb), "+v"(yb2));
#pragma unroll
        for (int j = k + 1; j < MA; ++j) ar[j] = fma(-t, cv[j], ar[j]);
        if (E && k == S - 1) dc = ar[MA - 1];  // the constant row's pivot (its lane only)
        __builtin_amdgcn_sched_barrier(0);
    }
    // unknown k of each group to all lanes of the group (v_readlane per group)
    auto bcast = [&](double v, int k) -> double {
        const double y0 = readlane_d(v, k), y1 = readlane_d(v, LS + k);
        if constexpr (NS == 2) {
            return slot ? y1 : y0;
        } else {
            const double y2 = readlane_d(v, 2 * LS + k);
            return slot == 2 ? y2 : (slot ? y1 : y0);
        }
    };
    double dinv2 = dinv;
    if constexpr (E > 0) {
        const double ic = pivot_rsqrt(bcast(dc, M1));
        dinv2 = (me == M1) ? ic : dinv;
        yb2 = (me == M1) ? rb2 * ic : yb2;
    }
    // pivot checks: lane k holds step k's 1 / sqrt(pivot) (and lane M0 the
    // constant's), which is finite exactly when the pivot is positive and
    // finite; fma(v, 0, 0) is 0 exactly when v is finite
    const uint64_t bad1 = __ballot(me < M1 && fma(dinv, 0.0, 0.0) != 0.0);
    const uint64_t bad2 = __ballot((me < S || (E > 0 && me == M1)) && fma(dinv2, 0.0, 0.0) != 0.0);
    const uint64_t grp = ((1ull << LS) - 1ull) << base;
    ok1 = (bad1 & grp) == 0ull;
    ok2 = (bad2 & grp) == 0ull;
    // strictly lower L rows to LDS (row stride MS): the substitutions read
    // column `me`
#pragma unroll
    for (int k = 0; k < MA; ++k) ar[k] = k < me ? ar[k] : 0.0;
    constexpr int MS = MA + (MA & 1);
    double* lm = lmat + slot * lm_slot_stride(MA);
    if (me < MA) {
#pragma unroll
        for (int k = 0; k < MA; ++k) lm[me * MS + k] = ar[k];
    }
    wave_lds_sync();
    // backward substitutions L^T x = y: operator 0 over its M0 unknowns,
    // operator 1 over the constant's (row M0) then the shared ones, which
    // share their L entries with operator 0's; lane i's sums are final from
    // its own step on
    auto back = [&](double& t1, double& t2) {
#pragma unroll
        for (int j = M1 - 1; j >= S; --j) t1 = fma(-lm[j * MS + li], bcast(t1 * dinv, j), t1);
        if constexpr (E > 0) t2 = fma(-lm[M1 * MS + li], bcast(t2 * dinv2, M1), t2);
#pragma unroll
        for (int j = S - 1; j >= 0; --j) {
            const double l = lm[j * MS + li];
            t1 = fma(-l, bcast(t1 * dinv, j), t1);
            t2 = fma(-l, bcast(t2 * dinv2, j), t2);
        }
        t1 *= dinv;
        t2 *= dinv2;
    };
    double x1 = yb, x2 = yb2;
    back(x1, x2);
    if (refine) {
        // residuals against the moments in solve order (the unknowns through
        // LDS, one per lane), then L y = r forward (row i of L in ar) and back
        sx[lane] = x1;
        sx2[lane] = x2;
        wave_lds_sync();
        double r1 = b1v, r2 = b2v;
#pragma unroll
        for (int k = 0; k < MA; ++k) {
            const double aik = mom[abase + grid_index<GW>(F.ex[k][0], F.ex[k][1], F.ex[k][2])];
            if (k < M1) r1 = fma(-aik, sx[base + k], r1);
            if (k < S || k == M1) r2 = fma(-aik, sx2[base + k], r2);
        }
#pragma unroll
        for (int k = 0; k < M1; ++k) r1 = fma(-ar[k], bcast(r1 * dinv, k), r1);
#pragma unroll
        for (int k = 0; k < S; ++k) r2 = fma(-ar[k], bcast(r2 * dinv2, k), r2);
        r1 *= dinv;
        r2 *= dinv2;
        back(r1, r2);
        x1 += r1;
        x2 += r2;
        wave_lds_sync();
    }
    if (p1 >= 0) sx[base + p1] = x1;
    if (q2 >= 0) sx2[base + q2] = x2;
    wave_lds_sync();
}

// Horner scheme of a basis polynomial sum_i x_i z^alpha_i: the monomials
// grouped by their (y, z) exponents (b, c); each group's x exponents are
// consecutive, a_min .. a_max, so the group is
//   y^b z^c x^a_min (x_max + x (... )) evaluated innermost first.
// idx lists the monomial indices of each group from the highest a down.
template <int DIM, int L, int AMIN>
struct HornerPlan {
    static constexpr int M = BasisShape<DIM, L, AMIN>::M;
    int ng;
    int b[M], c[M], cnt[M], amin[M], start[M];
    int idx[M];
    constexpr HornerPlan() : ng(0), b{}, c{}, cnt{}, amin{}, start{}, idx{} {
        const BasisTables<DIM, L, AMIN> E{};
        int pos = 0;
        for (int cc = 0; cc < (DIM == 3 ? L : 1); ++cc)
            for (int bb = 0; bb + cc < L; ++bb) {
                const int first = pos;
                int lo = 0, count = 0;
                for (int aa = L - 1; aa >= 0; --aa)
                    for (int k = 0; k < M; ++k)
                        if (E.e[k][0] == aa && E.e[k][1] == bb && E.e[k][2] == cc) {
                            idx[pos++] = k;
                            lo = aa;
                            ++count;
                        }
                if (count) {
                    b[ng] = bb;
                    c[ng] = cc;
                    cnt[ng] = count;
                    amin[ng] = lo;
                    start[ng] = first;
                    ++ng;
                }
            }
    }
};

// sum_i xs[i] z^alpha_i for the basis <DIM, L, AMIN>, from the powers
// px[k] = x^k, py[k] = y^k, pz[k] = z^k (k < L).
template <int DIM, int L, int AMIN>
__device__ __forceinline__ double basis_sum(const double* xs, const double* px, const double* py, const double* pz) {
    constexpr HornerPlan<DIM, L, AMIN> H{};
    double sum = 0.0;
#pragma unroll
    for (int g = 0; g < H.ng; ++g) {
        double h = xs[H.idx[H.start[g]]];
#pragma unroll
        for (int t = 1; t < H.cnt[g]; ++t) h = fma(h, px[1], xs[H.idx[H.start[g] + t]]);
        if (H.amin[g] > 0) h *= px[H.amin[g]];
        if (H.b[g] > 0 && H.c[g] > 0) h *= py[H.b[g]];
        if (H.c[g] > 0) sum = fma(h, pz[H.c[g]], sum);
        else if (H.b[g] > 0) sum = fma(h, py[H.b[g]], sum);
        else sum += h;
    }
    return sum;
}

// Kernel weights of the pass's operators for particle p (Dcpse.hpp:798-809):
// lanes own neighbours, K_pq = w_q * sum_i a_i z_q^alpha_i with each
// operator's solution a in xs1 / xs2 (wave-uniform), written with the
// eps^-|beta| pre-scaled copies. The displacement, weight w and powers of a
// neighbour are formed once for both operators; those of the first 64
// neighbours are the moment pass's: with DC from dsl[d * 64 + lane] and
// wsl[lane], otherwise the displacement from rdv (the lane's own neighbour,
// kept in registers) and the weight from rw when WR (else formed again by
// the moment pass's operations); later neighbours' are formed again. Operator
// k is assembled when use_k (enough neighbours;
// otherwise it is marked DCPSE_INFO_TOO_FEW with eps = 0) and solved when
// ok_k (otherwise DCPSE_INFO_NOT_SPD and zero weights).
template <int DIM, int L1, int A1, int L2, int A2, bool DC, bool WR, typename Disp>
__device__ __forceinline__ void asm_weights(const AsmArgs& a, int64_t p, int64_t rs, int n, double eps,
                                            double ieps, bool use1, bool ok1, bool use2, bool ok2,
                                            const double* xs1, const double* xs2, const double* dsl,
                                            const double* wsl, const double* rdv, double rw, Disp disp,
                                            int lane) {
    constexpr int LP = L1 > L2 ? L1 : L2;
    const AsmOut o1 = kernarg_out(0);
    const AsmOut o2 = L2 > 0 ? kernarg_out(1) : o1;
    const double eip1 = ipow_rt(ieps, o1.border);
    const double eip2 = L2 > 0 ? ipow_rt(ieps, o2.border) : 0.0;
    bool bad1 = false, bad2 = false;
    for (int j = lane; j < n; j += kWave) {
        double dv[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < DIM; ++d) dv[d] = DC ? dsl[d * kWave + lane] : rdv[d];
        if (j >= kWave) disp(j, dv);
        double z[3] = {0.0, 0.0, 0.0};
        double z2 = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            z[d] = dv[d] * ieps;
            z2 = fma(z[d], z[d], z2);  // as the moment pass forms it
        }
        const double w = j < kWave && (DC || WR) ? (DC ? wsl[j] : rw) : exp(-z2);
        double px[LP], py[LP], pz[LP];
        px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
        for (int k = 1; k < LP; ++k) {
            px[k] = px[k - 1] * z[0];
            py[k] = py[k - 1] * z[1];
            pz[k] = pz[k - 1] * z[2];
        }
        {
            const double K = ok1 ? basis_sum<DIM, L1, A1>(xs1, px, py, pz) * w : 0.0;
            bad1 |= !isfinite(K);
            if (o1.K64) o1.K64[rs + j] = K;
            if (o1.W64) o1.W64[rs + j] = K * eip1;
            if (o1.W32) o1.W32[rs + j] = (float)(K * eip1);
        }
        if constexpr (L2 > 0) {
            const double K = ok2 ? basis_sum<DIM, L2, A2>(xs2, px, py, pz) * w : 0.0;
            bad2 |= !isfinite(K);
            if (o2.K64) o2.K64[rs + j] = K;
            if (o2.W64) o2.W64[rs + j] = K * eip2;
            if (o2.W32) o2.W32[rs + j] = (float)(K * eip2);
        }
    }
    const bool anybad1 = __ballot(bad1) != 0ull;
    const bool anybad2 = __ballot(bad2) != 0ull;
    if (lane == 0) {
        auto mark = [&](const AsmOut& op, bool use, bool ok, bool anybad, double eip) {
            if (op.eps) op.eps[p] = use ? eps : 0.0;
            if (op.eps_inv_pow) op.eps_inv_pow[p] = use ? eip : 0.0;
            if (op.info)
                op.info[p] = !use  ? DCPSE_INFO_TOO_FEW
                             : !ok ? DCPSE_INFO_NOT_SPD
                                   : (anybad ? DCPSE_INFO_NONFINITE : DCPSE_INFO_OK);
        };
        mark(o1, use1, ok1, anybad1, eip1);
        if constexpr (L2 > 0) mark(o2, use2, ok2, anybad2, eip2);
    }
}

// Waves per SIMD each instantiation is compiled for: the largest bases need
// the whole register file of one wave; one operator of up to 10 monomials
// fits three waves.
template <int DIM, int L1, int A1, int L2, int A2>
constexpr int asm_waves() {
    constexpr int m1 = BasisShape<DIM, L1, A1>::M;
    constexpr int m2 = L2 > 0 ? BasisShape<DIM, (L2 > 0 ? L2 : 2), A2>::M : 0;
    constexpr int m = m1 > m2 ? m1 : m2;
    return m > 24 ? 1 : (m > 10 || L2 > 0 ? 2 : 3);
}

// One wavefront walks a.ppb particles (at most 63). Operator 0 has basis
// <DIM, L1, A1>; with L2 > 0 a second operator of basis <DIM, L2, A2> on the
// same support is assembled in the same pass, sharing ε, the neighbour
// displacements and the moments (the moment set covers both). ε and the
// moments are computed one particle at a time by the whole wave; when the
// bases fit in half a wave (m <= 32), two particles' systems are factored and
// solved together, one per half, and a nested pair (FusePlan) factors one
// system for both operators.
template <int DIM, int L1, int A1, int L2 = 0, int A2 = 0, int CH = 32, bool PQ = false>
__global__ __launch_bounds__(64, (asm_waves<DIM, L1, A1, L2, A2>())) void k_assemble(const AsmArgs a) {
    constexpr int P1 = 2 * (L1 - 1), P2 = L2 > 0 ? 2 * (L2 - 1) : 0;
    constexpr int P = P1 > P2 ? P1 : P2;
    constexpr int QMIN = L2 > 0 ? (A1 < A2 ? 2 * A1 : 2 * A2) : 2 * A1;
    constexpr int GW = P + 1, MG = grid_size(DIM, GW);  // moment grid (grid_index)
    constexpr int MGP = MG + (MG & 1);
    constexpr int M1 = BasisShape<DIM, L1, A1>::M;
    constexpr int M2 = L2 > 0 ? BasisShape<DIM, (L2 > 0 ? L2 : 2), A2>::M : 0;
    constexpr int MX = M1 > M2 ? M1 : M2;
    constexpr int MN = L2 > 0 ? (M1 < M2 ? M1 : M2) : M1;
    // nested pairs (FusePlan) share one factorisation of M1 + E rows; three
    // particles are factored together when that system (or a single
    // operator's) has 16 to 21 rows (the factorisation then outweighs the
    // epilogue inputs the three-slot kernels give up keeping in LDS:
    // Laplacian + Dx in 3D 5.77 -> 5.65 ms even with those recomputed, while
    // Dx + Dy, 9 rows, went from 4.14 to 4.41 ms), two when a basis fits half
    // a wave
    using FP = FusePlan<DIM, L1, A1, L2, A2>;
    constexpr FP FPL{};
    constexpr bool FUSE = L2 > 0 && FPL.ok && M1 + FPL.E <= kWave / 2;
    constexpr int MF = M1 + FPL.E;
    constexpr bool S3 = FUSE ? (MF >= 16 && MF <= kWave / 3) : (L2 == 0 && M1 >= 16 && M1 <= kWave / 3);
    constexpr int SLOTS = S3 ? 3 : (MX <= kWave / 2 ? 2 : 1);  // particles factored together
    constexpr int LS = kWave / SLOTS;
    // the first 64 displacements and weights of each slot's particle are kept
    // in LDS for the epilogue, except with three slots (three copies would not
    // fit the LDS of two blocks per SIMD): there each lane keeps its own
    // neighbour's in registers (the weight with chunks of 64; with chunks of
    // 32 another lane formed it, so the epilogue forms it again by the same
    // operations)
    constexpr bool DCACHE = SLOTS < 3;
    constexpr int PW = P + 1;
    using Seg = SegTable<DIM, P>;
    constexpr int ST = Seg::T;    // x powers per lane segment
    constexpr int NX = Seg::NX;   // x^T y^b slots
    // power-table row per staged neighbour: [x][y^0..y^P][x^T y^0..x^T y^(P-T)]
    // [w z^0..w z^P] (2D: [x][y^0..y^P][x^T y^b..][w])
    constexpr int ZW = DIM == 3 ? PW : 1;
    constexpr int TROW = 1 + PW + NX + ZW;
    static_assert(CH == 32 || CH == 64, "neighbour chunk of 32 or 64");
    constexpr int LT = FUSE && lm_slot_stride(MF) > lm_slot_stride(MX) ? SLOTS * lm_slot_stride(MF)
                                                                        : SLOTS * lm_slot_stride(MX);
    constexpr int TAB = CH * TROW > LT ? CH * TROW : LT;
    // lanes per neighbour (the segments of the moment pass fit them); a pass
    // covers HALVES = 64 / LPN neighbours
    constexpr int NSEG = Seg::N;
    constexpr int LPN = Seg::LPN;
    constexpr int HALVES = kWave / LPN;
    static_assert(CH % (2 * HALVES) == 0, "a chunk holds whole passes");
    static_assert(NSEG <= LPN, "moment segments exceed the lanes of a neighbour");
    __shared__ __align__(16) double tab[TAB];     // staged powers; L factors in the solve
    __shared__ double mom[SLOTS * MGP];           // moments of each slot's particle
    __shared__ double2 lbuf[2 * kWave];           // broadcast buffers of the factorisation
    __shared__ double sx[kWave];                  // solutions of operator 0, one entry per lane
    __shared__ double sx2[L2 > 0 ? kWave : 1];    // and of operator 1
    __shared__ double dsl[DCACHE ? SLOTS * 3 * kWave : 1];  // first 64 displacements per slot
    __shared__ double wsl[DCACHE ? SLOTS * kWave : 1];      // and their weights exp(-|z|^2)
    const int lane = threadIdx.x;

    // the moment segment this lane accumulates, and which neighbour of a pass
    const int half = lane / LPN;
    const int pl = lane % LPN;
    const bool pv = pl < NSEG;
    const int sgi = pv ? pl : 0;
    const int pys = kSeg<DIM, P>.yslot[sgi], pzs = kSeg<DIM, P>.zslot[sgi];
    const int pb = kSeg<DIM, P>.b[sgi], pc = kSeg<DIM, P>.c[sgi];
    const int pa0 = kSeg<DIM, P>.a0[sgi], plen = kSeg<DIM, P>.len[sgi];

    // lane slot * LS + i: row i of each operator's system (right-hand side
    // entry and offset of A row i in its slot's moment grid); with three
    // slots lane 63 is row 21 of the last
    const int lslot = min(lane / LS, SLOTS - 1), lrow = lane - lslot * LS;
    const double b1 = lrow < M1 ? a.op[0].b[lrow] : 0.0;
    const double b2 = L2 > 0 && lrow < M2 ? a.op[1].b[lrow] : 0.0;
    const int ab1 = lslot * MGP + mom_row_base<DIM, L1, A1, GW>(lrow);
    const int ab2 = L2 > 0 ? lslot * MGP + mom_row_base<DIM, (L2 > 0 ? L2 : 2), A2, GW>(lrow) : 0;
    // nested pair: solve row lrow's basis indices, right-hand sides and moment row
    int fp1 = -1, fq2 = -1, fab = 0;
    double fb1 = 0.0, fb2 = 0.0;
    if constexpr (FUSE) {
        if (lrow < MF) {
            fp1 = kFuse<DIM, L1, A1, L2, A2>.p1[lrow];
            fq2 = kFuse<DIM, L1, A1, L2, A2>.q2[lrow];
            const int e2 = DIM == 3 ? kFuse<DIM, L1, A1, L2, A2>.ex[lrow][2] : 0;
            fab = grid_index<GW>(kFuse<DIM, L1, A1, L2, A2>.ex[lrow][0], kFuse<DIM, L1, A1, L2, A2>.ex[lrow][1], e2);
        }
        fb1 = fp1 >= 0 ? a.op[0].b[fp1] : 0.0;
        fb2 = fq2 >= 0 ? a.op[1].b[fq2] : 0.0;
        fab += lslot * MGP;
    }

    // Batches of a.ppb particles. Static grid (PQ false): batch = block, its
    // particles b * ppb + i. Persistent grid (PQ true): batches taken from a
    // counter until none is left (every wave leaves the loop once the counter
    // passes n_batches), particles order[b * ppb + i] when bucketed by row
    // length. The static kernel runs the loop body once.
    for (int64_t batch = blockIdx.x;;) {
    if constexpr (PQ) {
        unsigned bq = 0;
        if (lane == 0) bq = atomicAdd(a.queue, 1u);
        batch = (int64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)bq);  // wave-uniform (SGPR)
    }
    if (batch >= a.n_batches) break;
    // The batch's particles: lane i holds particle i's row bounds and position
    // (static: row_ptr[pbeg + i] for i <= pcnt; persistent: the particle's
    // index, row start and length), loaded once; a particle takes its values
    // by v_readlane. Neighbour data is software-pipelined: while particle i's
    // ε and moments are computed, the positions of the first 64 neighbours of
    // particle i + 1 and the columns of particle i + 2 are in flight, and are
    // first used one particle later.
    const int64_t pbeg = batch * a.ppb;
    const int pcnt = (int)max((int64_t)0, min((int64_t)a.ppb, a.n_dom - pbeg));
    // (particle counts are below 2^31, so indices and row lengths are 32-bit)
    int pidv = 0, nv = 0;
    int64_t rsv = 0;
    if constexpr (PQ) {
        pidv = lane < pcnt ? (a.order ? a.order[pbeg + lane] : (int)(pbeg + lane)) : 0;
        rsv = lane < pcnt ? a.row_ptr[pidv] : 0;
        nv = lane < pcnt ? (int)(a.row_ptr[pidv + 1] - rsv) : 0;
    } else {
        pidv = (int)(pbeg + lane);
        rsv = lane <= pcnt ? a.row_ptr[pbeg + lane] : 0;
    }
    double xpv[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) xpv[d] = lane < pcnt ? a.pos_dom[d * a.n_dom + pidv] : 0.0;
    auto row_bounds = [&](int i, int64_t& rs_, int& n_) {
        rs_ = readlane_i64(rsv, i);
        if constexpr (PQ) n_ = __builtin_amdgcn_readlane(nv, i);
        else n_ = (int)(readlane_i64(rsv, i + 1) - rs_);
    };
    auto row_of = [&](int i, int64_t& rs_, int& n_, double* xp_) {
        row_bounds(i, rs_, n_);
#pragma unroll
        for (int d = 0; d < 3; ++d) xp_[d] = d < DIM ? readlane_d(xpv[d], i) : 0.0;
    };
    // columns of particle i's first 64 neighbours (lanes past its count repeat
    // its last column)
    auto load_cols = [&](int i) -> unsigned {
        if (i >= pcnt) return 0u;
        int64_t rs_;
        int n_;
        row_bounds(i, rs_, n_);
        if (n_ <= 0) return 0u;
        return (unsigned)a.col[rs_ + min(lane, n_ - 1)];
    };
    auto load_disp = [&](int64_t rs_, int n_, const double* xp_, int j, double* dv) {
        if (j < n_) {
            const int q = a.col[rs_ + j];
#pragma unroll
            for (int d = 0; d < DIM; ++d)
                dv[d] = min_image_r(xp_[d] - a.pos_sup[d * a.n_sup + q], a.len[d], a.ilen[d], (a.periodic >> d) & 1u);
        }
    };
    int64_t rs = 0;
    int n = 0;
    double xp[3] = {0.0, 0.0, 0.0};
    double cdv[3] = {0.0, 0.0, 0.0};  // lane j: displacement of neighbour j (j < 64)
    unsigned q_n = 0;                 // lane j: column of neighbour j of the next particle
    if (pcnt > 0) {
        row_of(0, rs, n, xp);
        const unsigned q0 = load_cols(0);
        if (n > 0) {
#pragma unroll
            for (int d = 0; d < DIM; ++d)
                cdv[d] = min_image_r(xp[d] - a.pos_sup[(size_t)d * a.n_sup + q0], a.len[d], a.ilen[d],
                                    (a.periodic >> d) & 1u);
        }
        q_n = load_cols(1);
    }
    for (int i = 0; i < pcnt; i += SLOTS) {
        // per slot: row, position, ε and whether its operators are assembled
        int64_t rs_s[SLOTS], pid_s[SLOTS];
        int n_s[SLOTS];
        double xp_s[SLOTS][3];
        double eps_s[SLOTS], ieps_s[SLOTS];
        bool go_s[SLOTS];
        // three slots: the lane's own neighbour's displacement and weight, in
        // registers instead of the LDS copies
        double rdv_s[SLOTS][3], rw_s[SLOTS];
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
            rs_s[s] = rs;
            n_s[s] = n;
#pragma unroll
            for (int d = 0; d < 3; ++d) xp_s[s][d] = xp[d];
            eps_s[s] = ieps_s[s] = 0.0;
            go_s[s] = false;
            rw_s[s] = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) rdv_s[s][d] = cdv[d];
            pid_s[s] = 0;
            const int ii = i + s;
            if (ii >= pcnt) {
                n_s[s] = 0;
                continue;
            }
            // (no scalar load here: its wait would also hold back the LDS work)
            int64_t p = pbeg + ii;
            if constexpr (PQ) p = a.order ? (int64_t)__builtin_amdgcn_readlane(pidv, ii) : pbeg + ii;
            pid_s[s] = p;
            int64_t rs_n = 0;
            int n_n = 0;
            double xp_n[3] = {0.0, 0.0, 0.0};
            if (ii + 1 < pcnt) row_of(ii + 1, rs_n, n_n, xp_n);
            double pq_n[3] = {0.0, 0.0, 0.0};
            if (n_n > 0) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) pq_n[d] = a.pos_sup[(size_t)d * a.n_sup + q_n];
            }
            const unsigned q_nn = load_cols(ii + 2);
            auto disp = [&](int j, double* dv) { load_disp(rs, n, xp, j, dv); };

            // ---- eps -------------------------------------------------------------
            double l1 = 0.0;
            for (int j = lane; j < n; j += kWave) {
                double dv[3] = {cdv[0], cdv[1], cdv[2]};
                if (j >= kWave) disp(j, dv);
                double sa = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) sa += fabs(dv[d]);
                l1 += sa;
            }
            l1 = wave_sum(l1);
            // ε and 1/ε as two independent divisions (1/ε within an ulp of
            // 1 / ε, and the same value wherever this kernel uses it)
            const double dn = (double)n * a.h_over_eps;
            const double eps = n > 0 ? l1 / dn : 0.0;
            const double ieps_n = dn / l1;
            const bool eps_bad = !(eps > 0.0) || !isfinite(eps);
            if constexpr (DCACHE) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) dsl[(s * 3 + d) * kWave + lane] = cdv[d];
            }
            if (n < MN || eps_bad) {  // no operator of the pass can be assembled
                const int st0 = n < M1 ? DCPSE_INFO_TOO_FEW : DCPSE_INFO_EPS_ZERO;
                asm_fail(kernarg_out(0), p, rs, n, st0, lane);
                if (L2 > 0) asm_fail(kernarg_out(1), p, rs, n, n < M2 ? DCPSE_INFO_TOO_FEW : DCPSE_INFO_EPS_ZERO, lane);
            } else {
                const double ieps = ieps_n;
                eps_s[s] = eps;
                ieps_s[s] = ieps;
                go_s[s] = true;

                // ---- moments -----------------------------------------------------------
                // M(a, b, c) = sum_q w_q x_q^a y_q^b z_q^c (x, y, z = scaled displacement).
                // Lanes own (b, c) pairs and read y^b and w z^c of the staged neighbour
                // from LDS, forming the x powers in registers; a pass covers HALVES
                // neighbours (one per group of LPN lanes), so each LDS read serves
                // the whole moment column of its pair.
                double acc[ST];
#pragma unroll
                for (int k = 0; k < ST; ++k) acc[k] = 0.0;
                for (int c0 = 0; c0 < n; c0 += CH) {
                    const int cnt = min(CH, n - c0);
                    const int jl = lane & (CH - 1);
                    double dv[3];
#pragma unroll
                    for (int d = 0; d < DIM; ++d)
                        dv[d] = CH == kWave ? cdv[d] : __shfl(cdv[d], (c0 + jl) & (kWave - 1), kWave);
                    // rows past cnt are written as zero neighbours (x = 0, w = 0):
                    // the passes then read every row of the chunk unconditionally
                    const bool vj = jl < cnt;
                    if (vj && c0 + jl >= kWave) disp(c0 + jl, dv);
                    double z[3] = {0.0, 0.0, 0.0};
#pragma unroll
                    for (int d = 0; d < DIM; ++d) z[d] = vj ? dv[d] * ieps : 0.0;
                    // neighbours 2i and 2i + 1 share a row of (even, odd) pairs
                    double* row = tab + (jl >> 1) * 2 * TROW + (jl & 1);
                    // a chunk of 32: lanes 0-31 stage the y powers, lanes 32-63
                    // the w z powers of the same neighbour; of 64: every lane both
                    if (CH == kWave || lane < CH) {
                        row[0] = z[0];
                        double pw = 1.0;
#pragma unroll
                        for (int k = 0; k <= P; ++k) {
                            row[2 * (1 + k)] = pw;
                            pw *= z[1];
                        }
                        if (NX > 0) {
                            double xt = 1.0;
#pragma unroll
                            for (int k = 0; k < ST; ++k) xt *= z[0];
#pragma unroll
                            for (int k = 0; k < NX; ++k) {
                                row[2 * (1 + PW + k)] = xt;
                                xt *= z[1];
                            }
                        }
                    }
                    if (CH == kWave || lane >= CH) {
                        double z2 = 0.0;
#pragma unroll
                        for (int d = 0; d < DIM; ++d) z2 = fma(z[d], z[d], z2);  // explicit: no contraction choice
                        double pw = vj ? exp(-z2) : 0.0;
                        if (DCACHE && c0 + jl < kWave) wsl[s * kWave + c0 + jl] = pw;
                        if (!DCACHE && CH == kWave && c0 == 0) rw_s[s] = pw;
#pragma unroll
                        for (int k = 0; k < ZW; ++k) {
                            row[2 * (1 + PW + NX + k)] = pw;
                            pw *= z[DIM - 1];
                        }
                    }
                    wave_lds_sync();
                    // a pass covers two neighbours per group; lanes of an
                    // invalid pair accumulate pair 0 and are not stored
                    auto pass = [&](double2 x, double2 yb, double2 wc) {
                        double t0 = yb.x * wc.x, t1 = yb.y * wc.y;
#pragma unroll
                        for (int k = 0; k < ST; ++k) {
                            acc[k] += t0 + t1;
                            t0 *= x.x;
                            t1 *= x.y;
                        }
                    };
                    // four passes per iteration, all rows read before any is
                    // accumulated (the later ones at immediate offsets)
                    const double2* rw0 = reinterpret_cast<const double2*>(tab) + half * TROW;
                    int j = 0;
                    for (; j + 8 * HALVES <= cnt; j += 8 * HALVES) {
                        const double2* ra = rw0 + (j >> 1) * TROW;
                        double2 xs[4], ys[4], ws[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            xs[u] = ra[u * HALVES * TROW];
                            ys[u] = ra[u * HALVES * TROW + pys];
                            ws[u] = ra[u * HALVES * TROW + pzs];
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) pass(xs[u], ys[u], ws[u]);
                    }
                    // at most four passes left: two at a time while two are
                    // needed (rows past cnt are zero neighbours of the chunk)
                    while (j < cnt) {
                        const double2* ra = rw0 + (j >> 1) * TROW;
                        if (j + 2 * HALVES < cnt) {
                            const double2* rb = ra + HALVES * TROW;
                            const double2 xa = ra[0], yba = ra[pys], wca = ra[pzs];
                            const double2 xb = rb[0], ybb = rb[pys], wcb = rb[pzs];
                            pass(xa, yba, wca);
                            pass(xb, ybb, wcb);
                            j += 4 * HALVES;
                        } else {
                            pass(ra[0], ra[pys], ra[pzs]);
                            j += 2 * HALVES;
                        }
                    }
                    wave_lds_sync();
                }
#pragma unroll
                for (int off = kWave / 2; off >= LPN; off /= 2) {
#pragma unroll
                    for (int k = 0; k < ST; ++k)
                        acc[k] = off == kWave / 2 ? half_sum(acc[k]) : acc[k] + __shfl_down(acc[k], off, kWave);
                }
                if (pv && half == 0) {
                    double* ms = mom + s * MGP;
#pragma unroll
                    for (int k = 0; k < ST; ++k) {
                        const int a = pa0 + k;
                        if (k < plen && a + pb + pc >= QMIN) ms[grid_index<GW>(a, pb, pc)] = acc[k];
                    }
                }
                wave_lds_sync();
            }
            // next particle: displacements of its first 64 neighbours
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                cdv[d] = d < DIM ? min_image_r(xp_n[d] - pq_n[d], a.len[d], a.ilen[d], (a.periodic >> d) & 1u) : 0.0;
                xp[d] = xp_n[d];
            }
            rs = rs_n;
            n = n_n;
            q_n = q_nn;
        }

        // ---- per operator: factorisation and solve of every slot, weights ----
        // the solution of each slot's system, per operator (sx, sx2)
        auto solve = [&](auto tag, int m, int abase, double b_lane, double* xs) -> uint64_t {
            constexpr int LL = decltype(tag)::L, AA = decltype(tag)::A;
            bool any = false;
#pragma unroll
            for (int s = 0; s < SLOTS; ++s) any |= go_s[s] && n_s[s] >= m;
            if (!any) return 0ull;  // no slot has enough neighbours for this operator
            bool okl;
            asm_factor_solve<DIM, LL, AA, GW, SLOTS>(mom, abase, b_lane, a.refine, tab, lbuf, xs, lane, okl);
            return __ballot(okl);
        };
        uint64_t okm1 = 0ull, okm2 = 0ull;
        if constexpr (FUSE) {
            // one factorisation for both operators
            bool any = false;
#pragma unroll
            for (int s = 0; s < SLOTS; ++s) any |= go_s[s] && n_s[s] >= MN;
            if (any) {
                bool o1, o2;
                asm_factor_solve_fused<DIM, L1, A1, L2, A2, GW, SLOTS>(mom, fab, fb1, fb2, fp1, fq2, a.refine, tab,
                                                                       lbuf, sx, sx2, lane, o1, o2);
                okm1 = __ballot(o1);
                okm2 = __ballot(o2);
            }
        } else {
            okm1 = solve(BasisTag<L1, A1>{}, M1, ab1, b1, sx);
            if constexpr (L2 > 0) okm2 = solve(BasisTag<(L2 > 0 ? L2 : 2), A2>{}, M2, ab2, b2, sx2);
        }
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
            if (!go_s[s]) continue;
            const bool use1 = n_s[s] >= M1, use2 = L2 > 0 && n_s[s] >= M2;
            const bool ok1 = use1 && ((okm1 >> (s * LS)) & 1ull), ok2 = use2 && ((okm2 >> (s * LS)) & 1ull);
            auto disp = [&](int j, double* dv) { load_disp(rs_s[s], n_s[s], xp_s[s], j, dv); };
            asm_weights<DIM, L1, A1, L2, A2, DCACHE, CH == kWave>(
                a, pid_s[s], rs_s[s], n_s[s], eps_s[s], ieps_s[s], use1, ok1, use2, ok2, sx + s * LS,
                sx2 + s * LS, dsl + (DCACHE ? s * 3 * kWave : 0), wsl + (DCACHE ? s * kWave : 0), rdv_s[s],
                rw_s[s], disp, lane);
        }
        wave_lds_sync();
    }
    if constexpr (!PQ) break;
    }  // batches
}

// ---------------------------------------------------------------------------
// generic assembly for bases without a compiled kernel (any order, up to 64
// monomials): one wavefront per particle. Lane i owns monomial i: it sums row
// i of A = sum_q w_q v_q v_q^T over the neighbours (taken one at a time, the
// values v_q broadcast through LDS), the wave factors A by Cholesky in LDS,
// solves A a = b (optionally one refinement step against A), and lanes take
// neighbours again for K_q = w_q sum_i a_i v_i(z_q). The same mathematics as
// k_assemble, without its moment grid and compile-time tables.
// ---------------------------------------------------------------------------
struct GenericArgs {
    const double* pos_sup;
    int64_t n_sup;
    const double* pos_dom;
    int64_t n_dom;
    const int64_t* row_ptr;
    const int32_t* col;
    double len[3], ilen[3];
    unsigned periodic;
    double h_over_eps;
    int refine;
    int M;
    const uint32_t* exps;  // M * dim exponents (device)
    AsmOp op;
};

__device__ __forceinline__ double mono_rt(const uint32_t* e, const double* z, int dim) {
    double v = 1.0;
    for (int d = 0; d < dim; ++d)
        for (uint32_t k = 0; k < e[d]; ++k) v *= z[d];
    return v;
}

template <int DIM, int MB>
__global__ __launch_bounds__(64) void k_assemble_generic(const GenericArgs a) {
    extern __shared__ double gsh[];
    double* A = gsh;              // M x M, row-major
    double* vq = gsh + MB * MB;   // broadcast values, then the solution
    uint32_t* ex = reinterpret_cast<uint32_t*>(vq + kWave);  // M * DIM exponents
    const int lane = threadIdx.x;
    const int M = a.M;
    const bool act = lane < M;
    for (int i = lane; i < M * DIM; i += kWave) ex[i] = a.exps[i];
    wave_lds_sync();
    const double bi = act ? a.op.b[lane] : 0.0;
    for (int64_t p = blockIdx.x; p < a.n_dom; p += gridDim.x) {
        const int64_t rs = a.row_ptr[p];
        const int n = (int)(a.row_ptr[p + 1] - rs);
        double xp[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < DIM; ++d) xp[d] = a.pos_dom[d * a.n_dom + p];
        auto disp = [&](int j, double* x) {
            const int q = a.col[rs + j];
#pragma unroll
            for (int d = 0; d < DIM; ++d)
                x[d] = min_image_r(xp[d] - a.pos_sup[d * a.n_sup + q], a.len[d], a.ilen[d], (a.periodic >> d) & 1u);
        };
        // eps = mean |x_pq|_1 / h_over_eps (Vandermonde.hpp:63-81), one division
        double l1 = 0.0;
        for (int j = lane; j < n; j += kWave) {
            double x[3];
            disp(j, x);
#pragma unroll
            for (int d = 0; d < DIM; ++d) l1 += fabs(x[d]);
        }
        l1 = wave_sum(l1);
        const double eps = n > 0 ? l1 / ((double)n * a.h_over_eps) : 0.0;
        if (n < M || !(eps > 0.0) || !isfinite(eps)) {
            asm_fail(a.op, p, rs, n, n < M ? DCPSE_INFO_TOO_FEW : DCPSE_INFO_EPS_ZERO, lane);
            continue;
        }
        const double ieps = 1.0 / eps;
        double arow[MB];
#pragma unroll
        for (int j = 0; j < MB; ++j) arow[j] = 0.0;
        for (int q = 0; q < n; ++q) {
            double x[3] = {0.0, 0.0, 0.0}, z[3] = {0.0, 0.0, 0.0}, z2 = 0.0;
            disp(q, x);
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                z[d] = x[d] * ieps;
                z2 = fma(z[d], z[d], z2);
            }
            const double w = exp(-z2);
            const double v = act ? mono_rt(ex + lane * DIM, z, DIM) : 0.0;
            vq[lane] = v;
            wave_lds_sync();
            const double wv = w * v;
#pragma unroll
            for (int j = 0; j < MB; ++j)
                if (j < M) arow[j] = fma(wv, vq[j], arow[j]);
            wave_lds_sync();
        }
        if (act) {
#pragma unroll
            for (int j = 0; j < MB; ++j)
                if (j < M) A[lane * M + j] = arow[j];
        }
        wave_lds_sync();
        // Cholesky in place (lower triangle); a pivot that is not positive
        // and finite ends it (no weights, DCPSE_INFO_NOT_SPD)
        bool ok = true;
        for (int k = 0; k < M; ++k) {
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
        if (!ok) {
            asm_fail(a.op, p, rs, n, DCPSE_INFO_NOT_SPD, lane);
            wave_lds_sync();
            continue;
        }
        auto solve = [&](double r) {
            for (int k = 0; k < M; ++k) {
                if (lane == k) r /= A[k * M + k];
                const double yk = readlane_d(r, k);
                if (lane > k && lane < M) r = fma(-A[lane * M + k], yk, r);
            }
            for (int k = M - 1; k >= 0; --k) {
                if (lane == k) r /= A[k * M + k];
                const double xk = readlane_d(r, k);
                if (lane < k) r = fma(-A[k * M + lane], xk, r);
            }
            return act ? r : 0.0;
        };
        double xa = solve(bi);
        if (a.refine) {
            // one step against the unfactored A (row i still in arow)
            double rr = bi;
            for (int j = 0; j < M; ++j) {
                const double xj = readlane_d(xa, j);
#pragma unroll
                for (int jj = 0; jj < MB; ++jj)
                    if (jj == j) rr = fma(-arow[jj], xj, rr);
            }
            xa += solve(act ? rr : 0.0);
        }
        vq[lane] = xa;
        wave_lds_sync();
        // kernel weights K_q = w_q sum_i a_i v_i(z_q) (Dcpse.hpp:798-809)
        const double eip = ipow_rt(ieps, a.op.border);
        bool bad = false;
        for (int j = lane; j < n; j += kWave) {
            double x[3] = {0.0, 0.0, 0.0}, z[3] = {0.0, 0.0, 0.0}, z2 = 0.0;
            disp(j, x);
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                z[d] = x[d] * ieps;
                z2 = fma(z[d], z[d], z2);
            }
            double sum = 0.0;
            for (int i = 0; i < M; ++i) sum = fma(vq[i], mono_rt(ex + i * DIM, z, DIM), sum);
            const double K = sum * exp(-z2);
            bad |= !isfinite(K);
            if (a.op.K64) a.op.K64[rs + j] = K;
            if (a.op.W64) a.op.W64[rs + j] = K * eip;
            if (a.op.W32) a.op.W32[rs + j] = (float)(K * eip);
        }
        const bool anybad = __ballot(bad) != 0ull;
        if (lane == 0) {
            if (a.op.eps) a.op.eps[p] = eps;
            if (a.op.eps_inv_pow) a.op.eps_inv_pow[p] = eip;
            if (a.op.info) a.op.info[p] = anybad ? DCPSE_INFO_NONFINITE : DCPSE_INFO_OK;
        }
        wave_lds_sync();
    }
}

template <int DIM, int MB>
static hipError_t launch_generic(const GenericArgs& a, int num_cus, hipStream_t s) {
    const int64_t blocks = std::min<int64_t>(a.n_dom, (int64_t)num_cus * 16);
    if (blocks <= 0) return hipSuccess;
    const size_t lds = (size_t)(MB * MB + kWave) * sizeof(double) + (size_t)MB * DIM * sizeof(uint32_t);
    k_assemble_generic<DIM, MB><<<(unsigned)blocks, kWave, lds, s>>>(a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// host-side basis rules (the device templates enumerate the same way)
// ---------------------------------------------------------------------------
struct HostBasis {
    int dim = 0, L = 0, amin = 0, m = 0;
    std::vector<uint32_t> e;
};

static HostBasis host_basis(int dim, const uint32_t* beta, uint32_t order) {
    HostBasis B;
    B.dim = dim;
    uint32_t s = 0;
    for (int d = 0; d < dim; ++d) s += beta[d];
    B.L = (int)(s + order);
    B.amin = (s % 2 == 0 && s != 0) ? 1 : 0;
    long total = 1;
    for (int d = 0; d < dim; ++d) total *= B.L;
    for (long lin = 0; lin < total; ++lin) {
        uint32_t k[3] = {0, 0, 0};
        long r = lin;
        int sum = 0;
        for (int d = 0; d < dim; ++d) {
            k[d] = (uint32_t)(r % B.L);
            r /= B.L;
            sum += (int)k[d];
        }
        if (sum >= B.amin && sum < B.L) {
            for (int d = 0; d < dim; ++d) B.e.push_back(k[d]);
            ++B.m;
        }
    }
    return B;
}

// b_i = sign * (d^beta m_i)(0), sign = +1 for even |beta| (DcpseRhs.hpp:27-60):
// nonzero only for alpha_i == beta, where it is sign * beta!.
static void host_rhs(const HostBasis& B, const uint32_t* beta, double* b) {
    uint32_t s = 0;
    for (int d = 0; d < B.dim; ++d) s += beta[d];
    const double sign = (s % 2 == 0) ? 1.0 : -1.0;
    for (int i = 0; i < B.m; ++i) {
        bool eq = true;
        double fact = 1.0;
        for (int d = 0; d < B.dim; ++d) {
            if (B.e[i * B.dim + d] != beta[d]) eq = false;
            for (uint32_t k = 2; k <= beta[d]; ++k) fact *= k;
        }
        b[i] = eq ? sign * fact : 0.0;
    }
}

// Static grid: one block (wavefront) per batch. Persistent grid: as many
// wavefronts as the GPU holds at once (the kernel's occupancy times the CU
// count), each taking batches from args.queue until none is left.
template <int DIM, int L1, int A1, int L2, int A2, int CH>
static hipError_t launch_asm(const AsmArgs& args, int num_cus, hipStream_t s) {
    int64_t blocks = args.n_batches;
    if (args.queue) {
        int per_cu = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_assemble<DIM, L1, A1, L2, A2, CH, true>, kWave, 0);
        if (e != hipSuccess) return e;
        blocks = std::min<int64_t>(blocks, (int64_t)std::max(per_cu, 1) * num_cus);
        if (blocks <= 0) return hipSuccess;
        k_assemble<DIM, L1, A1, L2, A2, CH, true><<<(unsigned)blocks, kWave, 0, s>>>(args);
        return hipGetLastError();
    }
    if (blocks <= 0) return hipSuccess;
    k_assemble<DIM, L1, A1, L2, A2, CH, false><<<(unsigned)blocks, kWave, 0, s>>>(args);
    return hipGetLastError();
}

typedef hipError_t (*AsmLauncher)(const AsmArgs&, int, hipStream_t);

// Single-operator kernels for every basis the library supports, and
// two-operator kernels for the pairs assembled together most often
// (second derivatives or the Laplacian with first derivatives, and two
// operators of one basis), at convergence order 2. Neighbour chunks of 32
// for every kernel, of 64 for the order-2 ones.
static AsmLauncher pick_launcher(int dim, int L, int amin, int L2 = 0, int amin2 = 0, int ch = 32) {
#define DCPSE_ASM_CASE(D, LL, A) \
    if (dim == D && L == LL && amin == A && L2 == 0 && ch == 32) return &launch_asm<D, LL, A, 0, 0, 32>;
#define DCPSE_ASM_CASE64(D, LL, A) \
    DCPSE_ASM_CASE(D, LL, A)       \
    if (dim == D && L == LL && amin == A && L2 == 0 && ch == 64) return &launch_asm<D, LL, A, 0, 0, 64>;
    DCPSE_ASM_CASE(2, 2, 0) DCPSE_ASM_CASE(2, 2, 1)
    DCPSE_ASM_CASE64(2, 3, 0) DCPSE_ASM_CASE(2, 3, 1)
    DCPSE_ASM_CASE(2, 4, 0) DCPSE_ASM_CASE64(2, 4, 1)
    DCPSE_ASM_CASE(2, 5, 0) DCPSE_ASM_CASE(2, 5, 1)
    DCPSE_ASM_CASE(2, 6, 0) DCPSE_ASM_CASE(2, 6, 1)
    DCPSE_ASM_CASE(3, 2, 0) DCPSE_ASM_CASE(3, 2, 1)
    DCPSE_ASM_CASE64(3, 3, 0) DCPSE_ASM_CASE(3, 3, 1)
    DCPSE_ASM_CASE(3, 4, 0) DCPSE_ASM_CASE64(3, 4, 1)
    DCPSE_ASM_CASE(3, 5, 0) DCPSE_ASM_CASE(3, 5, 1)
#undef DCPSE_ASM_CASE64
#undef DCPSE_ASM_CASE
#define DCPSE_ASM_PAIR(D, LA, AA, LB, AB)                                                  \
    if (dim == D && L == LA && amin == AA && L2 == LB && amin2 == AB)                     \
        return ch == 64 ? &launch_asm<D, LA, AA, LB, AB, 64> : &launch_asm<D, LA, AA, LB, AB, 32>;
    DCPSE_ASM_PAIR(2, 4, 1, 3, 0) DCPSE_ASM_PAIR(2, 3, 0, 3, 0) DCPSE_ASM_PAIR(2, 4, 1, 4, 1)
    DCPSE_ASM_PAIR(3, 4, 1, 3, 0) DCPSE_ASM_PAIR(3, 3, 0, 3, 0) DCPSE_ASM_PAIR(3, 4, 1, 4, 1)
#undef DCPSE_ASM_PAIR
    return nullptr;
}

// ---------------------------------------------------------------------------
// batches bucketed by row length: particles of longer rows first, so the
// persistent grid ends on short rows (longest-first scheduling); a stable
// counting sort by tiles, bucket-major, keeps the particles of a bucket in
// index order within each tile
// ---------------------------------------------------------------------------
constexpr int kBucketTile = 2048;

__global__ __launch_bounds__(256) void k_asm_bucket_hist(const int64_t* __restrict__ row_ptr, int64_t n, int width,
                                                         int nb, int64_t n_tiles, int32_t* __restrict__ counts) {
    extern __shared__ int32_t h[];
    for (int b = threadIdx.x; b < nb; b += blockDim.x) h[b] = 0;
    __syncthreads();
    const int64_t t = blockIdx.x;
    for (int64_t p = t * kBucketTile + threadIdx.x; p < min(n, (t + 1) * kBucketTile); p += blockDim.x) {
        const int len = (int)(row_ptr[p + 1] - row_ptr[p]);
        const int b = nb - 1 - min(len / width, nb - 1);  // longest rows in bucket 0
        atomicAdd(&h[b], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x) counts[(int64_t)b * n_tiles + t] = h[b];
}

// One wavefront per tile writes its particles bucket by bucket in index
// order: the tile's bucket numbers are staged in LDS once, and a ballot per
// bucket and round of 64 particles gives each lane its rank among the tile's
// particles of that bucket.
__global__ __launch_bounds__(64) void k_asm_bucket_scatter(const int64_t* __restrict__ row_ptr, int64_t n, int width,
                                                           int nb, int64_t n_tiles, const int64_t* __restrict__ start,
                                                           int32_t* __restrict__ order) {
    __shared__ uint16_t bk[kBucketTile];
    const int64_t t = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t p0 = t * kBucketTile, p1 = min(n, (t + 1) * kBucketTile);
    const int cnt = (int)(p1 - p0);
    for (int i = lane; i < cnt; i += kWave) {
        const int len = (int)(row_ptr[p0 + i + 1] - row_ptr[p0 + i]);
        bk[i] = (uint16_t)(nb - 1 - min(len / width, nb - 1));
    }
    wave_lds_sync();
    for (int b = 0; b < nb; ++b) {
        int64_t at = start[(int64_t)b * n_tiles + t];
        const int64_t end = start[(int64_t)b * n_tiles + t + 1];
        if (at == end) continue;
        for (int i0 = 0; i0 < cnt; i0 += kWave) {
            const int i = i0 + lane;
            const bool mine = i < cnt && bk[i] == b;
            const unsigned long long m = __ballot(mine);
            if (mine) order[at + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)(p0 + i);
            at += __popcll(m);
        }
    }
}

// The basis and summed right-hand side of one operator (n_sig signatures
// sharing one basis) into op.b; returns false if the signatures do not share it.
static bool setup_op(int dim, int n_sig, const uint32_t* betas, const double* coeffs, uint32_t order, AsmOp& op,
                     HostBasis& B0) {
    B0 = host_basis(dim, betas, order);
    uint32_t border = 0;
    for (int d = 0; d < dim; ++d) border += betas[d];
    std::vector<double> bk(B0.m);
    for (int k = 0; k < n_sig; ++k) {
        HostBasis Bk = host_basis(dim, betas + k * dim, order);
        uint32_t bo = 0;
        for (int d = 0; d < dim; ++d) bo += betas[k * dim + d];
        if (Bk.L != B0.L || Bk.amin != B0.amin || bo != border) return false;
        host_rhs(Bk, betas + k * dim, bk.data());
        const double c = coeffs ? coeffs[k] : 1.0;
        for (int i = 0; i < B0.m && i < 64; ++i) op.b[i] += c * bk[i];
    }
    op.border = border;
    return true;
}

}  // namespace dcpse

using namespace dcpse;

extern "C" {

int32_t dcpse_basis_size(int32_t dim, const uint32_t* beta, uint32_t order) {
    if ((dim != 2 && dim != 3) || !beta) return -1;
    return host_basis(dim, beta, order).m;
}

int32_t dcpse_basis(int32_t dim, const uint32_t* beta, uint32_t order, uint32_t* exps, int32_t cap) {
    if ((dim != 2 && dim != 3) || !beta) return 0;
    HostBasis B = host_basis(dim, beta, order);
    if (B.m > cap || !exps) return -B.m;
    std::copy(B.e.begin(), B.e.end(), exps);
    return B.m;
}

int32_t dcpse_rhs(int32_t dim, const uint32_t* beta, uint32_t order, double* b, int32_t cap) {
    if ((dim != 2 && dim != 3) || !beta) return 0;
    HostBasis B = host_basis(dim, beta, order);
    if (B.m > cap || !b) return -B.m;
    host_rhs(B, beta, b);
    return B.m;
}

static dcpse_status assemble_ops(dcpse_ctx* ctx, int32_t dim, const double* pos_sup, int64_t n_sup,
                                 const double* pos_dom, int64_t n_dom, const double* box_lo, const double* box_hi,
                                 uint32_t periodic_mask, const int64_t* row_ptr, const int32_t* col, int32_t n_ops,
                                 const int32_t* n_sig, const uint32_t* betas, const double* coeffs, uint32_t order,
                                 double h_over_eps, uint32_t flags, double* const* eps, double* const* eps_inv_pow,
                                 double* const* K64, float* const* W32, double* const* W64, int32_t* const* info) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (dim != 2 && dim != 3) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "dim must be 2 or 3");
    // positions of an empty set may be null (a particle set without storage)
    if ((!pos_sup && n_sup > 0) || !row_ptr || n_sup < 0 || n_dom < 0 || n_ops < 1 || !n_sig || !betas)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    if (!(h_over_eps > 0.0)) return fail(ctx, DCPSE_ERR_INVALID, "h_over_eps must be > 0");
    if (n_sup >= ((int64_t)1 << 31) || n_dom >= ((int64_t)1 << 31))
        return fail(ctx, DCPSE_ERR_RANGE, "particle counts must be < 2^31");
    for (int o = 0; o < n_ops; ++o)
        if (n_sig[o] < 1) return fail(ctx, DCPSE_ERR_INVALID, "operator %d has no signature", o);
    // a NULL domain is the support set itself: its row count is n_sup (a
    // smaller n_dom would size the outputs for fewer rows than are written)
    if (pos_dom == nullptr && n_dom != n_sup)
        return fail(ctx, DCPSE_ERR_INVALID, "pos_dom NULL (the support set as domain) needs n_dom == n_sup");
    if (pos_dom == nullptr) pos_dom = pos_sup;
    if (periodic_mask && (!box_lo || !box_hi))
        return fail(ctx, DCPSE_ERR_INVALID, "periodic boundaries need box_lo/box_hi");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));

    std::vector<AsmOp> ops(n_ops);
    std::vector<HostBasis> bases(n_ops);
    int64_t off_sig = 0, off_c = 0;
    for (int o = 0; o < n_ops; ++o) {
        AsmOp& op = ops[o];
        memset(&op, 0, sizeof(op));
        if (!setup_op(dim, n_sig[o], betas + off_sig * dim, coeffs ? coeffs + off_c : nullptr, order, op, bases[o]))
            return fail(ctx, DCPSE_ERR_INVALID, "combined signatures must share |beta| (one basis)");
        if (bases[o].m > 64)
            return fail(ctx, DCPSE_ERR_UNSUPPORTED,
                        "basis of %d monomials exceeds 64 (dim=%d, |beta|+order=%d; up to 6 in 3D, 10 in 2D)",
                        bases[o].m, dim, bases[o].L);
        op.eps = eps ? eps[o] : nullptr;
        op.eps_inv_pow = eps_inv_pow ? eps_inv_pow[o] : nullptr;
        op.K64 = K64 ? K64[o] : nullptr;
        op.W32 = W32 ? W32[o] : nullptr;
        op.W64 = W64 ? W64[o] : nullptr;
        op.info = info ? info[o] : nullptr;
        off_sig += n_sig[o];
        off_c += n_sig[o];
    }
    AsmArgs args;
    memset(&args, 0, sizeof(args));
    args.pos_sup = pos_sup;
    args.n_sup = n_sup;
    args.pos_dom = pos_dom;
    args.n_dom = n_dom;
    args.row_ptr = row_ptr;
    args.col = col;
    for (int d = 0; d < 3; ++d) {
        args.len[d] = (d < dim && box_lo && box_hi) ? box_hi[d] - box_lo[d] : 1.0;
        args.ilen[d] = 1.0 / args.len[d];
    }
    args.periodic = periodic_mask;
    args.h_over_eps = h_over_eps;
    args.refine = (flags & DCPSE_ASM_REFINE) ? 1 : 0;
    // one lane per particle of a batch; the static grid also reads the row
    // start of the particle after the last from lane ppb, so at most 63
    args.ppb = std::min(ctx->asm_persistent ? 64 : 63, std::max(1, ctx->asm_ppb));
    if (n_dom == 0) return DCPSE_OK;
    args.n_batches = (n_dom + args.ppb - 1) / args.ppb;
    const int ch = ctx->asm_chunk;
    // bases without a compiled kernel: the generic kernel, one operator at a time
    std::vector<int> fast;
    for (int k = 0; k < n_ops; ++k) {
        if (pick_launcher(dim, bases[k].L, bases[k].amin)) {
            fast.push_back(k);
            continue;
        }
        GenericArgs g;
        memset(&g, 0, sizeof(g));
        g.pos_sup = pos_sup;
        g.n_sup = n_sup;
        g.pos_dom = pos_dom;
        g.n_dom = n_dom;
        g.row_ptr = row_ptr;
        g.col = col;
        for (int d = 0; d < 3; ++d) {
            g.len[d] = args.len[d];
            g.ilen[d] = args.ilen[d];
        }
        g.periodic = periodic_mask;
        g.h_over_eps = h_over_eps;
        g.refine = args.refine;
        g.M = bases[k].m;
        g.op = ops[k];
        Workspace gw;
        dcpse_status gst = reserve_workspace(ctx, (size_t)g.M * dim * 4 + 1024, &gw);
        if (gst) return gst;
        uint32_t* d_ex = gw.take<uint32_t>((size_t)g.M * dim);
        DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(d_ex, bases[k].e.data(), (size_t)g.M * dim * 4, hipMemcpyHostToDevice,
                                            ctx->stream));
        g.exps = d_ex;
        hipError_t e;
        if (dim == 2) e = g.M <= 32 ? launch_generic<2, 32>(g, ctx->num_cus, ctx->stream)
                                    : launch_generic<2, 64>(g, ctx->num_cus, ctx->stream);
        else e = g.M <= 32 ? launch_generic<3, 32>(g, ctx->num_cus, ctx->stream)
                           : launch_generic<3, 64>(g, ctx->num_cus, ctx->stream);
        DCPSE_HIP_CHECK(ctx, e);
        // the exponents are read before the next call reuses the workspace
        DCPSE_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (fast.size() < (size_t)n_ops) {
        std::vector<AsmOp> fops;
        std::vector<HostBasis> fbases;
        for (int k : fast) {
            fops.push_back(ops[k]);
            fbases.push_back(bases[k]);
        }
        ops.swap(fops);
        bases.swap(fbases);
        n_ops = (int32_t)ops.size();
        if (n_ops == 0) return DCPSE_OK;
    }
    // launch shape (dcpse_set_assembly_launch): persistent grid with a batch
    // counter, and particles bucketed by row length
    const bool persistent = ctx->asm_persistent != 0 && n_ops <= 64;
    const int bw = persistent ? ctx->asm_bucket : 0;  // the static grid runs index order
    const int nb = bw > 0 ? std::max(1, (DCPSE_MAX_ROW + bw) / bw) : 0;
    const int64_t n_tiles = (n_dom + kBucketTile - 1) / kBucketTile;
    if (persistent || bw > 0) {
        Workspace w;
        const size_t need = 64 * 4 + (bw > 0 ? (size_t)n_dom * 4 + ((size_t)nb * n_tiles + 1) * 12 +
                                                   scan_workspace_bytes((int64_t)nb * n_tiles) + 1024
                                             : 0) + 1024;
        dcpse_status st = reserve_workspace(ctx, need, &w);
        if (st) return st;
        unsigned* queues = w.take<unsigned>(64);
        if (!queues) return fail(ctx, DCPSE_ERR_INVALID, "assembly workspace sizing error");
        if (persistent) {
            fill_i32(ctx->stream, reinterpret_cast<int32_t*>(queues), 64, 0);
            args.queue = queues;
        }
        if (bw > 0) {
            int32_t* order = w.take<int32_t>((size_t)n_dom);
            int32_t* counts = w.take<int32_t>((size_t)nb * n_tiles + 1);
            int64_t* start = w.take<int64_t>((size_t)nb * n_tiles + 1);
            if (!start) return fail(ctx, DCPSE_ERR_INVALID, "assembly workspace sizing error");
            k_asm_bucket_hist<<<(unsigned)n_tiles, 256, (size_t)nb * 4, ctx->stream>>>(row_ptr, n_dom, bw, nb,
                                                                                      n_tiles, counts);
            DCPSE_LAUNCH_CHECK(ctx);
            if ((st = scan_counts(ctx, counts, (int64_t)nb * n_tiles, start, &w))) return st;
            k_asm_bucket_scatter<<<(unsigned)n_tiles, kWave, 0, ctx->stream>>>(row_ptr, n_dom, bw, nb, n_tiles, start,
                                                                              order);
            DCPSE_LAUNCH_CHECK(ctx);
            args.order = order;
        }
    }
    // operators in pairs sharing one pass where a two-operator kernel exists
    // (larger basis first), one at a time otherwise; a basis without a kernel
    // for the chosen chunk runs with chunks of 32
    int o = 0, pass = 0;
    while (o < n_ops) {
        AsmLauncher pair = nullptr;
        int first = o, second = o + 1;
        if (o + 1 < n_ops) {
            if (bases[o + 1].L > bases[o].L || (bases[o + 1].L == bases[o].L && bases[o + 1].amin > bases[o].amin))
                std::swap(first, second);
            pair = pick_launcher(dim, bases[first].L, bases[first].amin, bases[second].L, bases[second].amin, ch);
        }
        AsmLauncher launch;
        if (pair) {
            args.op[0] = ops[first];
            args.op[1] = ops[second];
            launch = pair;
            o += 2;
        } else {
            args.op[0] = ops[o];
            memset(&args.op[1], 0, sizeof(AsmOp));
            launch = pick_launcher(dim, bases[o].L, bases[o].amin, 0, 0, ch);
            if (!launch) launch = pick_launcher(dim, bases[o].L, bases[o].amin, 0, 0, 32);
            o += 1;
        }
        if (args.queue && pass > 0) args.queue += 1;  // a fresh counter per pass
        DCPSE_HIP_CHECK(ctx, launch(args, ctx->num_cus, ctx->stream));
        ++pass;
    }
    return DCPSE_OK;
}

dcpse_status dcpse_assemble(dcpse_ctx* ctx, int32_t dim, const double* pos_sup, int64_t n_sup,
                            const double* pos_dom, int64_t n_dom, const double* box_lo, const double* box_hi,
                            uint32_t periodic_mask, const int64_t* row_ptr, const int32_t* col, int32_t n_sig,
                            const uint32_t* betas, const double* coeffs, uint32_t order, double h_over_eps,
                            uint32_t flags, double* eps, double* eps_inv_pow, double* K64, float* W32,
                            double* W64, int32_t* info) {
    return assemble_ops(ctx, dim, pos_sup, n_sup, pos_dom, n_dom, box_lo, box_hi, periodic_mask, row_ptr, col, 1,
                        &n_sig, betas, coeffs, order, h_over_eps, flags, &eps, &eps_inv_pow, &K64, &W32, &W64, &info);
}

dcpse_status dcpse_assemble_multi(dcpse_ctx* ctx, int32_t dim, const double* pos_sup, int64_t n_sup,
                                  const double* pos_dom, int64_t n_dom, const double* box_lo, const double* box_hi,
                                  uint32_t periodic_mask, const int64_t* row_ptr, const int32_t* col, int32_t n_ops,
                                  const int32_t* n_sig, const uint32_t* betas, const double* coeffs, uint32_t order,
                                  double h_over_eps, uint32_t flags, double* const* eps, double* const* eps_inv_pow,
                                  double* const* K64, float* const* W32, double* const* W64, int32_t* const* info) {
    return assemble_ops(ctx, dim, pos_sup, n_sup, pos_dom, n_dom, box_lo, box_hi, periodic_mask, row_ptr, col,
                        n_ops, n_sig, betas, coeffs, order, h_over_eps, flags, eps, eps_inv_pow, K64, W32, W64, info);
}

}  // extern "C"
