// Shared device helpers and compile-time DC-PSE basis tables for gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dcpse {

constexpr int kWave = 64;

// ---- wave-level primitives --------------------------------------------------

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)x, lane);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ int64_t readlane_i64(int64_t v, int lane) {
    const int lo = __builtin_amdgcn_readlane((int)v, lane);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
    return ((int64_t)hi << 32) | (uint32_t)lo;
}

// v from another lane of the same 16-lane row by a DPP pattern (both halves).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// v of lane l plus v of its partner l ^ 32 in the other half, in every lane
// (lower half's value first): two v_permlane32_swap per 32-bit word, no LDS.
__device__ __forceinline__ double half_sum(double v) {
    const long long x = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)x, (unsigned)x, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(x >> 32), (unsigned)(x >> 32), false, false);
    const double a = __longlong_as_double(((long long)hi[0] << 32) | (unsigned)lo[0]);
    const double b = __longlong_as_double(((long long)hi[1] << 32) | (unsigned)lo[1]);
    return a + b;
}

// Sum over the 64 lanes (all active), wave-uniform: within each 16-lane row
// by DPP exchanges (lane ^ 1, lane ^ 2, mirrored half-rows, mirrored rows;
// every lane of a row ends with the same bits), then the four row sums by
// v_readlane.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_d<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_d<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_d<0x141>(v);  // row_half_mirror
    v += dpp_d<0x140>(v);  // row_mirror
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

// Inclusive prefix sum over the 64 lanes (all active) by DPP: a window of
// four lanes (row_shr 1..3), then row_shr 4 and 8 inside each 16-lane row
// (upper banks only), then row_bcast 15 and 31 across rows.
__device__ __forceinline__ int wave_incl_scan(int v) {
    int s = v + __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true);
    s += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true);
    s += __builtin_amdgcn_update_dpp(0, v, 0x113, 0xF, 0xF, true);
    s += __builtin_amdgcn_update_dpp(0, s, 0x114, 0xF, 0xE, false);
    s += __builtin_amdgcn_update_dpp(0, s, 0x118, 0xF, 0xC, false);
    s += __builtin_amdgcn_update_dpp(0, s, 0x142, 0xA, 0xF, false);
    s += __builtin_amdgcn_update_dpp(0, s, 0x143, 0xC, 0xF, false);
    return s;
}

// Inclusive prefix maximum over the 64 lanes (all active) of values >= -1,
// by the same DPP pattern as wave_incl_scan (lanes without a source keep -1).
__device__ __forceinline__ int wave_incl_max(int v) {
    int s = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x111, 0xF, 0xF, false));
    s = max(s, __builtin_amdgcn_update_dpp(-1, v, 0x112, 0xF, 0xF, false));
    s = max(s, __builtin_amdgcn_update_dpp(-1, v, 0x113, 0xF, 0xF, false));
    s = max(s, __builtin_amdgcn_update_dpp(-1, s, 0x114, 0xF, 0xE, false));
    s = max(s, __builtin_amdgcn_update_dpp(-1, s, 0x118, 0xF, 0xC, false));
    s = max(s, __builtin_amdgcn_update_dpp(-1, s, 0x142, 0xA, 0xF, false));
    s = max(s, __builtin_amdgcn_update_dpp(-1, s, 0x143, 0xC, 0xF, false));
    return s;
}

// Minimum-image displacement component, v - len * rint(v / len) as the
// oracle computes it, with the reciprocal box length: rint(v * ilen) equals
// rint(v / len) unless v / len lies within an ulp of a half-integer, i.e. a
// separation of half the box, which no support of radius below half the box
// contains, so the displacement is the same double.
__device__ __forceinline__ double min_image_r(double v, double len, double ilen, bool periodic) {
    return periodic ? v - len * rint(v * ilen) : v;
}

__device__ __forceinline__ int lane_prefix(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// Orders LDS traffic of the lanes of one wavefront (all lanes that share the
// wave's private LDS slice) without a workgroup barrier.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- compile-time basis / moment tables ---------------------------------------
//
// The monomial basis of a DC-PSE operator with signature beta and convergence
// order r is every alpha in [0, L)^DIM with AMIN <= |alpha| < L, L = |beta| + r,
// AMIN = 1 for even nonzero |beta| and 0 otherwise, enumerated with dimension 0
// varying fastest (openfpm grid iteration order). The moment matrix
// A_ij = sum_q w_q z_q^(alpha_i + alpha_j) is a function of the moments
// M(gamma) = sum_q w_q z_q^gamma for 2*AMIN <= |gamma| <= 2(L-1).

constexpr int ipow_c(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

template <int DIM, int L, int AMIN>
struct BasisShape {
    static constexpr int count_m() {
        int c = 0;
        for (int lin = 0; lin < ipow_c(L, DIM); ++lin) {
            int s = 0, r = lin;
            for (int d = 0; d < DIM; ++d) { s += r % L; r /= L; }
            if (s >= AMIN && s < L) ++c;
        }
        return c;
    }
    static constexpr int P = 2 * (L - 1);  // largest moment exponent
    static constexpr int GW = P + 1;        // moment grid width per dimension
    static constexpr int count_g() {
        int c = 0;
        for (int lin = 0; lin < ipow_c(GW, DIM); ++lin) {
            int s = 0, r = lin;
            for (int d = 0; d < DIM; ++d) { s += r % GW; r /= GW; }
            if (s >= 2 * AMIN && s <= P) ++c;
        }
        return c;
    }
    static constexpr int M = count_m();
    static constexpr int G = count_g();
    static constexpr int MG = ipow_c(GW, DIM);
};

template <int DIM, int L, int AMIN>
struct BasisTables {
    using S = BasisShape<DIM, L, AMIN>;
    int e[S::M][3];  // basis exponents
    int g[S::G][3];  // moment exponents
    constexpr BasisTables() : e{}, g{} {
        int c = 0;
        for (int lin = 0; lin < ipow_c(L, DIM); ++lin) {
            int k[3] = {0, 0, 0}, s = 0, r = lin;
            for (int d = 0; d < DIM; ++d) { k[d] = r % L; r /= L; s += k[d]; }
            if (s >= AMIN && s < L) {
                e[c][0] = k[0]; e[c][1] = k[1]; e[c][2] = k[2];
                ++c;
            }
        }
        c = 0;
        for (int lin = 0; lin < ipow_c(S::GW, DIM); ++lin) {
            int k[3] = {0, 0, 0}, s = 0, r = lin;
            for (int d = 0; d < DIM; ++d) { k[d] = r % S::GW; r /= S::GW; s += k[d]; }
            if (s >= 2 * AMIN && s <= S::P) {
                g[c][0] = k[0]; g[c][1] = k[1]; g[c][2] = k[2];
                ++c;
            }
        }
    }
};

}  // namespace dcpse
