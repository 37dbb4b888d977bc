// Sliced-ELL (SELL-64) execution layout of a DC-PSE operator, FP32.
//
// The canonical operator storage is the reference's CSR (kerOffsets +
// calcKernels, src/DCPSE/Dcpse.hpp:59-63). For the FP32 apply the same
// nonzeros are also laid out one wavefront per slice of 64 consecutive rows:
// slice s holds rows [64s, 64s+64), its width w_s is the longest row rounded up
// to 4. Inside a slice the rows are ordered by decreasing length (ties by
// index): lane l of slice s serves row sell_row[64s+l] (-1 for the lanes of the
// last slice that have no row), and slot j of that lane sits at
//     slice_ptr[s] + (j/4)*256 + 4*l + (j%4).
// A wavefront then reads each stream as 1 KB contiguous per load instruction
// (lane l takes 4 consecutive slots of its own row with one 16-byte load), the
// f gathers of neighbouring lanes hit the same cache lines (the lanes' rows are
// the same 64 consecutive particles, taking the same neighbour slot), and every
// lane owns its row's sum outright (no cross-lane reduction). Because lengths
// decrease with the lane, the padding of a slot group sits in the highest
// lanes: a 128-byte line whose 8 lanes are all past their rows is never
// fetched (the kernel re-reads group 0 there), so the padding costs memory
// footprint but hardly any HBM traffic. Padded slots hold column 0 and
// weight 0 and are masked out by the row length.
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "dcpse_common.hpp"
#include "dcpse_internal.hpp"

namespace dcpse {

typedef int32_t vi4a __attribute__((ext_vector_type(4)));
typedef int16_t vs4a __attribute__((ext_vector_type(4)));
typedef float vf4a __attribute__((ext_vector_type(4)));

constexpr int kSellWaves = 4;  // slices (wavefronts) per block

// counts[s] = 64 * (longest row of slice s, rounded up to 4); sell_row: the
// rows of the slice by decreasing length (ties by index), -1 for lanes without
// a row. One wavefront per slice; a lane's rank is found by comparing its key
// with every other lane's.
__global__ __launch_bounds__(256) void k_sell_width(int64_t n_rows, const int64_t* __restrict__ row_ptr,
                                                   int32_t* __restrict__ counts, int32_t* __restrict__ sell_row,
                                                   int32_t* __restrict__ overflow) {
    const int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t n_slices = (n_rows + 63) / 64;
    if (s >= n_slices) return;  // whole wavefront
    const int64_t row = s * 64 + lane;
    const int len = row < n_rows ? (int)(row_ptr[row + 1] - row_ptr[row]) : -1;
    int rank = 0;
    for (int m = 0; m < 64; ++m) {
        const int lm = __shfl(len, m, kWave);
        rank += (lm > len) || (lm == len && m < lane);
    }
    sell_row[s * 64 + rank] = row < n_rows ? (int32_t)row : -1;
    int mx = len;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off, kWave));
    mx = max(mx, 0);
    if (lane == 0) {
        if (mx > DCPSE_MAX_ROW) atomicOr(overflow, 1);
        counts[s] = 64 * ((mx + 3) & ~3);
    }
}

// Thread layout of the layout builds: a block of 256 threads walks slices
// (grid-stride); thread t takes lane t % 64 of the slice and its slot groups
// t / 64, t / 64 + 4, ... (4 slots each), so the loads of a row's entries are
// independent and every store is a 16-byte (or 8-byte) vector store at
// consecutive addresses across the lanes.
struct SliceLane {
    int32_t row;
    int64_t rs;
    int len;
};

__device__ __forceinline__ SliceLane slice_lane(const int64_t* row_ptr, const int32_t* sell_row, int64_t i) {
    SliceLane r;
    r.row = sell_row[i];
    r.rs = r.row >= 0 ? row_ptr[r.row] : 0;
    r.len = r.row >= 0 ? (int)(row_ptr[r.row + 1] - r.rs) : 0;
    return r;
}

// Every slot is written, padding included, also for the lanes of the last
// slice that have no row, so no slot of the layout is ever left
// uninitialised.
template <bool COLS, typename TW = float>
__global__ __launch_bounds__(256) void k_sell_fill(int64_t n_rows, const int64_t* __restrict__ row_ptr,
                                                   const int32_t* __restrict__ col, const TW* __restrict__ W,
                                                   const int64_t* __restrict__ slice_ptr,
                                                   const int32_t* __restrict__ sell_row,
                                                   int32_t* __restrict__ sell_col, TW* __restrict__ sell_W) {
    typedef TW vw4 __attribute__((ext_vector_type(4)));
    const int64_t n_slices = (n_rows + 63) / 64;
    const int l = threadIdx.x & 63, gsub = threadIdx.x >> 6;
    for (int64_t s = blockIdx.x; s < n_slices; s += gridDim.x) {
        const SliceLane r = slice_lane(row_ptr, sell_row, s * 64 + l);
        const int64_t base = slice_ptr[s];
        const int ng = (int)((slice_ptr[s + 1] - base) >> 8);
        for (int g = gsub; g < ng; g += 4) {
            const int64_t dst = base + (int64_t)g * 256 + 4 * l;
            if constexpr (COLS) {
                vi4a v;
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) v[jj] = 4 * g + jj < r.len ? col[r.rs + 4 * g + jj] : 0;
                *reinterpret_cast<vi4a*>(sell_col + dst) = v;
            } else {
                vw4 v;
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) v[jj] = 4 * g + jj < r.len ? W[r.rs + 4 * g + jj] : TW(0);
                *reinterpret_cast<vw4*>(sell_W + dst) = v;
            }
        }
    }
}

// Interleaved ("bundle") form of the layout for n_ops operators sharing one
// support: slot group g of slice s is one block of (1 + n_ops) * 256 words,
// [columns of the 64 lanes][W_0 of the 64 lanes]...[W_{n_ops-1}], starting at
// (1 + n_ops) * slice_ptr[s] (weights stored as their float bits). A
// wavefront then reads one contiguous (1 + n_ops) KB span per slot group
// instead of 1 KB from each of 1 + n_ops arrays. Every word is written.
__global__ __launch_bounds__(256) void k_sell_fill_bundle(int64_t n_rows, const int64_t* __restrict__ row_ptr,
                                                          const int32_t* __restrict__ col, int n_ops,
                                                          const float* W0, const float* W1, const float* W2,
                                                          const float* W3, const int64_t* __restrict__ slice_ptr,
                                                          const int32_t* __restrict__ sell_row,
                                                          int32_t* __restrict__ bundle) {
    const int64_t n_slices = (n_rows + 63) / 64;
    const float* W[4] = {W0, W1, W2, W3};
    const int l = threadIdx.x & 63, gsub = threadIdx.x >> 6;
    for (int64_t s = blockIdx.x; s < n_slices; s += gridDim.x) {
        const SliceLane r = slice_lane(row_ptr, sell_row, s * 64 + l);
        const int64_t base = slice_ptr[s];
        const int ng = (int)((slice_ptr[s + 1] - base) >> 8);
        for (int g = gsub; g < ng; g += 4) {
            const int64_t dst = (int64_t)(1 + n_ops) * (base + (int64_t)g * 256) + 4 * l;
            vi4a v;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) v[jj] = 4 * g + jj < r.len ? col[r.rs + 4 * g + jj] : 0;
            *reinterpret_cast<vi4a*>(bundle + dst) = v;
            for (int o = 0; o < n_ops; ++o) {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj)
                    v[jj] = 4 * g + jj < r.len ? __float_as_int(W[o][r.rs + 4 * g + jj]) : 0;
                *reinterpret_cast<vi4a*>(bundle + dst + (1 + o) * 256) = v;
            }
        }
    }
}

// The same interleaved layout with 16-bit columns, for operators whose support
// and domain are one set of n particles: a column q of row p is stored as the
// offset d = q - p taken modulo n into [-n/2, n/2). Slot group g of slice s is
// one block of 128 + 256 * n_ops words - the 64 lanes' 4 offsets (8 bytes per
// lane), then each operator's weights - starting at
// (128 + 256 * n_ops) * slice_ptr[s] / 256. Padding and lanes without a row
// hold offset 0.
// A neighbour further than 32767 positions (in a lattice-ordered periodic set,
// the rare pairs whose offset combines a wrap with plane steps) is stored as
// offset 0 - so in the gather it contributes W * (f_p - f_p) = 0 - and kept in
// a per-lane exception list (exc_ptr over lanes, exc_col, exc_W[o * n_exc + k])
// that the apply adds after the slot loop; lanes with exceptions are marked
// in sell_row16 by kExcFlag.
constexpr int32_t kExcFlag = 1 << 30;

__device__ __forceinline__ int64_t col_offset(int32_t q, int32_t p, int64_t n) {
    int64_t d = (int64_t)q - p;
    if (2 * d >= n) d -= n;
    if (2 * d < -n) d += n;
    return d;
}

__device__ __forceinline__ bool far16(int64_t d) { return d > 32767 || d < -32768; }

// exc_count[i] = far neighbours of the row of lane i: the four threads of a
// lane count their slot groups and the first adds the four counts.
__global__ __launch_bounds__(256) void k_bundle16_count(int64_t n_rows, const int64_t* __restrict__ row_ptr,
                                                        const int32_t* __restrict__ col,
                                                        const int32_t* __restrict__ sell_row,
                                                        int32_t* __restrict__ exc_count) {
    __shared__ int cnt[256];
    const int64_t n_slices = (n_rows + 63) / 64;
    const int l = threadIdx.x & 63, gsub = threadIdx.x >> 6;
    for (int64_t s = blockIdx.x; s < n_slices; s += gridDim.x) {
        const SliceLane r = slice_lane(row_ptr, sell_row, s * 64 + l);
        int c = 0;
        for (int g = gsub; 4 * g < r.len; g += 4)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj)
                if (4 * g + jj < r.len) c += far16(col_offset(col[r.rs + 4 * g + jj], r.row, n_rows));
        cnt[threadIdx.x] = c;
        __syncthreads();
        if (gsub == 0) exc_count[s * 64 + l] = cnt[l] + cnt[64 + l] + cnt[128 + l] + cnt[192 + l];
        __syncthreads();
    }
}

// The 16-bit bundle, one block per slice, through an LDS tile of 32 slots x
// 64 lanes: the four waves read 32 entries of two rows at a time (columns and
// each operator's weights, coalesced along the row) into the tile, then each
// thread writes its lane's slot groups of the tile to the bundle, 16-bit
// column offsets (0 where the offset does not fit: the lane's exception list
// holds that neighbour) and the weights, padding slots zero.
constexpr int kB16Chunk = 32;  // slots per tile
constexpr int kB16Pitch = 65;  // tile row pitch in words (no bank conflicts)
__global__ __launch_bounds__(256) void k_sell_fill_bundle16(
    int64_t n_rows, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col, int n_ops,
    const float* W0, const float* W1, const float* W2, const float* W3, const int64_t* __restrict__ slice_ptr,
    const int32_t* __restrict__ sell_row, const int64_t* __restrict__ exc_ptr, int32_t* __restrict__ bundle,
    int32_t* __restrict__ sell_row16) {
    __shared__ int32_t tc[kB16Chunk * kB16Pitch];      // column offsets
    __shared__ float tw[4][kB16Chunk * kB16Pitch];     // weights per operator
    __shared__ int32_t srow[64];
    __shared__ int64_t srs[64];
    __shared__ int32_t slen[64];
    const int64_t n_slices = (n_rows + 63) / 64;
    const float* W[4] = {W0, W1, W2, W3};
    const int64_t blk = 128 + 256 * (int64_t)n_ops;
    int16_t* cols16 = reinterpret_cast<int16_t*>(bundle);
    const int tid = threadIdx.x;
    const int l = tid & 63, gsub = tid >> 6;
    for (int64_t s = blockIdx.x; s < n_slices; s += gridDim.x) {
        __syncthreads();  // the previous slice's tile and lane table are consumed
        if (tid < 64) {
            const int64_t i = s * 64 + tid;
            const SliceLane r = slice_lane(row_ptr, sell_row, i);
            srow[tid] = r.row;
            srs[tid] = r.rs;
            slen[tid] = r.len;
            sell_row16[i] = r.row < 0 ? r.row : (exc_ptr[i + 1] > exc_ptr[i] ? (r.row | kExcFlag) : r.row);
        }
        __syncthreads();
        const int64_t base = slice_ptr[s];
        const int ng = (int)((slice_ptr[s + 1] - base) >> 8);
        for (int g0 = 0; g0 < ng; g0 += kB16Chunk / 4) {
            // load: wave w takes lanes 16 w .. 16 w + 15, two rows at a time
            const int e = tid & 31;
            const int j = 4 * g0 + e;
            for (int it = 0; it < 8; ++it) {
                const int ll = 16 * gsub + 2 * it + ((tid >> 5) & 1);
                const int len = slen[ll];
                const int64_t rs = srs[ll];
                int32_t d = 0;
                if (j < len) {
                    const int64_t dj = col_offset(col[rs + j], srow[ll], n_rows);
                    d = far16(dj) ? 0 : (int32_t)dj;
                }
                tc[e * kB16Pitch + ll] = d;
                for (int o = 0; o < n_ops; ++o) tw[o][e * kB16Pitch + ll] = j < len ? W[o][rs + j] : 0.0f;
            }
            __syncthreads();
            // write: thread (lane l, gsub) writes the tile's groups gsub, gsub + 4
            const int gn = min(kB16Chunk / 4, ng - g0);
            for (int gl = gsub; gl < gn; gl += 4) {
                const int64_t gb = (base / 256 + g0 + gl) * blk;  // word offset of the slot group's block
                vs4a dv;
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) dv[jj] = (int16_t)tc[(4 * gl + jj) * kB16Pitch + l];
                *reinterpret_cast<vs4a*>(cols16 + 2 * gb + 4 * l) = dv;
                for (int o = 0; o < n_ops; ++o) {
                    vi4a v;
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) v[jj] = __float_as_int(tw[o][(4 * gl + jj) * kB16Pitch + l]);
                    *reinterpret_cast<vi4a*>(bundle + gb + 128 + o * 256 + 4 * l) = v;
                }
            }
            __syncthreads();
        }
    }
}

// The exception lists: one thread per lane whose row has far neighbours
// walks the row in order.
template <typename WT>
__global__ void k_bundle16_exc(int64_t n_rows, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                               int n_ops, const WT* W0, const WT* W1, const WT* W2, const WT* W3,
                               const int32_t* __restrict__ sell_row, const int64_t* __restrict__ exc_ptr,
                               int64_t n_exc, int32_t* __restrict__ exc_col, WT* __restrict__ exc_W) {
    const int64_t n_lanes = ((n_rows + 63) / 64) * 64;
    const WT* W[4] = {W0, W1, W2, W3};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_lanes;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t e = exc_ptr[i];
        if (exc_ptr[i + 1] == e) continue;
        const SliceLane r = slice_lane(row_ptr, sell_row, i);
        for (int j = 0; j < r.len; ++j) {
            const int32_t q = col[r.rs + j];
            if (far16(col_offset(q, r.row, n_rows))) {
                exc_col[e] = q;
                for (int o = 0; o < n_ops; ++o) exc_W[o * n_exc + e] = W[o][r.rs + j];
                ++e;
            }
        }
    }
}

// The FP64 16-bit bundle: slot group g of slice s is one block of
// 128 + 512 * n_ops words (the 64 lanes' four 16-bit offsets, then each
// operator's four FP64 weights per lane, 32 bytes) at word offset
// (128 + 512 n_ops) * slice_ptr[s] / 256 + g (128 + 512 n_ops). One block per
// slice, through LDS tiles of 32 slots x 64 lanes as in k_sell_fill_bundle16:
// the column offsets, then each operator's weights in turn, read coalesced
// along the rows and written 32 contiguous bytes per lane.
__global__ __launch_bounds__(256) void k_sell_fill_bundle16_f64(
    int64_t n_rows, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col, int n_ops,
    const double* W0, const double* W1, const double* W2, const double* W3, const int64_t* __restrict__ slice_ptr,
    const int32_t* __restrict__ sell_row, const int64_t* __restrict__ exc_ptr, int32_t* __restrict__ bundle,
    int32_t* __restrict__ sell_row16) {
    __shared__ int32_t tc[kB16Chunk * kB16Pitch];  // column offsets
    __shared__ double tw[kB16Chunk * kB16Pitch];   // one operator's weights
    __shared__ int32_t srow[64];
    __shared__ int64_t srs[64];
    __shared__ int32_t slen[64];
    const int64_t n_slices = (n_rows + 63) / 64;
    const double* W[4] = {W0, W1, W2, W3};
    const int64_t blk = 128 + 512 * (int64_t)n_ops;
    int16_t* cols16 = reinterpret_cast<int16_t*>(bundle);
    const int tid = threadIdx.x;
    const int l = tid & 63, gsub = tid >> 6, e = tid & 31;
    for (int64_t s = blockIdx.x; s < n_slices; s += gridDim.x) {
        __syncthreads();  // the previous slice's tiles and lane table are consumed
        if (tid < 64) {
            const int64_t i = s * 64 + tid;
            const SliceLane r = slice_lane(row_ptr, sell_row, i);
            srow[tid] = r.row;
            srs[tid] = r.rs;
            slen[tid] = r.len;
            sell_row16[i] = r.row < 0 ? r.row : (exc_ptr[i + 1] > exc_ptr[i] ? (r.row | kExcFlag) : r.row);
        }
        __syncthreads();
        const int64_t base = slice_ptr[s];
        const int ng = (int)((slice_ptr[s + 1] - base) >> 8);
        for (int g0 = 0; g0 < ng; g0 += kB16Chunk / 4) {
            const int j = 4 * g0 + e;
            const int gn = min(kB16Chunk / 4, ng - g0);
            for (int o = -1; o < n_ops; ++o) {  // -1: the column offsets
                // load: wave w takes lanes 16 w .. 16 w + 15, two rows at a time
                for (int it = 0; it < 8; ++it) {
                    const int ll = 16 * gsub + 2 * it + ((tid >> 5) & 1);
                    const int len = slen[ll];
                    const int64_t rs = srs[ll];
                    if (o < 0) {
                        int32_t d = 0;
                        if (j < len) {
                            const int64_t dj = col_offset(col[rs + j], srow[ll], n_rows);
                            d = far16(dj) ? 0 : (int32_t)dj;
                        }
                        tc[e * kB16Pitch + ll] = d;
                    } else {
                        tw[e * kB16Pitch + ll] = j < len ? W[o][rs + j] : 0.0;
                    }
                }
                __syncthreads();
                // write: thread (lane l, gsub) writes the tile's groups gsub, gsub + 4
                for (int gl = gsub; gl < gn; gl += 4) {
                    const int64_t gb = (base / 256 + g0 + gl) * blk;  // word offset of the slot group's block
                    if (o < 0) {
                        vs4a dv;
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj) dv[jj] = (int16_t)tc[(4 * gl + jj) * kB16Pitch + l];
                        *reinterpret_cast<vs4a*>(cols16 + 2 * gb + 4 * l) = dv;
                    } else {
                        double2* wb = reinterpret_cast<double2*>(bundle + gb + 128 + o * 512) + 2 * l;
                        wb[0] = make_double2(tw[(4 * gl) * kB16Pitch + l], tw[(4 * gl + 1) * kB16Pitch + l]);
                        wb[1] = make_double2(tw[(4 * gl + 2) * kB16Pitch + l], tw[(4 * gl + 3) * kB16Pitch + l]);
                    }
                }
                __syncthreads();
            }
        }
    }
}

struct SellArgs {
    int64_t n_rows;
    const int64_t* row_ptr;
    const int64_t* slice_ptr;
    const int32_t* sell_row;
    const int32_t* sell_col;
    const int32_t* bundle;  // interleaved layout (see k_sell_fill_bundle), or null
    const int64_t* exc_ptr;  // 16-bit bundle: far neighbours per lane
    const int32_t* exc_col;
    const float* exc_W;
    int64_t n_exc;
    const float* W[4];
    const float* f;
    const float* f_self;
    const float* diag;
    float* out[4];
    float alpha, beta;
    // MODE 2 (explicit time-integration stage, one operator): k = alpha * D f,
    // out_a = y_a + ca * k and out_b = y_b + cb * k (each skipped when null)
    const float* y_a;
    const float* y_b;
    float* out_a;
    float* out_b;
    float ca, cb;
    int64_t blocks_needed, per_xcd;
};

// Row epilogue shared by the SELL kernels: MODE 0 forward, 1 transposed
// (explicit diagonal), 2 time-integration stage.
// The row sums are accumulated in FP64 by default (the FP32 products are
// exact in FP64): at high resolution the difference-form terms cancel to
// first order, and FP32 accumulation rounding, not the FP32 weights,
// dominates the error. The forward kernels can also accumulate in FP32
// (ACC32, dcpse_set_apply_accumulation), differences included.
template <int NOPS, int MODE>
__device__ __forceinline__ void sell_epilogue(const SellArgs& a, int32_t row, float fp, const float* acc) {
    static_assert(MODE == 0, "FP32 accumulation is a forward-apply option");
#pragma unroll
    for (int o = 0; o < NOPS; ++o) {
        float val = acc[o] * a.alpha;
        if (a.beta != 0.0f) val = fmaf(a.beta, a.out[o][row], val);
        a.out[o][row] = val;
    }
}

template <int NOPS, int MODE>
__device__ __forceinline__ void sell_epilogue(const SellArgs& a, int32_t row, float fp, const double* acc) {
    if (MODE == 2) {  // both inputs are read before either output is written
        const double k = (double)a.alpha * acc[0];
        const double ya = a.out_a ? (double)a.y_a[row] : 0.0;
        const double yb = a.out_b ? (double)a.y_b[row] : 0.0;
        if (a.out_a) a.out_a[row] = (float)fma((double)a.ca, k, ya);
        if (a.out_b) a.out_b[row] = (float)fma((double)a.cb, k, yb);
        return;
    }
#pragma unroll
    for (int o = 0; o < NOPS; ++o) {
        double val = acc[o];
        if (MODE == 1) val = fma((double)a.diag[row], (double)fp, val);
        val *= (double)a.alpha;
        if (a.beta != 0.0f) val = fma((double)a.beta, (double)a.out[o][row], val);
        a.out[o][row] = (float)val;
    }
}

template <int NOPS, int UNR, bool C16 = false>
struct SellBatch {
    typename std::conditional<C16, vs4a, vi4a>::type q[UNR];  // columns, or 16-bit offsets
    vf4a w[NOPS][UNR];
};

// Groups [g0, g0+UNR) of the lane's row: the 16-byte col/W loads. Groups past
// the row (or the slice) re-read group 0, so every load is issued
// unconditionally and the wait counts stay exact; sell_fma masks them out.
// Every slot a lane can read is written by k_sell_fill: padding and the lanes
// without a row hold column 0, so every gather stays inside f. The loaded
// columns are used as they are (no select on them), so nothing waits for a
// batch before its gathers are due.
template <int NOPS, int UNR, bool C16 = false>
__device__ __forceinline__ void sell_load(const int32_t* cp, const float* const* wp, int64_t gs, int g0, int len,
                                          SellBatch<NOPS, UNR, C16>& b) {
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
        const int g = g0 + u;
        const int64_t off = 4 * g < len ? (int64_t)g * gs : 0;
        if constexpr (C16)  // cp points at this lane's 8 bytes of offsets
            b.q[u] = *(const vs4a*)(reinterpret_cast<const int16_t*>(cp) + 2 * off);
        else
            b.q[u] = *(const vi4a*)(cp + off);
#pragma unroll
        for (int o = 0; o < NOPS; ++o) b.w[o][u] = *(const vf4a*)(wp[o] + off);
    }
}

// Where lane `lane` of slice s finds its columns and weights, and the element
// stride between its consecutive slot groups: separate arrays (stride 256) or
// the interleaved bundle (stride (1 + NOPS) * 256).
template <int NOPS, bool C16 = false>
__device__ __forceinline__ int64_t sell_streams(const SellArgs& a, int64_t base, int lane, const int32_t*& cp,
                                                const float* (&wp)[NOPS]) {
    if constexpr (C16) {
        const int64_t blk = 128 + 256 * NOPS;
        const int32_t* bp = a.bundle + (base / 256) * blk;
        cp = bp + 2 * lane;  // 8 bytes of offsets per lane
#pragma unroll
        for (int o = 0; o < NOPS; ++o) wp[o] = reinterpret_cast<const float*>(bp + 128 + o * 256) + 4 * lane;
        return blk;
    }
    if (a.bundle) {
        const int32_t* bp = a.bundle + (int64_t)(1 + NOPS) * base + 4 * lane;
        cp = bp;
#pragma unroll
        for (int o = 0; o < NOPS; ++o) wp[o] = reinterpret_cast<const float*>(bp + (1 + o) * 256);
        return (int64_t)(1 + NOPS) * 256;
    }
    cp = a.sell_col + base + 4 * lane;
#pragma unroll
    for (int o = 0; o < NOPS; ++o) wp[o] = a.W[o] + base + 4 * lane;
    return 256;
}

// The f gathers of groups [g0, g0+UNR) (columns of a loaded batch).
template <int UNR>
struct SellGathered {
    float fq[UNR][4];
};

// Column of a 16-bit offset d of row rb in a set of n particles.
__device__ __forceinline__ int32_t sell_col16(int32_t rb, int16_t d, int32_t n) {
    int32_t q = rb + (int32_t)d;
    q += q < 0 ? n : 0;
    q -= q >= n ? n : 0;
    return q;
}

template <int NOPS, int UNR, bool C16 = false>
__device__ __forceinline__ void sell_gather(const SellArgs& a, const SellBatch<NOPS, UNR, C16>& b,
                                            SellGathered<UNR>& g, int32_t rb = 0) {
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
        if constexpr (C16) {
            const int32_t n = (int32_t)a.n_rows;
            g.fq[u][0] = a.f[sell_col16(rb, b.q[u].x, n)];
            g.fq[u][1] = a.f[sell_col16(rb, b.q[u].y, n)];
            g.fq[u][2] = a.f[sell_col16(rb, b.q[u].z, n)];
            g.fq[u][3] = a.f[sell_col16(rb, b.q[u].w, n)];
        } else {
            g.fq[u][0] = a.f[b.q[u].x];
            g.fq[u][1] = a.f[b.q[u].y];
            g.fq[u][2] = a.f[b.q[u].z];
            g.fq[u][3] = a.f[b.q[u].w];
        }
    }
}

// The FMAs of groups [g0, g0+UNR); slots past the row are masked. Acc is
// double (default) or float (every operation in FP32).
template <int NOPS, int UNR, int MODE, bool C16, typename Acc>
__device__ __forceinline__ void sell_fma(int g0, int len, float fp, const SellBatch<NOPS, UNR, C16>& b,
                                         const SellGathered<UNR>& g, Acc* acc) {
#pragma unroll
    for (int u = 0; u < UNR; ++u)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const bool v = 4 * (g0 + u) + i < len;
            const Acc d = MODE == 1 ? (Acc)g.fq[u][i] : (Acc)g.fq[u][i] - (Acc)fp;
#pragma unroll
            for (int o = 0; o < NOPS; ++o) acc[o] = v ? fma((Acc)b.w[o][u][i], d, acc[o]) : acc[o];
        }
}

// One wavefront per slice. Two register batches alternate. Issue order per
// half step: the gathers of the loaded batch, then the stream loads of the
// other batch, then the FMAs. Vector-memory loads complete in issue order, so
// waiting for the gathers does not wait for the stream loads issued after
// them: the next batch's col/W loads stay in flight across the FMAs, and its
// gathers go out as soon as its columns arrive.
template <int NOPS, int UNR, int MODE, bool C16 = false, bool ACC32 = false>
__global__ __launch_bounds__(kSellWaves * kWave) void k_apply_sell(const SellArgs a) {
    using Acc = typename std::conditional<ACC32, float, double>::type;
    const int64_t hb = blockIdx.x;
    const int64_t lb = (hb & 7) * a.per_xcd + (hb >> 3);
    if (lb >= a.blocks_needed) return;
    const int lane = threadIdx.x & 63;
    const int64_t s = lb * kSellWaves + (threadIdx.x >> 6);
    if (s * 64 >= a.n_rows) return;  // whole wavefront
    const int32_t row_raw = a.sell_row[s * 64 + lane];
    const bool live = row_raw >= 0;
    const bool has_exc = C16 && live && (row_raw & kExcFlag);
    const int32_t row = C16 && live ? (row_raw & ~kExcFlag) : row_raw;
    const int len = live ? (int)(a.row_ptr[row + 1] - a.row_ptr[row]) : 0;
    const float fp = live ? a.f_self[row] : 0.0f;
    const int64_t base = a.slice_ptr[s];
    const int ng = (int)((a.slice_ptr[s + 1] - base) >> 8);  // groups of 4 slots
    const int32_t* cp;
    const float* wp[NOPS];
    const int64_t gs = sell_streams<NOPS, C16>(a, base, lane, cp, wp);
    const int32_t rb = live ? row : 0;  // 16-bit offsets are relative to the lane's row
    Acc acc[NOPS];
#pragma unroll
    for (int o = 0; o < NOPS; ++o) acc[o] = 0;
    if (ng > 0) {  // a slice of empty rows has no slots to read
        SellBatch<NOPS, UNR, C16> bA, bB;
        SellGathered<UNR> gA, gB;
        sell_load<NOPS, UNR, C16>(cp, wp, gs, 0, len, bA);
        for (int g0 = 0; g0 < ng; g0 += 2 * UNR) {
            sell_gather<NOPS, UNR, C16>(a, bA, gA, rb);
            __builtin_amdgcn_sched_barrier(0);
            sell_load<NOPS, UNR, C16>(cp, wp, gs, g0 + UNR, len, bB);
            __builtin_amdgcn_sched_barrier(0);
            sell_fma<NOPS, UNR, MODE, C16, Acc>(g0, len, fp, bA, gA, acc);
            __builtin_amdgcn_sched_barrier(0);
            sell_gather<NOPS, UNR, C16>(a, bB, gB, rb);
            __builtin_amdgcn_sched_barrier(0);
            sell_load<NOPS, UNR, C16>(cp, wp, gs, g0 + 2 * UNR, len, bA);
            __builtin_amdgcn_sched_barrier(0);
            sell_fma<NOPS, UNR, MODE, C16, Acc>(g0 + UNR, len, fp, bB, gB, acc);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    if constexpr (C16) {
        if (has_exc) {  // the far neighbours of this row (rare; after every pipelined load)
            const int64_t i = s * 64 + lane;
            for (int64_t k = a.exc_ptr[i]; k < a.exc_ptr[i + 1]; ++k) {
                const Acc d = MODE == 1 ? (Acc)a.f[a.exc_col[k]] : (Acc)a.f[a.exc_col[k]] - (Acc)fp;
#pragma unroll
                for (int o = 0; o < NOPS; ++o) acc[o] = fma((Acc)a.exc_W[o * a.n_exc + k], d, acc[o]);
            }
        }
    }
    if (live) sell_epilogue<NOPS, MODE>(a, row, fp, acc);
}

// FP64 forward apply from the sliced-ELL layout: the columns of the FP32
// layout (sell_col) and FP64 weights in the same slot order. The pipeline is
// that of k_apply_sell (two register batches; the gathers of the loaded batch
// before the next batch's stream loads); a slot group is 16 bytes of columns
// and 32 bytes of weights per operator and lane.
typedef double vd4a __attribute__((ext_vector_type(4)));

struct Sell64Args {
    int64_t n_rows;
    const int64_t* row_ptr;
    const int64_t* slice_ptr;
    const int32_t* sell_row;  // 16-bit bundle: sell_row16 (bit 30: the lane has far neighbours)
    const int32_t* sell_col;
    const double* W[4];
    const int32_t* bundle;    // 16-bit bundle (k_sell_fill_bundle16_f64), or null
    const int64_t* exc_ptr;   // 16-bit bundle: far neighbours per lane
    const int32_t* exc_col;
    const double* exc_W;
    int64_t n_exc;
    const double* f;
    const double* f_self;
    double* out[4];
    double alpha, beta;
    int64_t blocks_needed, per_xcd;
};

template <int NOPS, int UNR, bool C16 = false>
struct Sell64Batch {
    typename std::conditional<C16, vs4a, vi4a>::type q[UNR];  // columns, or 16-bit offsets
    vd4a w[NOPS][UNR];
};

// gs: element stride between a lane's consecutive slot groups (the columns'
// unit: int32 for the separate arrays, int16 for the bundle; weights double)
template <int NOPS, int UNR, bool C16 = false>
__device__ __forceinline__ void sell64_load(const void* cp, const double* const* wp, int64_t gsc, int64_t gsw,
                                            int g0, int len, Sell64Batch<NOPS, UNR, C16>& b) {
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
        const int g = g0 + u;
        const int64_t gg = 4 * g < len ? (int64_t)g : 0;  // past the row: group 0 again
        if constexpr (C16)
            b.q[u] = *(const vs4a*)(reinterpret_cast<const int16_t*>(cp) + gg * gsc);
        else
            b.q[u] = *(const vi4a*)(reinterpret_cast<const int32_t*>(cp) + gg * gsc);
#pragma unroll
        for (int o = 0; o < NOPS; ++o) b.w[o][u] = *(const vd4a*)(wp[o] + gg * gsw);
    }
}

template <int NOPS, int UNR, bool C16 = false>
__device__ __forceinline__ void sell64_gather(const Sell64Args& a, const Sell64Batch<NOPS, UNR, C16>& b,
                                              double (&fq)[UNR][4], int32_t rb) {
#pragma unroll
    for (int u = 0; u < UNR; ++u)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if constexpr (C16) fq[u][i] = a.f[sell_col16(rb, b.q[u][i], (int32_t)a.n_rows)];
            else fq[u][i] = a.f[b.q[u][i]];
        }
}

template <int NOPS, int UNR, bool C16 = false>
__device__ __forceinline__ void sell64_fma(int g0, int len, double fp, const Sell64Batch<NOPS, UNR, C16>& b,
                                           const double (&fq)[UNR][4], double* acc) {
#pragma unroll
    for (int u = 0; u < UNR; ++u)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const bool v = 4 * (g0 + u) + i < len;
            const double d = fq[u][i] - fp;
#pragma unroll
            for (int o = 0; o < NOPS; ++o) acc[o] = v ? fma(b.w[o][u][i], d, acc[o]) : acc[o];
        }
}

// C16: the 16-bit bundle (k_sell_fill_bundle16_f64); offsets are relative to
// the lane's row, and a lane flagged in sell_row16 adds its far neighbours
// from the exception list after the slot loop.
template <int NOPS, int UNR, bool C16 = false>
__global__ __launch_bounds__(kSellWaves * kWave) void k_apply_sell64(const Sell64Args a) {
    const int64_t hb = blockIdx.x;
    const int64_t lb = (hb & 7) * a.per_xcd + (hb >> 3);
    if (lb >= a.blocks_needed) return;
    const int lane = threadIdx.x & 63;
    const int64_t s = lb * kSellWaves + (threadIdx.x >> 6);
    if (s * 64 >= a.n_rows) return;  // whole wavefront
    const int32_t row_raw = a.sell_row[s * 64 + lane];
    const bool live = row_raw >= 0;
    const bool has_exc = C16 && live && (row_raw & kExcFlag);
    const int32_t row = C16 && live ? (row_raw & ~kExcFlag) : row_raw;
    const int len = live ? (int)(a.row_ptr[row + 1] - a.row_ptr[row]) : 0;
    const double fp = live ? a.f_self[row] : 0.0;
    const int64_t base = a.slice_ptr[s];
    const int ng = (int)((a.slice_ptr[s + 1] - base) >> 8);
    const void* cp;
    const double* wp[NOPS];
    int64_t gsc, gsw;
    if constexpr (C16) {
        const int64_t blk = 128 + 512 * NOPS;  // words per slot group
        const int32_t* bp = a.bundle + (base / 256) * blk;
        cp = reinterpret_cast<const int16_t*>(bp) + 4 * lane;
#pragma unroll
        for (int o = 0; o < NOPS; ++o) wp[o] = reinterpret_cast<const double*>(bp + 128 + o * 512) + 4 * lane;
        gsc = 2 * blk;  // int16 units
        gsw = blk / 2;  // double units
    } else {
        cp = a.sell_col + base + 4 * lane;
#pragma unroll
        for (int o = 0; o < NOPS; ++o) wp[o] = a.W[o] + base + 4 * lane;
        gsc = gsw = 256;
    }
    const int32_t rb = live ? row : 0;
    double acc[NOPS];
#pragma unroll
    for (int o = 0; o < NOPS; ++o) acc[o] = 0.0;
    if (ng > 0) {
        Sell64Batch<NOPS, UNR, C16> bA, bB;
        double gA[UNR][4], gB[UNR][4];
        sell64_load<NOPS, UNR, C16>(cp, wp, gsc, gsw, 0, len, bA);
        for (int g0 = 0; g0 < ng; g0 += 2 * UNR) {
            sell64_gather<NOPS, UNR, C16>(a, bA, gA, rb);
            __builtin_amdgcn_sched_barrier(0);
            sell64_load<NOPS, UNR, C16>(cp, wp, gsc, gsw, g0 + UNR, len, bB);
            __builtin_amdgcn_sched_barrier(0);
            sell64_fma<NOPS, UNR, C16>(g0, len, fp, bA, gA, acc);
            __builtin_amdgcn_sched_barrier(0);
            sell64_gather<NOPS, UNR, C16>(a, bB, gB, rb);
            __builtin_amdgcn_sched_barrier(0);
            sell64_load<NOPS, UNR, C16>(cp, wp, gsc, gsw, g0 + 2 * UNR, len, bA);
            __builtin_amdgcn_sched_barrier(0);
            sell64_fma<NOPS, UNR, C16>(g0 + UNR, len, fp, bB, gB, acc);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    if constexpr (C16) {
        if (has_exc) {  // the far neighbours of this row (rare; after every pipelined load)
            const int64_t i = s * 64 + lane;
            for (int64_t k = a.exc_ptr[i]; k < a.exc_ptr[i + 1]; ++k) {
                const double d = a.f[a.exc_col[k]] - fp;
#pragma unroll
                for (int o = 0; o < NOPS; ++o) acc[o] = fma(a.exc_W[o * a.n_exc + k], d, acc[o]);
            }
        }
    }
    if (!live) return;
#pragma unroll
    for (int o = 0; o < NOPS; ++o) {
        double val = acc[o] * a.alpha;
        if (a.beta != 0.0) val = fma(a.beta, a.out[o][row], val);
        a.out[o][row] = val;
    }
}

template <int NOPS, int UNR, bool C16 = false>
static void launch_sell64(Sell64Args a, hipStream_t st) {
    const int64_t n_slices = (a.n_rows + 63) / 64;
    a.blocks_needed = (n_slices + kSellWaves - 1) / kSellWaves;
    a.per_xcd = (a.blocks_needed + 7) / 8;
    k_apply_sell64<NOPS, UNR, C16><<<(unsigned)(a.per_xcd * 8), kSellWaves * kWave, 0, st>>>(a);
}

template <int UNR, bool C16 = false>
static bool dispatch_sell64(const Sell64Args& a, int nops, hipStream_t st) {
    switch (nops) {
        case 1: launch_sell64<1, UNR, C16>(a, st); return true;
        case 2: launch_sell64<2, UNR, C16>(a, st); return true;
        case 3: launch_sell64<3, UNR, C16>(a, st); return true;
        case 4: launch_sell64<4, UNR, C16>(a, st); return true;
    }
    return false;
}

template <int NOPS, int UNR, int MODE, bool ACC32 = false>
static void launch_sell(SellArgs a, hipStream_t st) {
    const int64_t n_slices = (a.n_rows + 63) / 64;
    a.blocks_needed = (n_slices + kSellWaves - 1) / kSellWaves;
    a.per_xcd = (a.blocks_needed + 7) / 8;
    k_apply_sell<NOPS, UNR, MODE, false, ACC32><<<(unsigned)(a.per_xcd * 8), kSellWaves * kWave, 0, st>>>(a);
}

template <int NOPS, int UNR, int MODE = 0, bool ACC32 = false>
static void launch_sell16(SellArgs a, hipStream_t st) {
    const int64_t n_slices = (a.n_rows + 63) / 64;
    a.blocks_needed = (n_slices + kSellWaves - 1) / kSellWaves;
    a.per_xcd = (a.blocks_needed + 7) / 8;
    k_apply_sell<NOPS, UNR, MODE, true, ACC32><<<(unsigned)(a.per_xcd * 8), kSellWaves * kWave, 0, st>>>(a);
}

// 16-bit-offset bundles: forward apply, two-stage pipeline (groups 2 or 4).
template <bool ACC32>
static bool dispatch_sell16_acc(const SellArgs& a, int nops, int unr, hipStream_t st) {
    const bool u4 = unr == 4;
    switch (nops) {
        case 1: u4 ? launch_sell16<1, 4, 0, ACC32>(a, st) : launch_sell16<1, 2, 0, ACC32>(a, st); return true;
        case 2: u4 ? launch_sell16<2, 4, 0, ACC32>(a, st) : launch_sell16<2, 2, 0, ACC32>(a, st); return true;
        case 3: u4 ? launch_sell16<3, 4, 0, ACC32>(a, st) : launch_sell16<3, 2, 0, ACC32>(a, st); return true;
        case 4: u4 ? launch_sell16<4, 4, 0, ACC32>(a, st) : launch_sell16<4, 2, 0, ACC32>(a, st); return true;
    }
    return false;
}

static bool dispatch_sell16(const SellArgs& a, int nops, int unr, bool acc32, hipStream_t st) {
    return acc32 ? dispatch_sell16_acc<true>(a, nops, unr, st) : dispatch_sell16_acc<false>(a, nops, unr, st);
}

template <int UNR, int MODE, bool ACC32 = false>
static bool dispatch_sell_ops(const SellArgs& a, int nops, hipStream_t st) {
    if constexpr (MODE == 2) {  // stages apply one operator
        if (nops != 1) return false;
        launch_sell<1, UNR, MODE>(a, st);
        return true;
    }
    switch (nops) {
        case 1: launch_sell<1, UNR, MODE, ACC32>(a, st); return true;
        case 2: launch_sell<2, UNR, MODE, ACC32>(a, st); return true;
        case 3: launch_sell<3, UNR, MODE, ACC32>(a, st); return true;
        case 4: launch_sell<4, UNR, MODE, ACC32>(a, st); return true;
    }
    return false;
}

template <int MODE, bool ACC32 = false>
static bool dispatch_sell(const SellArgs& a, int nops, int unr, hipStream_t st) {
    switch (unr) {
        case 2: return dispatch_sell_ops<2, MODE, ACC32>(a, nops, st);
        case 4: return dispatch_sell_ops<4, MODE, ACC32>(a, nops, st);
    }
    return false;
}

// Scratch of dcpse_sell_count_noalloc: the per-slice widths, the overflow
// flag and the scan's tile sums live in this file-scope device array, so that
// entry point allocates nothing. Each device holds its own copy for the life
// of the process; host code never passes the array's name to a kernel (on the
// host it is only a stand-in for the symbol) but the device address that
// hipGetSymbolAddress returns, looked up once per device by
// noalloc_scratch_init (called from dcpse_ctx_create with the device current).
// Every no-alloc call on a device shares that device's copy: calls that
// overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxScratchDevices] = {};

dcpse_status noalloc_scratch_init(dcpse_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxScratchDevices)
        return fail(ctx, DCPSE_ERR_RANGE, "device index %d beyond the no-alloc scratch table", ctx->device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* ptr = nullptr;
        DCPSE_HIP_CHECK(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(ctx, DCPSE_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_ptr[ctx->device] = ptr;
    }
    ctx->scratch = g_scratch_ptr[ctx->device];
    return DCPSE_OK;
}

// Bytes the slice-structure pass takes from its scratch: the same bump
// allocations as sell_count_impl, then the scan's share.
static size_t sell_count_scratch_bytes(int64_t n_slices) {
    Workspace probe;
    probe.cap = ~size_t(0) >> 1;
    (void)probe.take<int32_t>((size_t)n_slices + 1);
    (void)probe.take<int32_t>(16);
    return probe.used + 256 + scan_workspace_bytes(n_slices);
}

// Slice structure of the sliced-ELL layout with its scratch taken from w.
static dcpse_status sell_count_impl(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, int64_t* slice_ptr,
                                    int32_t* sell_row, int64_t* n_slots, Workspace& w) {
    const int64_t n_slices = (n_rows + 63) / 64;
    dcpse_status st;
    int32_t* counts = w.take<int32_t>(n_slices + 1);
    int32_t* flags = w.take<int32_t>(16);
    if (!flags) return fail(ctx, DCPSE_ERR_INVALID, "sliced-ELL count: scratch sizing error");
    fill_i32(ctx->stream, flags, 16, 0);
    if (n_slices > 0) {
        k_sell_width<<<(unsigned)((n_slices * 64 + 255) / 256), 256, 0, ctx->stream>>>(n_rows, row_ptr, counts,
                                                                                        sell_row, flags);
        DCPSE_LAUNCH_CHECK(ctx);
    }
    if ((st = scan_counts(ctx, counts, n_slices, slice_ptr, &w))) return st;
    int64_t total = 0;
    int32_t ovf = 0;
    DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(&total, slice_ptr + n_slices, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(&ovf, flags, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    DCPSE_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ovf) return fail(ctx, DCPSE_ERR_RANGE, "a row exceeds DCPSE_MAX_ROW=%d", DCPSE_MAX_ROW);
    if (n_slots) *n_slots = total;
    return DCPSE_OK;
}

}  // namespace dcpse

using namespace dcpse;

extern "C" {

dcpse_status dcpse_sell_count(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, int64_t* slice_ptr,
                              int32_t* sell_row, int64_t* n_slots) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !slice_ptr || !sell_row || n_rows < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int64_t n_slices = (n_rows + 63) / 64;
    Workspace w;
    dcpse_status st = reserve_workspace(ctx, sell_count_scratch_bytes(n_slices) + 1024, &w);
    if (st) return st;
    return sell_count_impl(ctx, n_rows, row_ptr, slice_ptr, sell_row, n_slots, w);
}

dcpse_status dcpse_sell_count_noalloc(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, int64_t* slice_ptr,
                                      int32_t* sell_row, int64_t* n_slots) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !slice_ptr || !sell_row || n_rows < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    if (n_rows > DCPSE_NOALLOC_MAX_ROWS)
        return fail(ctx, DCPSE_ERR_RANGE, "dcpse_sell_count_noalloc: %lld rows exceed DCPSE_NOALLOC_MAX_ROWS=%d",
                    (long long)n_rows, DCPSE_NOALLOC_MAX_ROWS);
    if (!ctx->scratch) return fail(ctx, DCPSE_ERR_HIP, "no-alloc scratch address was not set up for this device");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int64_t n_slices = (n_rows + 63) / 64;
    if (sell_count_scratch_bytes(n_slices) > kNoAllocScratchBytes)
        return fail(ctx, DCPSE_ERR_RANGE, "dcpse_sell_count_noalloc: scratch of %zu bytes too small",
                    kNoAllocScratchBytes);
    Workspace w;
    w.base = static_cast<char*>(ctx->scratch);
    w.cap = kNoAllocScratchBytes;
    return sell_count_impl(ctx, n_rows, row_ptr, slice_ptr, sell_row, n_slots, w);
}

dcpse_status dcpse_sell_fill_cols(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                                  const int64_t* slice_ptr, const int32_t* sell_row, int32_t* sell_col) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !col || !slice_ptr || !sell_row || !sell_col || n_rows < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0) return DCPSE_OK;
    const int64_t b = std::min<int64_t>((n_rows + 63) / 64, 65536);
    k_sell_fill<true, float><<<(unsigned)b, 256, 0, ctx->stream>>>(n_rows, row_ptr, col, nullptr, slice_ptr, sell_row,
                                                            sell_col, nullptr);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

dcpse_status dcpse_sell_fill_weights_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const float* W,
                                         const int64_t* slice_ptr, const int32_t* sell_row, float* sell_W) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !W || !slice_ptr || !sell_row || !sell_W || n_rows < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0) return DCPSE_OK;
    const int64_t b = std::min<int64_t>((n_rows + 63) / 64, 65536);
    k_sell_fill<false, float><<<(unsigned)b, 256, 0, ctx->stream>>>(n_rows, row_ptr, nullptr, W, slice_ptr, sell_row,
                                                             nullptr, sell_W);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

dcpse_status dcpse_sell_fill_weights_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const double* W,
                                         const int64_t* slice_ptr, const int32_t* sell_row, double* sell_W) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !W || !slice_ptr || !sell_row || !sell_W || n_rows < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0) return DCPSE_OK;
    const int64_t b = std::min<int64_t>((n_rows + 63) / 64, 65536);
    k_sell_fill<false, double><<<(unsigned)b, 256, 0, ctx->stream>>>(n_rows, row_ptr, nullptr, W, slice_ptr,
                                                                     sell_row, nullptr, sell_W);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

dcpse_status dcpse_apply_sell_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int64_t* slice_ptr,
                                  const int32_t* sell_row, const int32_t* sell_col, int32_t n_ops,
                                  const double* const* sell_W, const double* f, const double* f_self,
                                  double* const* out, double alpha, double beta) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !slice_ptr || !sell_row || !sell_col || !sell_W || !f || !out || n_rows < 0 || n_ops < 1)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    if (n_ops > 4) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "at most 4 operators per pass");
    for (int o = 0; o < n_ops; ++o)
        if (!sell_W[o] || !out[o]) return fail(ctx, DCPSE_ERR_INVALID, "null operator pointer");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0) return DCPSE_OK;
    Sell64Args a{};
    a.n_rows = n_rows;
    a.row_ptr = row_ptr;
    a.slice_ptr = slice_ptr;
    a.sell_row = sell_row;
    a.sell_col = sell_col;
    for (int o = 0; o < 4; ++o) {
        a.W[o] = o < n_ops ? sell_W[o] : nullptr;
        a.out[o] = o < n_ops ? out[o] : nullptr;
    }
    a.f = f;
    a.f_self = f_self ? f_self : f;
    a.alpha = alpha;
    a.beta = beta;
    // two slot groups per batch (437 us for the fused Laplacian + Dx at 128^3,
    // one group 445 us; the CSR kernel takes 512 us)
    if (!dispatch_sell64<2>(a, n_ops, ctx->stream)) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "no SELL kernel for %d operators", n_ops);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

dcpse_status dcpse_sell_fill_bundle16_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                          const int32_t* col, int32_t n_ops, const double* const* W,
                                          const int64_t* slice_ptr, const int32_t* sell_row,
                                          const int64_t* exc_ptr, int64_t n_exc, int32_t* bundle,
                                          int32_t* sell_row16, int32_t* exc_col, double* exc_W) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !col || !W || !slice_ptr || !sell_row || !exc_ptr || !bundle || !sell_row16 || n_rows < 0 ||
        n_exc < 0 || (n_exc > 0 && (!exc_col || !exc_W)))
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    if (n_ops < 1 || n_ops > 4) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "1 to 4 operators per bundle");
    for (int o = 0; o < n_ops; ++o)
        if (!W[o]) return fail(ctx, DCPSE_ERR_INVALID, "null weight array");
    if (n_rows >= ((int64_t)1 << 30)) return fail(ctx, DCPSE_ERR_RANGE, "16-bit bundles need n_rows < 2^30");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0) return DCPSE_OK;
    const int64_t n_slices = (n_rows + 63) / 64;
    const int64_t b = std::min<int64_t>(n_slices, (int64_t)ctx->num_cus * 16);
    k_sell_fill_bundle16_f64<<<(unsigned)b, 256, 0, ctx->stream>>>(
        n_rows, row_ptr, col, n_ops, W[0], n_ops > 1 ? W[1] : nullptr, n_ops > 2 ? W[2] : nullptr,
        n_ops > 3 ? W[3] : nullptr, slice_ptr, sell_row, exc_ptr, bundle, sell_row16);
    DCPSE_LAUNCH_CHECK(ctx);
    if (n_exc > 0) {
        const int64_t be = std::min<int64_t>((n_rows + 63 + 255) / 256, 65536);
        k_bundle16_exc<double><<<(unsigned)be, 256, 0, ctx->stream>>>(
            n_rows, row_ptr, col, n_ops, W[0], n_ops > 1 ? W[1] : nullptr, n_ops > 2 ? W[2] : nullptr,
            n_ops > 3 ? W[3] : nullptr, sell_row, exc_ptr, n_exc, exc_col, exc_W);
    }
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

dcpse_status dcpse_apply_sell_bundle16_f64(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                           const int64_t* slice_ptr, const int32_t* sell_row16, int32_t n_ops,
                                           const int32_t* bundle, const int64_t* exc_ptr, const int32_t* exc_col,
                                           const double* exc_W, int64_t n_exc, const double* f,
                                           const double* f_self, double* const* out, double alpha, double beta) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !slice_ptr || !sell_row16 || !bundle || !exc_ptr || !f || !out || n_rows < 0 || n_ops < 1 ||
        n_exc < 0 || (n_exc > 0 && (!exc_col || !exc_W)))
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    if (n_ops > 4) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "at most 4 operators per pass");
    for (int o = 0; o < n_ops; ++o)
        if (!out[o]) return fail(ctx, DCPSE_ERR_INVALID, "null output pointer");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0) return DCPSE_OK;
    Sell64Args a{};
    a.n_rows = n_rows;
    a.row_ptr = row_ptr;
    a.slice_ptr = slice_ptr;
    a.sell_row = sell_row16;
    a.bundle = bundle;
    a.exc_ptr = exc_ptr;
    a.exc_col = exc_col;
    a.exc_W = exc_W;
    a.n_exc = n_exc;
    for (int o = 0; o < 4; ++o) a.out[o] = o < n_ops ? out[o] : nullptr;
    a.f = f;
    a.f_self = f_self ? f_self : f;
    a.alpha = alpha;
    a.beta = beta;
    if (!dispatch_sell64<2, true>(a, n_ops, ctx->stream))
        return fail(ctx, DCPSE_ERR_UNSUPPORTED, "no SELL kernel for %d operators", n_ops);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

static dcpse_status sell_apply(dcpse_ctx* ctx, int mode, int64_t n_rows, const int64_t* row_ptr,
                               const int64_t* slice_ptr, const int32_t* sell_row, const int32_t* sell_col,
                               const int32_t* bundle, int32_t n_ops, const float* const* sell_W, const float* f,
                               const float* f_self, const float* diag, float* const* out, float alpha, float beta) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !slice_ptr || !sell_row || !(bundle || (sell_col && sell_W)) || !f || !out || n_rows < 0 ||
        n_ops < 1)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    if (n_ops > 4) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "at most 4 operators per pass");
    for (int o = 0; o < n_ops; ++o)
        if ((!bundle && !sell_W[o]) || !out[o]) return fail(ctx, DCPSE_ERR_INVALID, "null operator pointer");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0) return DCPSE_OK;
    SellArgs a{};
    a.n_rows = n_rows;
    a.row_ptr = row_ptr;
    a.slice_ptr = slice_ptr;
    a.sell_row = sell_row;
    a.sell_col = sell_col;
    a.bundle = bundle;
    for (int o = 0; o < 4; ++o) {
        a.W[o] = o < n_ops && sell_W ? sell_W[o] : nullptr;
        a.out[o] = o < n_ops ? out[o] : nullptr;
    }
    a.f = f;
    a.f_self = f_self ? f_self : f;
    a.diag = diag;
    a.alpha = alpha;
    a.beta = beta;
    a.y_a = a.y_b = nullptr;
    a.out_a = a.out_b = nullptr;
    a.ca = a.cb = 0.0f;
    const bool ok = mode == 1               ? dispatch_sell<1>(a, n_ops, ctx->sell_groups, ctx->stream)
                    : ctx->apply_acc == 32 ? dispatch_sell<0, true>(a, n_ops, ctx->sell_groups, ctx->stream)
                                           : dispatch_sell<0>(a, n_ops, ctx->sell_groups, ctx->stream);
    if (!ok) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "no SELL kernel for %d operators", n_ops);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

dcpse_status dcpse_sell_fill_bundle_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                                        int32_t n_ops, const float* const* W, const int64_t* slice_ptr,
                                        const int32_t* sell_row, int32_t* bundle) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !col || !W || !slice_ptr || !sell_row || !bundle || n_rows < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    if (n_ops < 1 || n_ops > 4) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "1 to 4 operators per bundle");
    for (int o = 0; o < n_ops; ++o)
        if (!W[o]) return fail(ctx, DCPSE_ERR_INVALID, "null weight pointer");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0) return DCPSE_OK;
    const int64_t b = std::min<int64_t>((n_rows + 63) / 64, 65536);
    k_sell_fill_bundle<<<(unsigned)b, 256, 0, ctx->stream>>>(n_rows, row_ptr, col, n_ops, W[0],
                                                             n_ops > 1 ? W[1] : nullptr, n_ops > 2 ? W[2] : nullptr,
                                                             n_ops > 3 ? W[3] : nullptr, slice_ptr, sell_row, bundle);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

dcpse_status dcpse_apply_sell_bundle_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                         const int64_t* slice_ptr, const int32_t* sell_row, int32_t n_ops,
                                         const int32_t* bundle, const float* f, const float* f_self,
                                         float* const* out, float alpha, float beta) {
    if (!bundle) return fail(ctx, DCPSE_ERR_INVALID, "bundle is null");
    return sell_apply(ctx, 0, n_rows, row_ptr, slice_ptr, sell_row, nullptr, bundle, n_ops, nullptr, f, f_self,
                      nullptr, out, alpha, beta);
}

dcpse_status dcpse_sell_bundle16_count(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                                       const int32_t* sell_row, int64_t* exc_ptr, int64_t* n_exc) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !col || !sell_row || !exc_ptr || !n_exc || n_rows < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    if (n_rows >= kExcFlag) return fail(ctx, DCPSE_ERR_RANGE, "too many particles for the 16-bit bundle");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int64_t n_lanes = ((n_rows + 63) / 64) * 64;
    Workspace w;
    dcpse_status st = reserve_workspace(ctx, (size_t)(n_lanes + 1) * 4 + scan_workspace_bytes(n_lanes) + 1024, &w);
    if (st) return st;
    int32_t* counts = w.take<int32_t>(n_lanes + 1);
    if (n_lanes > 0) {
        const int64_t b = std::min<int64_t>(n_lanes / 64, 65536);
        k_bundle16_count<<<(unsigned)b, 256, 0, ctx->stream>>>(n_rows, row_ptr, col, sell_row, counts);
        DCPSE_LAUNCH_CHECK(ctx);
    }
    if ((st = scan_counts(ctx, counts, n_lanes, exc_ptr, &w))) return st;
    int64_t total = 0;
    DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(&total, exc_ptr + n_lanes, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    DCPSE_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *n_exc = total;
    return DCPSE_OK;
}

dcpse_status dcpse_sell_fill_bundle16_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                          const int32_t* col, int32_t n_ops, const float* const* W,
                                          const int64_t* slice_ptr, const int32_t* sell_row,
                                          const int64_t* exc_ptr, int64_t n_exc, int32_t* bundle,
                                          int32_t* sell_row16, int32_t* exc_col, float* exc_W) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !col || !W || !slice_ptr || !sell_row || !exc_ptr || !bundle || !sell_row16 || n_rows < 0 ||
        n_exc < 0 || (n_exc > 0 && (!exc_col || !exc_W)))
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    if (n_ops < 1 || n_ops > 4) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "1 to 4 operators per bundle");
    if (n_rows >= kExcFlag) return fail(ctx, DCPSE_ERR_RANGE, "too many particles for the 16-bit bundle");
    for (int o = 0; o < n_ops; ++o)
        if (!W[o]) return fail(ctx, DCPSE_ERR_INVALID, "null weight pointer");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0) return DCPSE_OK;
    const int64_t b = std::min<int64_t>((n_rows + 63) / 64, 65536);
    k_sell_fill_bundle16<<<(unsigned)b, 256, 0, ctx->stream>>>(
        n_rows, row_ptr, col, n_ops, W[0], n_ops > 1 ? W[1] : nullptr, n_ops > 2 ? W[2] : nullptr,
        n_ops > 3 ? W[3] : nullptr, slice_ptr, sell_row, exc_ptr, bundle, sell_row16);
    DCPSE_LAUNCH_CHECK(ctx);
    if (n_exc > 0) {
        const int64_t be = std::min<int64_t>((n_rows + 63 + 255) / 256, 65536);
        k_bundle16_exc<float><<<(unsigned)be, 256, 0, ctx->stream>>>(
            n_rows, row_ptr, col, n_ops, W[0], n_ops > 1 ? W[1] : nullptr, n_ops > 2 ? W[2] : nullptr,
            n_ops > 3 ? W[3] : nullptr, sell_row, exc_ptr, n_exc, exc_col, exc_W);
    }
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

dcpse_status dcpse_apply_sell_bundle16_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                           const int64_t* slice_ptr, const int32_t* sell_row16, int32_t n_ops,
                                           const int32_t* bundle, const int64_t* exc_ptr, const int32_t* exc_col,
                                           const float* exc_W, int64_t n_exc, const float* f, const float* f_self,
                                           float* const* out, float alpha, float beta) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !slice_ptr || !sell_row16 || !bundle || !exc_ptr || !f || !out || n_rows < 0 || n_ops < 1 ||
        n_exc < 0 || (n_exc > 0 && (!exc_col || !exc_W)))
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    if (n_ops > 4) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "at most 4 operators per pass");
    for (int o = 0; o < n_ops; ++o)
        if (!out[o]) return fail(ctx, DCPSE_ERR_INVALID, "null output pointer");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0) return DCPSE_OK;
    SellArgs a{};
    a.n_rows = n_rows;
    a.row_ptr = row_ptr;
    a.slice_ptr = slice_ptr;
    a.sell_row = sell_row16;
    a.bundle = bundle;
    a.exc_ptr = exc_ptr;
    a.exc_col = exc_col;
    a.exc_W = exc_W;
    a.n_exc = n_exc;
    for (int o = 0; o < 4; ++o) a.out[o] = o < n_ops ? out[o] : nullptr;
    a.f = f;
    a.f_self = f_self ? f_self : f;
    a.alpha = alpha;
    a.beta = beta;
    if (!dispatch_sell16(a, n_ops, ctx->sell_groups == 4 ? 4 : 2, ctx->apply_acc == 32, ctx->stream))
        return fail(ctx, DCPSE_ERR_UNSUPPORTED, "no SELL kernel for %d operators", n_ops);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

dcpse_status dcpse_stage_sell_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int64_t* slice_ptr,
                                  const int32_t* sell_row, const int32_t* sell_col, const float* sell_W,
                                  const float* u, float alpha, const float* y_a, float ca, float* out_a,
                                  const float* y_b, float cb, float* out_b) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !slice_ptr || !sell_row || !sell_col || !sell_W || !u || n_rows < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    if ((out_a && !y_a) || (out_b && !y_b)) return fail(ctx, DCPSE_ERR_INVALID, "an output without its input");
    if (out_a == u || out_b == u) return fail(ctx, DCPSE_ERR_INVALID, "a stage output may not alias its input u");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0 || (!out_a && !out_b)) return DCPSE_OK;
    SellArgs a{};
    a.n_rows = n_rows;
    a.row_ptr = row_ptr;
    a.slice_ptr = slice_ptr;
    a.sell_row = sell_row;
    a.sell_col = sell_col;
    a.bundle = nullptr;
    for (int o = 0; o < 4; ++o) {
        a.W[o] = o == 0 ? sell_W : nullptr;
        a.out[o] = nullptr;
    }
    a.f = u;
    a.f_self = u;
    a.diag = nullptr;
    a.alpha = alpha;
    a.beta = 0.0f;
    a.y_a = y_a;
    a.y_b = y_b;
    a.out_a = out_a;
    a.out_b = out_b;
    a.ca = ca;
    a.cb = cb;
    if (!dispatch_sell<2>(a, 1, ctx->sell_groups, ctx->stream))
        return fail(ctx, DCPSE_ERR_UNSUPPORTED, "no SELL stage kernel for this pipeline");
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

dcpse_status dcpse_stage_bundle16_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr,
                                      const int64_t* slice_ptr, const int32_t* sell_row16, const int32_t* bundle,
                                      const int64_t* exc_ptr, const int32_t* exc_col, const float* exc_W,
                                      int64_t n_exc, const float* u, float alpha, const float* y_a, float ca,
                                      float* out_a, const float* y_b, float cb, float* out_b) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !slice_ptr || !sell_row16 || !bundle || !exc_ptr || !u || n_rows < 0 || n_exc < 0 ||
        (n_exc > 0 && (!exc_col || !exc_W)))
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    if ((out_a && !y_a) || (out_b && !y_b)) return fail(ctx, DCPSE_ERR_INVALID, "an output without its input");
    if (out_a == u || out_b == u) return fail(ctx, DCPSE_ERR_INVALID, "a stage output may not alias its input u");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n_rows == 0 || (!out_a && !out_b)) return DCPSE_OK;
    SellArgs a{};
    a.n_rows = n_rows;
    a.row_ptr = row_ptr;
    a.slice_ptr = slice_ptr;
    a.sell_row = sell_row16;
    a.bundle = bundle;
    a.exc_ptr = exc_ptr;
    a.exc_col = exc_col;
    a.exc_W = exc_W;
    a.n_exc = n_exc;
    a.f = u;
    a.f_self = u;
    a.alpha = alpha;
    a.y_a = y_a;
    a.y_b = y_b;
    a.out_a = out_a;
    a.out_b = out_b;
    a.ca = ca;
    a.cb = cb;
    if (ctx->sell_groups == 4) launch_sell16<1, 4, 2>(a, ctx->stream);
    else launch_sell16<1, 2, 2>(a, ctx->stream);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

dcpse_status dcpse_apply_sell_f32(dcpse_ctx* ctx, int64_t n_rows, const int64_t* row_ptr, const int64_t* slice_ptr,
                                  const int32_t* sell_row, const int32_t* sell_col, int32_t n_ops,
                                  const float* const* sell_W, const float* f, const float* f_self, float* const* out,
                                  float alpha, float beta) {
    return sell_apply(ctx, 0, n_rows, row_ptr, slice_ptr, sell_row, sell_col, nullptr, n_ops, sell_W, f, f_self,
                      nullptr, out, alpha, beta);
}

dcpse_status dcpse_apply_transpose_sell_f32(dcpse_ctx* ctx, int64_t n_cols, const int64_t* rowT_ptr,
                                            const int64_t* slice_ptr, const int32_t* sell_rowT,
                                            const int32_t* sell_colT, const float* sell_WT, const float* diag,
                                            const float* g, float* out, float alpha, float beta) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!diag) return fail(ctx, DCPSE_ERR_INVALID, "diag is null");
    const float* Wp[1] = {sell_WT};
    float* Op[1] = {out};
    return sell_apply(ctx, 1, n_cols, rowT_ptr, slice_ptr, sell_rowT, sell_colT, nullptr, 1, Wp, g, nullptr, diag, Op,
                      alpha, beta);
}

dcpse_status dcpse_apply_transpose_bundle16_f32(dcpse_ctx* ctx, int64_t n, const int64_t* row_ptr,
                                                const int64_t* slice_ptr, const int32_t* sell_row16,
                                                const int32_t* bundle, const int64_t* exc_ptr,
                                                const int32_t* exc_col, const float* exc_W, int64_t n_exc,
                                                const float* diag, const float* g, float* out, float alpha,
                                                float beta) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !slice_ptr || !sell_row16 || !bundle || !exc_ptr || !diag || !g || !out || n < 0 || n_exc < 0 ||
        (n_exc > 0 && (!exc_col || !exc_W)))
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or bad size");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (n == 0) return DCPSE_OK;
    SellArgs a{};
    a.n_rows = n;
    a.row_ptr = row_ptr;
    a.slice_ptr = slice_ptr;
    a.sell_row = sell_row16;
    a.bundle = bundle;
    a.exc_ptr = exc_ptr;
    a.exc_col = exc_col;
    a.exc_W = exc_W;
    a.n_exc = n_exc;
    a.f = g;
    a.f_self = g;
    a.diag = diag;
    a.out[0] = out;
    a.alpha = alpha;
    a.beta = beta;
    if (ctx->sell_groups == 4) launch_sell16<1, 4, 1>(a, ctx->stream);
    else launch_sell16<1, 2, 1>(a, ctx->stream);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

dcpse_status dcpse_set_apply_accumulation(dcpse_ctx* ctx, int32_t bits) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (bits != 32 && bits != 64) return fail(ctx, DCPSE_ERR_INVALID, "accumulation must be 32 or 64 bits");
    ctx->apply_acc = bits;
    return DCPSE_OK;
}

dcpse_status dcpse_set_sell_pipeline(dcpse_ctx* ctx, int32_t groups, int32_t depth) {
    if (!ctx) return DCPSE_ERR_INVALID;
    const bool ok = depth == 2 && (groups == 2 || groups == 4);
    if (!ok) return fail(ctx, DCPSE_ERR_INVALID, "(groups, depth) must be (2, 2) or (4, 2)");
    ctx->sell_groups = groups;
    ctx->sell_depth = depth;
    return DCPSE_OK;
}

}  // extern "C"
