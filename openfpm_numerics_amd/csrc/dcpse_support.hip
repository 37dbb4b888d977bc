// Neighbour-support (Verlet CSR) builder, device scan and transposed-CSR
// builder for gfx950.
//
// Support semantics follow the reference's Verlet list for DC-PSE
// (VL_NON_SYMMETRIC | VL_SKIP_REF_PART: q != p, |x_p - x_q| < rCut), see
// src/DCPSE/SupportBuilder.cuh:15-61. Unlike that builder (one thread per
// particle, host-side serial prefix sum, size_t keys, 4 host round trips), the
// whole CSR is built on the device: a cell grid by counting sort, one
// wavefront per particle walking the 3^dim neighbour cells as one flattened
// candidate list, ballot compaction into LDS and an in-LDS rank sort so every
// row comes out in ascending particle index (deterministic, independent of
// atomic ordering).
#include <algorithm>
#include <cmath>

#include "dcpse_common.hpp"
#include "dcpse_internal.hpp"

namespace dcpse {

// ---------------------------------------------------------------------------
// scan
// ---------------------------------------------------------------------------
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ __launch_bounds__(kScanBlock) void k_scan_tilesum(const int32_t* __restrict__ in, int64_t n,
                                                             int64_t* __restrict__ tsum) {
    __shared__ long long red[kScanBlock / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    long long s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t k = base + (int64_t)i * kScanBlock + threadIdx.x;
        if (k < n) s += in[k];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < kScanBlock / kWave; ++w) t += red[w];
        tsum[blockIdx.x] = t;
    }
}

// Exclusive scan of tsum[0..nt) in place, tsum[nt] = total. One block.
__global__ __launch_bounds__(1024) void k_scan_tilesums(int64_t* tsum, int64_t nt) {
    __shared__ long long buf[1024];
    long long carry = 0;
    for (int64_t b = 0; b < nt; b += 1024) {
        int64_t i = b + threadIdx.x;
        long long v = i < nt ? tsum[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            long long t = threadIdx.x >= (unsigned)off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nt) tsum[i] = carry + buf[threadIdx.x] - v;
        long long tot = buf[1023];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) tsum[nt] = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const int32_t* __restrict__ in, int64_t n,
                                                           const int64_t* __restrict__ tsum,
                                                           int64_t* __restrict__ out) {
    __shared__ long long buf[kScanBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    long long v[kScanItems];
    long long tot = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t k = base + i;
        v[i] = k < n ? in[k] : 0;
        tot += v[i];
    }
    buf[threadIdx.x] = tot;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        long long t = threadIdx.x >= (unsigned)off ? buf[threadIdx.x - off] : 0;
        __syncthreads();
        buf[threadIdx.x] += t;
        __syncthreads();
    }
    long long run = tsum[blockIdx.x] + buf[threadIdx.x] - tot;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t k = base + i;
        if (k < n) out[k] = run;
        run += v[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = tsum[gridDim.x];
}

size_t scan_workspace_bytes(int64_t n) {
    int64_t nt = (n + kScanTile - 1) / kScanTile;
    return (size_t)(nt + 1) * sizeof(int64_t) + 256;
}

dcpse_status scan_counts(dcpse_ctx* ctx, const int32_t* in, int64_t n, int64_t* out, Workspace* w) {
    int64_t nt = (n + kScanTile - 1) / kScanTile;
    if (nt == 0) nt = 1;
    int64_t* tsum = w->take<int64_t>((size_t)nt + 1);
    if (!tsum) return fail(ctx, DCPSE_ERR_INVALID, "scan: workspace too small");
    k_scan_tilesum<<<(unsigned)nt, kScanBlock, 0, ctx->stream>>>(in, n, tsum);
    DCPSE_LAUNCH_CHECK(ctx);
    k_scan_tilesums<<<1, 1024, 0, ctx->stream>>>(tsum, nt);
    DCPSE_LAUNCH_CHECK(ctx);
    k_scan_tiles<<<(unsigned)nt, kScanBlock, 0, ctx->stream>>>(in, n, tsum, out);
    DCPSE_LAUNCH_CHECK(ctx);
    return DCPSE_OK;
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v) {
    if (n <= 0) return;
    int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    k_fill_i32<<<(unsigned)blocks, 256, 0, s>>>(p, n, v);
}

dcpse_status reserve_workspace(dcpse_ctx* ctx, size_t bytes, Workspace* w) {
    if (ctx->ws_size < bytes) {
        DCPSE_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->ws) DCPSE_HIP_CHECK(ctx, hipFree(ctx->ws));
        ctx->ws = nullptr;
        size_t nb = std::max(bytes, ctx->ws_size * 2);
        DCPSE_HIP_CHECK(ctx, hipMalloc(&ctx->ws, nb));
        ctx->ws_size = nb;
    }
    w->base = static_cast<char*>(ctx->ws);
    w->cap = ctx->ws_size;
    w->used = 0;
    return DCPSE_OK;
}

// ---------------------------------------------------------------------------
// cell grid
// ---------------------------------------------------------------------------
struct CellGrid {
    int nc[3];
    int kmax[3];  // cells scanned on each side: ceil(largest rcut / cs), capped so a periodic wrap visits no cell twice
    double lo[3], cs[3], len[3], ilen[3];
    unsigned periodic;
    int64_t ncell;
};

__device__ __forceinline__ int cell_of(double x, int d, const CellGrid& g) {
    double t = floor((x - g.lo[d]) / g.cs[d]);
    t = fmin(fmax(t, -1.0e9), 1.0e9);
    int v = (int)t;
    if ((g.periodic >> d) & 1u) {
        v %= g.nc[d];
        if (v < 0) v += g.nc[d];
    } else {
        v = min(max(v, 0), g.nc[d] - 1);
    }
    return v;
}

template <int DIM>
__global__ void k_cell_assign(const double* __restrict__ pos, int64_t n, CellGrid g,
                              int32_t* __restrict__ cell_id, int32_t* __restrict__ cell_cnt) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = 0, stride = 1;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            c += stride * cell_of(pos[d * n + q], d, g);
            stride *= g.nc[d];
        }
        cell_id[q] = (int32_t)c;
        atomicAdd(&cell_cnt[c], 1);
    }
}

// cell_part[slot] = particle, and a copy of its position in cell order
// (pos_cell[d * n + slot]) so the row kernels read candidate positions
// contiguously per cell, without waiting for the particle index first.
template <int DIM>
__global__ void k_cell_fill(const int32_t* __restrict__ cell_id, int64_t n, const double* __restrict__ pos,
                            const int64_t* __restrict__ cell_start, int32_t* __restrict__ cursor,
                            int32_t* __restrict__ cell_part, double* __restrict__ pos_cell) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        int32_t c = cell_id[q];
        int64_t at = cell_start[c] + atomicAdd(&cursor[c], 1);
        cell_part[at] = (int32_t)q;
#pragma unroll
        for (int d = 0; d < DIM; ++d) pos_cell[d * n + at] = pos[d * n + q];
    }
}

constexpr int kRowWaves = 4;  // wavefronts per block in the row kernels

// One wavefront per domain particle. FILL = false: count the support of p into
// counts[p] (and raise *overflow if it exceeds DCPSE_MAX_ROW). FILL = true:
// write the support of p, sorted ascending, to col[row_ptr[p] ..].
// The cells within k_d = ceil(rcut_p / cell edge) of p's cell in each
// dimension are scanned as rows along dimension 0: cells of one row are
// consecutive in the cell order, so a row is one contiguous candidate segment
// (two where a periodic wrap splits it). Up to 32 rows at a time, one segment
// per lane, are flattened into one candidate list by a prefix sum over the
// lanes.
template <int DIM, bool FILL>
__global__ __launch_bounds__(kRowWaves * kWave) void k_support_rows(
    const double* __restrict__ pos_sup, int64_t n_sup, const double* __restrict__ pos_dom,
    int64_t n_dom, int same_set, const double* __restrict__ rcut, int rcut_pp, CellGrid g,
    const int64_t* __restrict__ cell_start, const int32_t* __restrict__ cell_part,
    const double* __restrict__ pos_cell, int32_t* __restrict__ counts, int32_t* __restrict__ overflow,
    const int64_t* __restrict__ row_ptr, int32_t* __restrict__ col) {
    __shared__ int32_t buf[FILL ? kRowWaves : 1][FILL ? DCPSE_MAX_ROW : 1];
    __shared__ int32_t mark[kRowWaves][kWave];      // segment starting at a candidate of the round
    __shared__ long long seg_cst[kRowWaves][kWave];  // first cell slot of each segment
    __shared__ int32_t seg_ex[kRowWaves][kWave];     // candidates before each segment
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    for (int64_t p = (int64_t)blockIdx.x * kRowWaves + wv; p < n_dom;
         p += (int64_t)gridDim.x * kRowWaves) {
        double xp[3];
        int c[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            xp[d] = pos_dom[d * n_dom + p];
            c[d] = cell_of(xp[d], d, g);
        }
        const double r = rcut_pp ? rcut[p] : rcut[0];
        const double r2 = r * r;
        int kd[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; ++d) kd[d] = min((int)ceil(r / g.cs[d]), g.kmax[d]);
        const int ny = 2 * kd[1] + 1;
        const int rows = DIM == 3 ? ny * (2 * kd[2] + 1) : ny;
        int nacc = 0;
        for (int rb = 0; rb < rows; rb += kWave / 2) {
            // lane 2i (+1): row rb + i (its part past a periodic wrap)
            long long cst = 0;
            int ccnt = 0;
            const int row = rb + (lane >> 1), part = lane & 1;
            if (row < rows) {
                const int o1 = row % ny - kd[1];
                const int o2 = DIM == 3 ? row / ny - kd[2] : 0;
                int cy = c[1] + o1, cz = DIM == 3 ? c[2] + o2 : 0;
                bool ok = true;
                if ((g.periodic >> 1) & 1u) cy = (cy % g.nc[1] + g.nc[1]) % g.nc[1];
                else ok &= cy >= 0 && cy < g.nc[1];
                if (DIM == 3) {
                    if ((g.periodic >> 2) & 1u) cz = (cz % g.nc[2] + g.nc[2]) % g.nc[2];
                    else ok &= cz >= 0 && cz < g.nc[2];
                }
                int xa = c[0] - kd[0], xb = c[0] + kd[0];
                if (g.periodic & 1u) {
                    if (xa < 0) {
                        if (part == 0) xa += g.nc[0], xb = g.nc[0] - 1;
                        else xa = 0;
                    } else if (xb >= g.nc[0]) {
                        if (part == 0) xb = g.nc[0] - 1;
                        else xa = 0, xb -= g.nc[0];
                    } else if (part == 1) {
                        ok = false;
                    }
                } else {
                    xa = max(xa, 0);
                    xb = min(xb, g.nc[0] - 1);
                    ok &= part == 0;
                }
                if (ok && xa <= xb) {
                    const int64_t lin = (int64_t)g.nc[0] * ((int64_t)cy + (int64_t)g.nc[1] * cz);
                    cst = cell_start[lin + xa];
                    ccnt = (int)(cell_start[lin + xb + 1] - cst);
                }
            }
            // inclusive prefix over lanes of the candidate counts
            const int incl = wave_incl_scan(ccnt);
            const int total = __builtin_amdgcn_readlane(incl, 63);
            const int excl = incl - ccnt;
            seg_cst[wv][lane] = cst;
            seg_ex[wv][lane] = excl;
            // segment of candidate k: the last non-empty segment starting at
            // or before k. Each round marks the segments that start inside its
            // 64 candidates, a prefix maximum spreads them, and a candidate
            // before the round's first start continues the previous segment.
            int carry = 0;
            for (int base = 0; base < total; base += kWave) {
                const int k = min(base + lane, total - 1);
                mark[wv][lane] = -1;
                wave_lds_sync();
                if (ccnt > 0 && excl >= base && excl < base + kWave) mark[wv][excl - base] = lane;
                wave_lds_sync();
                const int seg = max(wave_incl_max(mark[wv][lane]), carry);
                carry = __builtin_amdgcn_readlane(seg, 63);
                const long long s_cst = seg_cst[wv][seg];
                const int s_ex = seg_ex[wv][seg];
                const int64_t slot = s_cst + (k - s_ex);
                const int q = cell_part[slot];
                double d[3];
#pragma unroll
                for (int dd = 0; dd < DIM; ++dd)
                    d[dd] = min_image_r(xp[dd] - pos_cell[dd * n_sup + slot], g.len[dd], g.ilen[dd],
                                        (g.periodic >> dd) & 1u);
                double d2 = d[DIM - 1] * d[DIM - 1];
#pragma unroll
                for (int dd = DIM - 2; dd >= 0; --dd) d2 = fma(d[dd], d[dd], d2);
                const bool take = (base + lane < total) && !(same_set && q == (int)p) && d2 < r2;
                const unsigned long long m = __ballot(take);
                if (FILL) {
                    const int at = nacc + lane_prefix(m);
                    if (take && at < DCPSE_MAX_ROW) buf[wv][at] = q;
                }
                nacc += __popcll(m);
            }
        }
        if (!FILL) {
            if (lane == 0) {
                counts[p] = nacc;
                if (nacc > DCPSE_MAX_ROW) atomicOr(overflow, 1);
            }
        } else {
            wave_lds_sync();
            const int64_t rs = row_ptr[p];
            const int n = min(nacc, DCPSE_MAX_ROW);
            if ((int64_t)nacc != row_ptr[p + 1] - rs) {
                if (lane == 0) atomicOr(overflow, 2);
            } else {
                for (int e = lane; e < n; e += kWave) {
                    const int key = buf[wv][e];
                    int rank = 0;
                    for (int j = 0; j < n; ++j) rank += buf[wv][j] < key;
                    col[rs + rank] = key;
                }
            }
            wave_lds_sync();
        }
    }
}

// Cell edge about rmax / K: smaller cells than the cut-off shrink the scanned
// volume around a particle ((2k+1)^dim cells of edge rmax/K instead of 3^dim
// of edge rmax), and let particles with a smaller cut-off scan fewer cells.
// A periodic dimension needs at least 2k+1 cells; where the box is too small
// for that, the edge falls back to rmax (k = 1, at least 3 cells).
static dcpse_status make_grid(dcpse_ctx* ctx, int dim, const double* box_lo, const double* box_hi,
                              unsigned periodic, double rmax, int K, CellGrid* g) {
    if (!(rmax > 0.0) || !std::isfinite(rmax)) return fail(ctx, DCPSE_ERR_INVALID, "rcut must be > 0");
    g->ncell = 1;
    g->periodic = periodic;
    for (int d = 0; d < 3; ++d) {
        if (d < dim) {
            double len = box_hi[d] - box_lo[d];
            if (!(len > 0.0)) return fail(ctx, DCPSE_ERR_INVALID, "box_hi must exceed box_lo");
            int n = (int)std::floor(len * K / rmax);
            if (n < 1) n = 1;
            // keep the cell grid bounded (cells beyond ~2^26 add nothing but memory)
            while ((double)g->ncell * n > 6.7e7 && n / 2 >= 3) n /= 2;
            int kmax = (int)std::ceil(rmax / (len / n));
            if (((periodic >> d) & 1u) && n < 2 * kmax + 1) {
                n = (int)std::floor(len / rmax);
                if (n < 3)
                    return fail(ctx, DCPSE_ERR_INVALID,
                                "periodic dimension %d has %d cells of edge >= rcut (need 3)", d, n);
                kmax = 1;
            }
            g->nc[d] = n;
            g->kmax[d] = kmax;
            g->lo[d] = box_lo[d];
            g->len[d] = len;
            g->ilen[d] = 1.0 / len;
            g->cs[d] = len / n;
            g->ncell *= n;
        } else {
            g->nc[d] = 1; g->kmax[d] = 0; g->lo[d] = 0; g->len[d] = 1; g->ilen[d] = 1; g->cs[d] = 1;
        }
    }
    return DCPSE_OK;
}

static dcpse_status support_pass(dcpse_ctx* ctx, int dim, const double* pos_sup, int64_t n_sup,
                                 const double* pos_dom, int64_t n_dom, const double* rcut, int rcut_pp,
                                 const double* box_lo, const double* box_hi, unsigned periodic,
                                 int64_t* row_ptr, int32_t* col, int64_t* nnz_out) {
    if (dim != 2 && dim != 3) return fail(ctx, DCPSE_ERR_UNSUPPORTED, "dim must be 2 or 3");
    if (!pos_sup || !rcut || !row_ptr || !box_lo || !box_hi || n_sup < 0 || n_dom < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    if (n_sup >= (int64_t)1 << 31 || n_dom >= (int64_t)1 << 31)
        return fail(ctx, DCPSE_ERR_RANGE, "particle counts must be < 2^31");
    const int same = pos_dom == nullptr;
    if (same) { pos_dom = pos_sup; n_dom = n_sup; }
    // the largest cut-off fixes the cell edge
    double rmax = 0.0, rmin = 0.0;
    if (rcut_pp) {
        // reduce on the host: one copy of n_dom doubles is cheap next to the build
        double* h = (double*)malloc(sizeof(double) * (size_t)std::max<int64_t>(n_dom, 1));
        if (!h) return fail(ctx, DCPSE_ERR_INVALID, "host allocation failed");
        hipError_t e = hipMemcpyAsync(h, rcut, sizeof(double) * n_dom, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) { free(h); DCPSE_HIP_CHECK(ctx, e); }
        rmin = n_dom > 0 ? h[0] : 0.0;
        for (int64_t i = 0; i < n_dom; ++i) {
            rmax = std::max(rmax, h[i]);
            rmin = std::min(rmin, h[i]);
        }
        free(h);
    } else {
        DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(&rmax, rcut, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        DCPSE_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        rmin = rmax;
    }
    // cells per cut-off: 2 (3D) or 4 (2D) for one cut-off; with a spread of
    // cut-offs, about rmax / rmin so the smallest scan only their own
    // neighbourhood (bounded: a wave handles 32 cell rows per round)
    const double spread = rmin > 0.0 ? rmax / rmin : 1.0;
    const int K = dim == 3 ? std::min(4, std::max(2, (int)std::ceil(spread - 1e-9)))
                           : std::min(8, std::max(4, (int)std::ceil(spread - 1e-9)));
    CellGrid g;
    dcpse_status st = make_grid(ctx, dim, box_lo, box_hi, periodic, rmax, K, &g);
    if (st) return st;

    Workspace w;
    size_t need = (size_t)n_sup * 4 * 2 + (size_t)dim * (n_sup + 1) * 8 + (size_t)(g.ncell + 1) * 8 +
                  (size_t)g.ncell * 4 * 2 + (size_t)n_dom * 4 + scan_workspace_bytes(std::max(g.ncell, n_dom)) +
                  64 * 256;
    if ((st = reserve_workspace(ctx, need, &w))) return st;
    int32_t* cell_id = w.take<int32_t>(n_sup + 1);
    int32_t* cell_part = w.take<int32_t>(n_sup + 1);
    double* pos_cell = w.take<double>((size_t)dim * (n_sup + 1));
    int32_t* cell_cnt = w.take<int32_t>(g.ncell + 1);
    int64_t* cell_start = w.take<int64_t>(g.ncell + 1);
    int32_t* counts = w.take<int32_t>(n_dom + 1);
    int32_t* flags = w.take<int32_t>(64);
    Workspace ws_scan = w;
    if (!flags) return fail(ctx, DCPSE_ERR_INVALID, "workspace sizing error");

    hipStream_t s = ctx->stream;
    fill_i32(s, cell_cnt, g.ncell + 1, 0);
    fill_i32(s, flags, 64, 0);
    const int64_t tb = std::max<int64_t>(1, std::min<int64_t>((n_sup + 255) / 256, 16384));
    if (n_sup > 0) {
        if (dim == 2) k_cell_assign<2><<<(unsigned)tb, 256, 0, s>>>(pos_sup, n_sup, g, cell_id, cell_cnt);
        else k_cell_assign<3><<<(unsigned)tb, 256, 0, s>>>(pos_sup, n_sup, g, cell_id, cell_cnt);
        DCPSE_LAUNCH_CHECK(ctx);
    }
    if ((st = scan_counts(ctx, cell_cnt, g.ncell, cell_start, &ws_scan))) return st;
    fill_i32(s, cell_cnt, g.ncell + 1, 0);  // reuse as cursor
    if (n_sup > 0) {
        if (dim == 2)
            k_cell_fill<2><<<(unsigned)tb, 256, 0, s>>>(cell_id, n_sup, pos_sup, cell_start, cell_cnt, cell_part,
                                                        pos_cell);
        else
            k_cell_fill<3><<<(unsigned)tb, 256, 0, s>>>(cell_id, n_sup, pos_sup, cell_start, cell_cnt, cell_part,
                                                        pos_cell);
        DCPSE_LAUNCH_CHECK(ctx);
    }
    const int64_t rb = std::max<int64_t>(1, std::min<int64_t>((n_dom + kRowWaves - 1) / kRowWaves,
                                                              (int64_t)ctx->num_cus * 16));
    if (col == nullptr) {
        if (n_dom > 0) {
            if (dim == 2)
                k_support_rows<2, false><<<(unsigned)rb, kRowWaves * kWave, 0, s>>>(
                    pos_sup, n_sup, pos_dom, n_dom, same, rcut, rcut_pp, g, cell_start, cell_part, pos_cell, counts,
                    flags, nullptr, nullptr);
            else
                k_support_rows<3, false><<<(unsigned)rb, kRowWaves * kWave, 0, s>>>(
                    pos_sup, n_sup, pos_dom, n_dom, same, rcut, rcut_pp, g, cell_start, cell_part, pos_cell, counts,
                    flags, nullptr, nullptr);
            DCPSE_LAUNCH_CHECK(ctx);
        }
        Workspace ws2 = ws_scan;
        if ((st = scan_counts(ctx, counts, n_dom, row_ptr, &ws2))) return st;
        int64_t nnz = 0;
        int32_t ovf = 0;
        DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(&nnz, row_ptr + n_dom, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(&ovf, flags, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        DCPSE_HIP_CHECK(ctx, hipStreamSynchronize(s));
        if (ovf & 1)
            return fail(ctx, DCPSE_ERR_RANGE, "a support row exceeds DCPSE_MAX_ROW=%d neighbours",
                        DCPSE_MAX_ROW);
        if (nnz >= ((int64_t)1 << 31) * 64)
            return fail(ctx, DCPSE_ERR_RANGE, "nnz out of range");
        if (nnz_out) *nnz_out = nnz;
    } else {
        if (n_dom > 0) {
            if (dim == 2)
                k_support_rows<2, true><<<(unsigned)rb, kRowWaves * kWave, 0, s>>>(
                    pos_sup, n_sup, pos_dom, n_dom, same, rcut, rcut_pp, g, cell_start, cell_part, pos_cell, counts,
                    flags, row_ptr, col);
            else
                k_support_rows<3, true><<<(unsigned)rb, kRowWaves * kWave, 0, s>>>(
                    pos_sup, n_sup, pos_dom, n_dom, same, rcut, rcut_pp, g, cell_start, cell_part, pos_cell, counts,
                    flags, row_ptr, col);
            DCPSE_LAUNCH_CHECK(ctx);
        }
        int32_t ovf = 0;
        DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(&ovf, flags, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        DCPSE_HIP_CHECK(ctx, hipStreamSynchronize(s));
        if (ovf)
            return fail(ctx, DCPSE_ERR_INVALID,
                        "support fill disagrees with row_ptr (positions or rcut changed since count?)");
    }
    return DCPSE_OK;
}

// ---------------------------------------------------------------------------
// transposed CSR
// ---------------------------------------------------------------------------
__global__ void k_tcount(const int32_t* __restrict__ col, int64_t nnz, int32_t* __restrict__ cnt) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz;
         j += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[col[j]], 1);
}

// One wavefront per source row p, a lane per entry: each entry takes its
// slot in transposed row col[j] by an atomic cursor (the order is undone by
// k_tsort) and the row sum for diag[p] is a wave sum in FP64.
template <typename T>
__global__ __launch_bounds__(kRowWaves * kWave) void k_tfill(int64_t n_rows, const int64_t* __restrict__ row_ptr,
                                                             const int32_t* __restrict__ col,
                                                             const T* __restrict__ W,
                                                             const int64_t* __restrict__ rowT_ptr,
                                                             int32_t* __restrict__ cursor,
                                                             int32_t* __restrict__ colT, T* __restrict__ WT,
                                                             T* __restrict__ diag) {
    const int lane = threadIdx.x & 63;
    for (int64_t p = (int64_t)blockIdx.x * kRowWaves + (threadIdx.x >> 6); p < n_rows;
         p += (int64_t)gridDim.x * kRowWaves) {
        const int64_t rs = row_ptr[p], re = row_ptr[p + 1];
        double acc = 0.0;
        for (int64_t j = rs + lane; j < re; j += kWave) {
            const int32_t q = col[j];
            const T wj = W[j];
            const int64_t at = rowT_ptr[q] + atomicAdd(&cursor[q], 1);
            colT[at] = (int32_t)p;
            WT[at] = wj;
            acc += (double)wj;
        }
        acc = wave_sum(acc);
        if (lane == 0) diag[p] = (T)(-acc);
    }
}

// Sorts each transposed row by ascending source row (rank sort in LDS); rows
// longer than DCPSE_MAX_ROW raise *overflow and are left unsorted.
template <typename T>
__global__ __launch_bounds__(kRowWaves * kWave) void k_tsort(int64_t n_cols, const int64_t* __restrict__ rowT_ptr,
                                                             int32_t* __restrict__ colT, T* __restrict__ WT,
                                                             int32_t* __restrict__ overflow) {
    __shared__ int32_t kbuf[kRowWaves][DCPSE_MAX_ROW];
    __shared__ T vbuf[kRowWaves][DCPSE_MAX_ROW];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    for (int64_t r = (int64_t)blockIdx.x * kRowWaves + wv; r < n_cols; r += (int64_t)gridDim.x * kRowWaves) {
        const int64_t rs = rowT_ptr[r];
        const int64_t n64 = rowT_ptr[r + 1] - rs;
        if (n64 > DCPSE_MAX_ROW) {
            if (lane == 0) atomicOr(overflow, 1);
            continue;
        }
        const int n = (int)n64;
        for (int e = lane; e < n; e += kWave) {
            kbuf[wv][e] = colT[rs + e];
            vbuf[wv][e] = WT[rs + e];
        }
        wave_lds_sync();
        for (int e = lane; e < n; e += kWave) {
            const int key = kbuf[wv][e];
            int rank = 0;
            for (int j = 0; j < n; ++j) rank += kbuf[wv][j] < key;
            colT[rs + rank] = key;
            WT[rs + rank] = vbuf[wv][e];
        }
        wave_lds_sync();
    }
}

template <typename T>
static dcpse_status transpose_fill(dcpse_ctx* ctx, int64_t n_rows, int64_t n_cols, const int64_t* row_ptr,
                                   const int32_t* col, const T* W, const int64_t* rowT_ptr, int32_t* colT,
                                   T* WT, T* diag) {
    if (!row_ptr || !col || !W || !rowT_ptr || !colT || !WT || !diag || n_rows < 0 || n_cols < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    if (n_rows != n_cols)
        return fail(ctx, DCPSE_ERR_UNSUPPORTED, "transpose needs the domain set to be the support set");
    Workspace w;
    dcpse_status st = reserve_workspace(ctx, (size_t)(n_cols + 64) * 4 + 1024, &w);
    if (st) return st;
    int32_t* cursor = w.take<int32_t>(n_cols + 1);
    int32_t* flags = w.take<int32_t>(16);
    hipStream_t s = ctx->stream;
    fill_i32(s, cursor, n_cols + 1, 0);
    fill_i32(s, flags, 16, 0);
    if (n_rows > 0) {
        int64_t b = std::max<int64_t>(1, std::min<int64_t>((n_rows + kRowWaves - 1) / kRowWaves,
                                                           (int64_t)ctx->num_cus * 32));
        k_tfill<T><<<(unsigned)b, kRowWaves * kWave, 0, s>>>(n_rows, row_ptr, col, W, rowT_ptr, cursor, colT, WT,
                                                              diag);
        DCPSE_LAUNCH_CHECK(ctx);
        int64_t rb = std::max<int64_t>(1, std::min<int64_t>((n_cols + kRowWaves - 1) / kRowWaves,
                                                            (int64_t)ctx->num_cus * 16));
        k_tsort<T><<<(unsigned)rb, kRowWaves * kWave, 0, s>>>(n_cols, rowT_ptr, colT, WT, flags);
        DCPSE_LAUNCH_CHECK(ctx);
    }
    int32_t ovf = 0;
    DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(&ovf, flags, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    DCPSE_HIP_CHECK(ctx, hipStreamSynchronize(s));
    if (ovf) return fail(ctx, DCPSE_ERR_RANGE, "a transposed row exceeds DCPSE_MAX_ROW=%d", DCPSE_MAX_ROW);
    return DCPSE_OK;
}

}  // namespace dcpse

using namespace dcpse;

extern "C" {

dcpse_status dcpse_support_count(dcpse_ctx* ctx, int32_t dim, const double* pos_sup, int64_t n_sup,
                                 const double* pos_dom, int64_t n_dom, const double* rcut,
                                 int32_t rcut_per_particle, const double* box_lo, const double* box_hi,
                                 uint32_t periodic_mask, int64_t* row_ptr, int64_t* nnz) {
    if (!ctx) return DCPSE_ERR_INVALID;
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return support_pass(ctx, dim, pos_sup, n_sup, pos_dom, n_dom, rcut, rcut_per_particle, box_lo, box_hi,
                        periodic_mask, row_ptr, nullptr, nnz);
}

dcpse_status dcpse_support_fill(dcpse_ctx* ctx, int32_t dim, const double* pos_sup, int64_t n_sup,
                                const double* pos_dom, int64_t n_dom, const double* rcut,
                                int32_t rcut_per_particle, const double* box_lo, const double* box_hi,
                                uint32_t periodic_mask, const int64_t* row_ptr, int32_t* col) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!col) return fail(ctx, DCPSE_ERR_INVALID, "col is null");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return support_pass(ctx, dim, pos_sup, n_sup, pos_dom, n_dom, rcut, rcut_per_particle, box_lo, box_hi,
                        periodic_mask, const_cast<int64_t*>(row_ptr), col, nullptr);
}

dcpse_status dcpse_transpose_count(dcpse_ctx* ctx, int64_t n_rows, int64_t n_cols, const int64_t* row_ptr,
                                   const int32_t* col, int64_t* rowT_ptr) {
    if (!ctx) return DCPSE_ERR_INVALID;
    if (!row_ptr || !col || !rowT_ptr || n_rows < 0 || n_cols < 0)
        return fail(ctx, DCPSE_ERR_INVALID, "null pointer or negative size");
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    int64_t nnz = 0;
    DCPSE_HIP_CHECK(ctx, hipMemcpyAsync(&nnz, row_ptr + n_rows, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    DCPSE_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    Workspace w;
    dcpse_status st = reserve_workspace(ctx, (size_t)(n_cols + 1) * 4 + scan_workspace_bytes(n_cols) + 1024, &w);
    if (st) return st;
    int32_t* cnt = w.take<int32_t>(n_cols + 1);
    fill_i32(ctx->stream, cnt, n_cols + 1, 0);
    if (nnz > 0) {
        int64_t b = std::max<int64_t>(1, std::min<int64_t>((nnz + 255) / 256, 65536));
        k_tcount<<<(unsigned)b, 256, 0, ctx->stream>>>(col, nnz, cnt);
        DCPSE_LAUNCH_CHECK(ctx);
    }
    return scan_counts(ctx, cnt, n_cols, rowT_ptr, &w);
}

dcpse_status dcpse_transpose_fill_f32(dcpse_ctx* ctx, int64_t n_rows, int64_t n_cols, const int64_t* row_ptr,
                                      const int32_t* col, const float* W, const int64_t* rowT_ptr,
                                      int32_t* colT, float* WT, float* diag) {
    if (!ctx) return DCPSE_ERR_INVALID;
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return transpose_fill<float>(ctx, n_rows, n_cols, row_ptr, col, W, rowT_ptr, colT, WT, diag);
}

dcpse_status dcpse_transpose_fill_f64(dcpse_ctx* ctx, int64_t n_rows, int64_t n_cols, const int64_t* row_ptr,
                                      const int32_t* col, const double* W, const int64_t* rowT_ptr,
                                      int32_t* colT, double* WT, double* diag) {
    if (!ctx) return DCPSE_ERR_INVALID;
    DCPSE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return transpose_fill<double>(ctx, n_rows, n_cols, row_ptr, col, W, rowT_ptr, colT, WT, diag);
}

}  // extern "C"
