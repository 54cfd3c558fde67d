// Voronoi landmark cells, ghost points and the batched per-cell epsilon query
// on one GPU.
//
//   landmarks   : selection sampling (include/utils.hpp:1-29) on the host, or
//                 farthest-point sampling on the device (exact FP32 distances)
//   assignment  : argmin over centers with ties to the later center
//                 (include/point.hpp:535-559), reference-order FP32 distances
//   cells       : members grouped by cell in ascending id, the center swapped
//                 to the front (include/point.hpp:721-754)
//   ghosts      : q is a ghost of cell b != cell(q) if L2(center_b, q) <=
//                 fl(delta_q + 2 eps) and some member p of b has L2(p, q) <= eps
//                 (include/point.hpp:894-918)
//   cell query  : the pairs the landmark driver evaluates (members x members
//                 and members x ghosts of every cell, drivers/landmark.cpp:
//                 203-210), deduplicated to undirected edges
//                 (include/graph.hpp:59-116)
//
// The cell query runs as one self join over the points in cell order (each
// cell padded to 32-point tiles) on the BF16 MFMA filter, restricted to the
// tiles of cell pairs that can hold a reference pair: p in cell a and q in
// cell b (a != b) are paired by the reference only if L2(p, q) <= eps, and so
// only if L2(c_a, c_b) <= R_a + R_b + eps with R the largest delta of a cell
// (triangle inequality) and L2(c_a, c_b) <= 2 (min(R_a, R_b) + eps) (each
// point's own center is its nearest).  Cell pairs beyond these bounds (with a
// 1e-3 relative margin, far above the FP32 distance error) are skipped.  Every evaluated pair is
// evaluated once; a pair within eps across cells is kept exactly when the
// reference keeps it (the exact FP32 center test of either side passes), and
// the passing sides are the ghosts: the ghost existence test of
// include/point.hpp:912 is the pair test itself.  When eps is large against
// the cell radii (every center test passes, as for unit vectors in 96 dims
// at mean degree 25) the pruning keeps every tile and the join is the full
// upper triangle; the reference's ghost lists then hold almost every point.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "mfma.hpp"
#include "mlp.hpp"

struct rnng_voronoi_s {
    const rnng_dataset_s* ds = nullptr;   // borrowed: must outlive the handle
    int64_t m = 0;
    std::vector<int> centers;             // center point ids, cell order
    rnng::DevBuf ctr;                     // int32 [m]
    rnng::DevBuf C;                       // float [m*d] center coordinates
    rnng::DevBuf cell;                    // int32 [n]
    rnng::DevBuf delta;                   // float [n] distance to own center
    rnng::DevBuf cptr;                    // int32 [m+1] cell CSR
    rnng::DevBuf cids;                    // int32 [n] members, coalesce order
    std::vector<int> cnt;                 // members per cell (host)
    // ghosts of the last graph build, (cell << 32 | point) sorted
    rnng::DevBuf gkeys;
    int64_t n_ghosts = 0;
    // cost-model features (m x 4) and counts (m x 2: candidates, edges) of the
    // last graph build
    rnng::DevBuf feat;
    rnng::DevBuf counts;
    // graph-build workspace, grown on demand and kept across builds
    rnng::LocalBuf rad, cnt_dev, adj_dev, slot, items, yd, counters, cand, cid, cflag, cw, tune_items;
    rnng::MfmaPrep prep;
    // filter configuration chosen by the autotuner on the first build
    int tuned_variant = -1, tuned_xcd = 1;
    int tuned_i8 = 0;   // int8 cell kernel code chosen by the tuner (0: the default, code 3 or 10)
    bool tuned_l16 = false;   // the fragment layout the configuration was chosen on
    double tune_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // schedule: row order (0 cells in landmark order, 1 cells along the
    // nearest-neighbour chain of the centers, 2 flat: the points in their own
    // order as one block, every pair in the work) and work form (0 runs of
    // column tiles, 1 tile lists, 2 tile lists on a persistent queue in order
    // of predicted cost); -1: the autotuner's choice.  forced_*: set by
    // rnng_voronoi_set_schedule.
    int order = -1, sched = -1, forced_order = -1, forced_sched = -1;
    std::vector<int> sched_cfg;           // tuner configurations (order << 8 | sched << 4 | variant slot)
    std::vector<double> sched_ms;         // and their times on every 8th work item
    std::vector<int> chain;               // positions -> cells along the chain (computed once)
    rnng::LocalBuf start, ctl, queue, cdist, nbc;
    // int8 rows of the centers (the exact integer recheck): center ids padded
    // to 32 with -1, x' rows and |x'|^2
    rnng::LocalBuf cmap, crow, chsq;
    // member chunks (cell, first member) of the feature counts, built once
    rnng::LocalBuf nbchunks;
    int64_t n_nbchunks = 0;
    bool rad_ready = false;   // rad (largest delta per cell) and cnt_dev hold this handle's cells
    // the row layout on the device (position -> cell, cell -> position, padded
    // row starts, tile -> position) and the work-list builder's workspace
    rnng::LocalBuf lay_perm, lay_pos, lay_rs, lay_tcell, flat_adj;
    rnng::LocalBuf wl_need, wl_cnt, wl_off, wl_icnt, wl_ioff, wl_stage, wl_stage2, wl_keys, wl_tot, wl_rbc;
    int use_graph = -1, autotune = -1;   // -1: from RNNG_CELL_GRAPH / RNNG_CELL_TUNE (default on)
    // captured fill pass (filter, id mapping, recheck, reference semantics,
    // per-cell counts) and the arguments it was captured with
    hipGraphExec_t fill_exec = nullptr;
    std::vector<uint64_t> fill_key;
    ~rnng_voronoi_s() {
        if (fill_exec) (void)hipGraphExecDestroy(fill_exec);
    }
};

namespace rnng {
namespace {

constexpr int AB = 128;   // rows / columns per FP32 tile
constexpr int AK = 16;    // dims staged per LDS round

__device__ __forceinline__ unsigned f2ord_v(float f) {
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f_v(unsigned u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Row index of micro-tile element r (0..7) of thread coordinate t (0..15).
__device__ __forceinline__ int micro(int t, int r) { return r < 4 ? t * 4 + r : 64 + t * 4 + r - 4; }

// Reference-order squared distances of a 128 x 128 tile: rows through rowp,
// columns through colp (nullptr for padding), 8 x 8 per thread.
template <class RowP, class ColP>
__device__ __forceinline__ void fp32_tile(int d, RowP rowp, ColP colp, float (&As)[AK][AB + 4],
                                          float (&Bs)[AK][AB + 4], float (&acc)[8][8]) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int lr = tid >> 1, lk = (tid & 1) * 8;
    const float* ra = rowp(lr);
    const float* cb = colp(lr);
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[r][c] = 0.0f;
    for (int k0 = 0; k0 < d; k0 += AK) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + lk + e;
            As[lk + e][lr] = (ra && k < d) ? ra[k] : 0.0f;
            Bs[lk + e][lr] = (cb && k < d) ? cb[k] : 0.0f;
        }
        __syncthreads();
        const int kend = min(AK, d - k0);
#pragma unroll 4
        for (int kk = 0; kk < kend; ++kk) {
            const float4 a0 = *(const float4*)&As[kk][ty * 4];
            const float4 a1 = *(const float4*)&As[kk][64 + ty * 4];
            const float4 b0 = *(const float4*)&Bs[kk][tx * 4];
            const float4 b1 = *(const float4*)&Bs[kk][64 + tx * 4];
            const float a[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
            const float b[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const float t = a[r] - b[c];
                    acc[r][c] = acc[r][c] + t * t;
                }
        }
        __syncthreads();
    }
}

// (dist, index) order of the reference's `if (dist <= best)` scan over
// ascending centers: the smaller distance wins, equal distances go to the
// larger index.
__device__ __forceinline__ bool beats(float d, int c, float bd, int bc) {
    return d < bd || (d == bd && c > bc);
}

// Exact nearest-center assignment.  One workgroup per 128 points, all centers
// streamed through in 128-center tiles.
__global__ __launch_bounds__(256) void k_assign(const float* __restrict__ X, int n,
                                                const float* __restrict__ C, int m, int d,
                                                int* __restrict__ cell, float* __restrict__ dist) {
    __shared__ float As[AK][AB + 4];
    __shared__ float Bs[AK][AB + 4];
    const int row0 = blockIdx.x * AB;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    float best[8];
    int bidx[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        best[r] = FLT_MAX;
        bidx[r] = 0;
    }
    for (int col0 = 0; col0 < m; col0 += AB) {
        float acc[8][8];
        fp32_tile(
            d, [&](int i) -> const float* { return row0 + i < n ? X + (size_t)(row0 + i) * d : nullptr; },
            [&](int j) -> const float* { return col0 + j < m ? C + (size_t)(col0 + j) * d : nullptr; }, As, Bs,
            acc);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int j = col0 + micro(tx, c);
            if (j < m) {
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const float dd = sqrtf(acc[r][c]);
                    if (beats(dd, j, best[r], bidx[r])) {
                        best[r] = dd;
                        bidx[r] = j;
                    }
                }
            }
        }
    }
    // the 16 threads of one row group are 16 consecutive lanes
#pragma unroll
    for (int r = 0; r < 8; ++r) {
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const float ob = __shfl_xor(best[r], o, 64);
            const int oi = __shfl_xor(bidx[r], o, 64);
            if (beats(ob, oi, best[r], bidx[r])) {
                best[r] = ob;
                bidx[r] = oi;
            }
        }
    }
    if (tx == 0) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int i = row0 + micro(ty, r);
            if (i < n) {
                cell[i] = bidx[r];
                dist[i] = best[r];
            }
        }
    }
}

__global__ void k_gather_rows(const float* __restrict__ X, int d, const int* __restrict__ ids, int m,
                              float* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (int64_t)m * d) {
        const int c = (int)(e / d), k = (int)(e % d);
        out[e] = X[(size_t)ids[c] * d + k];
    }
}

// ------------------------------------------------------------------------
// Farthest-point sampling
// ------------------------------------------------------------------------
// Step k: every point's distance to center k (point ctr[k]) updates its
// nearest-center distance and cell with the reference's `<=` rule, so after
// the last step (cell, mind) is the Voronoi assignment of the centers in
// selection order.  Each workgroup also writes the argmax of the updated
// distances (ties to the smaller point index) for the next selection.
//
// A point i in cell a can move to center k only if L2(c_k, i) <= mind_i,
// which needs L2(c_k, c_a) <= 2 mind_i by the triangle inequality.  Points
// with L2(c_k, c_a) >= 2 mind_i (1 + 1e-4) are skipped without reading their
// row: the margin is far above the FP32 error of the three distances
// (~(d + 2) 2^-24 relative each), so a skipped point's exact distance is > mind_i.

// Dc[a] = L2(c_k, c_a) for a < k (any summation order: only a bound), and
// the new center's row copied to C[k].
// Round-to-nearest-even BF16 bits of x (NaN stays NaN).
__device__ __forceinline__ unsigned bf16_bits(float x) {
    const unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

__global__ __launch_bounds__(256) void k_fps_center(const float* __restrict__ X, int d, const int* __restrict__ ctr,
                                                    int k, float* __restrict__ C, float* __restrict__ Dc) {
    const float* c = X + (size_t)ctr[k] * d;
    if (blockIdx.x == 0)
        for (int t = threadIdx.x; t < d; t += 256) C[(size_t)k * d + t] = c[t];
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a < k) {
        const float* ca = C + (size_t)a * d;
        float acc = 0.0f;
        for (int j = 0; j < d; ++j) {
            const float t = ca[j] - c[j];
            acc += t * t;
        }
        Dc[a] = sqrtf(acc);
    }
}

// XT is the points transposed (dimension-major, row stride ldt): lane i
// reads XT[k*ldt + i] for k = 0..d-1, so every dimension is one coalesced
// stream across the wave, and the FP32 sum keeps the reference's order.
//
// Before the exact sum, a point is screened with BF16 copies: XB packs the
// rounded coordinates two per word (dimension pairs, same layout), ex[i]
// bounds |bf16(x_i) - x_i| from above.  By the triangle inequality
// |x - c| >= |bf16(x) - bf16(c)| - ex[x] - ex[c]; when that lower bound
// exceeds md with a 2^-13 relative margin (every FP32 rounding of the screen
// and of the reference-order sum, d <= 144) and a 1e-20 absolute one (the
// reference's squares may underflow: each loses at most 2^-150, so its
// distance is at least the true one minus sqrt(d 2^-149) < 5e-22), the
// reference distance cannot be <= md and the point keeps its cell: the
// result is the exact one.
__global__ __launch_bounds__(256) void k_fps_update(const float* __restrict__ X, const float* __restrict__ XT,
                                                    const unsigned* __restrict__ XB, const float* __restrict__ ex,
                                                    int64_t ldt, int n, int d, const int* __restrict__ ctr, int k,
                                                    float* __restrict__ mind, int* __restrict__ cell,
                                                    const float* __restrict__ Dc,
                                                    unsigned long long* __restrict__ part) {
#pragma clang fp contract(off)
    extern __shared__ float fps_lds[];   // [d] center, then [(d+1)/2] its packed BF16 pairs
    float* cs = fps_lds;
    unsigned* cb = reinterpret_cast<unsigned*>(fps_lds + d);
    const int d2 = (d + 1) / 2;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int c = ctr[k];
    for (int t = tid; t < d; t += 256) cs[t] = X[(size_t)c * d + t];
    for (int t = tid; t < d2; t += 256) {
        const unsigned lo = bf16_bits(X[(size_t)c * d + 2 * t]);
        const unsigned hi = 2 * t + 1 < d ? bf16_bits(X[(size_t)c * d + 2 * t + 1]) : 0u;
        cb[t] = lo | (hi << 16);
    }
    const float ec = ex[c];
    __syncthreads();
    const int i = blockIdx.x * 256 + tid;
    float md = FLT_MAX;
    int cl = 0;
    bool need = false;
    if (i < n) {
        if (k == 0) {
            need = true;
        } else {
            md = mind[i];
            cl = cell[i];
            need = !((double)Dc[cl] >= 2.0 * (double)md * (1.0 + 1e-4) + 1e-30);
        }
    }
    if (need && k > 0) {
        float sq = 0.0f;
        const unsigned* xb = XB + i;
#pragma unroll 8
        for (int kk = 0; kk < d2; ++kk) {
            const unsigned w = xb[(size_t)kk * ldt], cw = cb[kk];
            const float a0 = __uint_as_float(w << 16) - __uint_as_float(cw << 16);
            const float a1 = __uint_as_float(w & 0xffff0000u) - __uint_as_float(cw & 0xffff0000u);
            sq = fmaf(a0, a0, sq);
            sq = fmaf(a1, a1, sq);
        }
        const float lower = sqrtf(sq) * (1.0f - 0x1p-13f) - ex[i] - ec;
        if (isfinite(sq) && lower * (1.0f - 0x1p-13f) - 1e-20f > md) need = false;
    }
    float acc = 0.0f;
    if (need) {
        const float* xt = XT + i;
#pragma unroll 8
        for (int kk = 0; kk < d; ++kk) {
            const float t = xt[(size_t)kk * ldt] - cs[kk];
            acc = acc + t * t;
        }
    }
    unsigned long long key = 0ull;
    if (i < n) {
        if (need) {
            const float dd = sqrtf(acc);
            if (dd <= md) {
                md = dd;
                cl = k;
            }
            mind[i] = md;
            cell[i] = cl;
        }
        key = ((unsigned long long)f2ord_v(md) << 32) | (unsigned)(0xffffffffu - (unsigned)i);
    }
    // block argmax: wave shuffles, then 4 waves through LDS
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long v = __shfl_xor(key, o, 64);
        key = v > key ? v : key;
    }
    __shared__ unsigned long long wk[4];
    if (lane == 0) wk[wave] = key;
    __syncthreads();
    if (tid == 0) {
        unsigned long long b = wk[0];
        for (int w = 1; w < 4; ++w) b = wk[w] > b ? wk[w] : b;
        part[blockIdx.x] = b;
    }
}

// XT[k*ldt + i] = X[i*d + k] and XB[p*ldt + i] = bf16 pair (2p, 2p+1) of
// point i: 32 x 32 tiles through LDS (coalesced both ways)
__global__ __launch_bounds__(256) void k_transpose(const float* __restrict__ X, int n, int d,
                                                   float* __restrict__ XT, unsigned* __restrict__ XB, int64_t ldt) {
    __shared__ float t[32][33];
    const int i0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // ty 0..7
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, kk = k0 + tx;
        t[r][tx] = (i < n && kk < d) ? X[(size_t)i * d + kk] : 0.0f;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int kk = k0 + r, i = i0 + tx;
        if (kk < d && i < n) XT[(size_t)kk * ldt + i] = t[tx][r];
    }
    for (int r = ty; r < 16; r += 8) {
        const int p = k0 / 2 + r, i = i0 + tx;
        if (2 * p < d && i < n) XB[(size_t)p * ldt + i] = bf16_bits(t[tx][2 * r]) | (bf16_bits(t[tx][2 * r + 1]) << 16);
    }
}

// ex[i] >= |bf16(x_i) - x_i|: each coordinate difference is exact in FP32,
// the sum runs in FP64 (no underflow of tiny differences) and the result is
// rounded up to float.
__global__ void k_bf16_error(const float* __restrict__ X, int n, int d, float* __restrict__ ex) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* x = X + (size_t)i * d;
    double s = 0.0;
    for (int k = 0; k < d; ++k) {
        const double e = (double)(__uint_as_float(bf16_bits(x[k]) << 16) - x[k]);
        s += e * e;
    }
    const double v = sqrt(s) * (1.0 + 1e-12);
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, INFINITY);
    ex[i] = f;
}

__global__ __launch_bounds__(1024) void k_fps_select(const unsigned long long* __restrict__ part, int nb,
                                                     int* __restrict__ ctr, int k) {
    unsigned long long key = 0ull;
    for (int b = threadIdx.x; b < nb; b += 1024) key = part[b] > key ? part[b] : key;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long v = __shfl_xor(key, o, 64);
        key = v > key ? v : key;
    }
    __shared__ unsigned long long wk[16];
    if ((threadIdx.x & 63) == 0) wk[threadIdx.x >> 6] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = wk[0];
        for (int w = 1; w < 16; ++w) b = wk[w] > b ? wk[w] : b;
        ctr[k + 1] = (int)(0xffffffffu - (unsigned)(b & 0xffffffffull));
    }
}

// ------------------------------------------------------------------------
// Cells
// ------------------------------------------------------------------------
__global__ void k_cell_keys(const int* __restrict__ cell, int n, uint64_t* __restrict__ keys,
                            int* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        keys[i] = ((uint64_t)(uint32_t)cell[i] << 32) | (uint32_t)i;
        atomicAdd(&cnt[cell[i]], 1);
    }
}

__global__ void k_low_words(const uint64_t* __restrict__ keys, int64_t n, int* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int)(keys[i] & 0xffffffffull);
}

// Members are in ascending id; the center (if it is a member) is swapped
// with the first member, as the reference does after gathering a cell.
__global__ void k_center_swap(const int* __restrict__ ctr, const int* __restrict__ cptr, int m,
                              int* __restrict__ ids) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m) return;
    int lo = cptr[c], hi = cptr[c + 1];
    const int first = lo, z = ctr[c];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ids[mid] < z) lo = mid + 1; else hi = mid;
    }
    if (lo < cptr[c + 1] && ids[lo] == z) {
        const int t = ids[first];
        ids[first] = z;
        ids[lo] = t;
    }
}

__global__ void k_cell_radius(const int* __restrict__ cell, const float* __restrict__ delta, int n,
                              unsigned* __restrict__ rad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicMax(&rad[cell[i]], f2ord_v(delta[i]));
}

// Cell adjacency bit matrix adj[a][b / 32] (m x words): cells a != b, both
// non-empty, with L2(c_a, c_b) <= min(R_a + R_b + eps, 2 min(R_a, R_b) +
// 2 eps)(1 + 1e-3) + 1e-15.  Any edge the reference emits joins p in cell a
// and q in cell b with L2(p, q) <= eps, so L2(c_a, c_b) <= L2(c_a, p) +
// L2(p, q) + L2(q, c_b) <= R_a + eps + R_b; and q is no nearer to c_a than to
// c_b (its cell is the nearest center), so L2(c_a, c_b) <= L2(c_a, q) +
// L2(q, c_b) <= 2 L2(c_a, q) <= 2 (R_a + eps), and likewise 2 (R_b + eps).
// The FP32 values here and in the reference's own distances
// are within a few 1e-6 relative of the true ones (sums of squares), far below
// the relative margin; the absolute one covers squares that underflow (about
// 1e-18 in a distance at the cell path's d <= 144, even with subnormals
// flushed).  This is tighter than the bound
// the ghost criterion alone gives (2 max(R_a, R_b) + 2 eps).
//
// With ADJ the adjacency bits (their 64 x 64 block of the matrix: two words
// of each of 64 rows), else the squared FP32 distances (an ordering aid) into
// out[a * m + b].  One workgroup per 64 x 64 block of center pairs, the two
// blocks' centers staged in LDS 32 dimensions at a time, each thread 4 x 4
// pairs; every pair's sum runs over the dimensions in order.
constexpr int CPB = 64;   // centers per block side
constexpr int CPK = 32;   // dimensions per LDS round
template <bool ADJ>
__global__ __launch_bounds__(256) void k_center_pairs(const float* __restrict__ C, int m, int d,
                                                      const int* __restrict__ cnt, const unsigned* __restrict__ rad,
                                                      float eps, int words, unsigned* __restrict__ adj,
                                                      float* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ float ca[CPK][CPB + 1], cb[CPK][CPB + 1];
    __shared__ unsigned bits[CPB][2];
    const int a0 = blockIdx.y * CPB, b0 = blockIdx.x * CPB;
    const int tid = threadIdx.x, ta = tid >> 4, tb = tid & 15;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
    for (int k0 = 0; k0 < d; k0 += CPK) {
        const int kn = min(CPK, d - k0);
        __syncthreads();
        for (int e = tid; e < CPB * CPK; e += 256) {
            const int r = e / CPK, k = e % CPK;
            const int a = a0 + r, b = b0 + r;
            ca[k][r] = (a < m && k < kn) ? C[(size_t)a * d + k0 + k] : 0.0f;
            cb[k][r] = (b < m && k < kn) ? C[(size_t)b * d + k0 + k] : 0.0f;
        }
        __syncthreads();
        for (int k = 0; k < kn; ++k) {
            float va[4], vb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) va[i] = ca[k][ta * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) vb[j] = cb[k][tb * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float t = va[i] - vb[j];
                    acc[i][j] = acc[i][j] + t * t;
                }
        }
    }
    if constexpr (!ADJ) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int a = a0 + ta * 4 + i, b = b0 + tb * 4 + j;
                if (a < m && b < m) out[(size_t)a * m + b] = acc[i][j];
            }
    } else {
        if (tid < 2 * CPB) bits[tid >> 1][tid & 1] = 0u;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int a = a0 + ta * 4 + i;
            if (a >= m || cnt[a] == 0) continue;
            const double ra = (double)ord2f_v(rad[a]);
            unsigned w = 0u;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int b = b0 + tb * 4 + j;
                if (b < m && b != a && cnt[b] > 0) {
                    const double rb = (double)ord2f_v(rad[b]);
                    const double bound =
                        fmin(ra + rb + (double)eps, 2.0 * (fmin(ra, rb) + (double)eps)) * (1.0 + 1e-3) + 1e-15;
                    if ((double)sqrtf(acc[i][j]) <= bound) w |= 1u << (((tb * 4) & 31) + j);
                }
            }
            if (w) atomicOr(&bits[ta * 4 + i][(tb * 4) >> 5], w);
        }
        __syncthreads();
        if (tid < 2 * CPB) {
            const int a = a0 + (tid >> 1), wd = (b0 >> 5) + (tid & 1);
            if (a < m && wd < words) adj[(size_t)a * words + wd] = bits[tid >> 1][tid & 1];
        }
    }
}

// Slot map of the flat order: slot i holds point i, the padding -1.
__global__ void k_slot_identity(int n, int n_pad, int* __restrict__ slot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_pad) slot[i] = i < n ? i : -1;
}

// Slot map of a cell layout: slots start[b] .. start[b] + n_b hold the
// members of cell b in coalesce order, the rest of its 32-padded range -1.
__global__ __launch_bounds__(256) void k_slot_map_at(const int* __restrict__ cptr, const int* __restrict__ cids,
                                                     const int* __restrict__ start, int* __restrict__ slot) {
    const int b = blockIdx.x;
    const int nb = cptr[b + 1] - cptr[b], nr = (nb + 31) & ~31;
    for (int k = threadIdx.x; k < nr; k += 256) slot[start[b] + k] = k < nb ? cids[cptr[b] + k] : -1;
}

// The fill pass's work per candidate.  Each wave takes a contiguous range of
// the candidates, RC_CELL per round, with their two point rows staged once in
// its own part of LDS:
//  * candidate slots -> point ids, written (min, max);
//  * the FP32 distance in the reference's order (sequential, not fused), the
//    edge decision acc <= T and the float weight; candidates within 1e-5 eps
//    of eps are band pairs (counters[2]), whose FP64 decision is compared
//    (flips: counters[3]);
//  * the reference's semantics for an edge (p, q) across cells
//    a = cell(p) != b = cell(q): the landmark driver emits it iff q is a
//    ghost of a (L2(c_a, q) <= fl(delta_q + 2 eps)) or p a ghost of b; the
//    flag is cleared otherwise and each ghost found is written
//    (cell << 32 | point); counters[4] = center tests, counters[5] = ghosts;
//  * per-cell (candidates, edges) by the cell of the row point p.
// Counters are kept per wave and added once: the ghost keys go through a
// per-wave LDS buffer, the per-cell counts accumulate while the row cell
// stays the same (consecutive candidates come from one row block).  Reads
// the candidate count from the device (the grid is sized by cap), so the
// pass can be captured.
constexpr int RC_CELL = 16;
static_assert(3 * RC_CELL <= 64 && (RC_CELL & (RC_CELL - 1)) == 0, "three sums per candidate in one wave");
constexpr int GBUF = 256;   // ghost keys buffered per wave
__host__ __device__ constexpr int cell_rows_ld(int d) { return ((d + 3) & ~3) + 4; }

//
// I8 (integer data, the int8 filter's preparation): one candidate per lane
// (64 per round), the three sums from the
// int8 x' rows by position (xrow, hsq) and of the centers (crow, chsq; nw8
// 16-byte words per row) as |u|^2 + |y|^2 - 2 u.y with 4-way int8 dot
// products.  Each is the exact integer the reference's FP32 sum yields on
// such data (every partial sum an integer below 2^24), so every decision and
// weight is the same as from the FP32 rows; no rows are staged, and the
// center sums are made only for edges across cells.
__device__ __forceinline__ int i8_sq_dist(const i32x4_t* __restrict__ u, const i32x4_t* __restrict__ y, int h,
                                          int nw8) {
    int dot = 0;
#pragma unroll 4
    for (int j = 0; j < nw8; ++j) {
        const i32x4_t uv = u[j], yv = y[j];
        dot = __builtin_amdgcn_sdot4(uv.x, yv.x, dot, false);
        dot = __builtin_amdgcn_sdot4(uv.y, yv.y, dot, false);
        dot = __builtin_amdgcn_sdot4(uv.z, yv.z, dot, false);
        dot = __builtin_amdgcn_sdot4(uv.w, yv.w, dot, false);
    }
    return h - 2 * dot;
}

template <bool VEC4, bool I8>
__global__ __launch_bounds__(256) void k_cell_recheck(
    const float* __restrict__ X, const float* __restrict__ C, int d, const int2* __restrict__ cand,
    unsigned long long cap, const int* __restrict__ slot, const int* __restrict__ cell,
    const float* __restrict__ delta, float T, float eps, int2* __restrict__ ids, uint8_t* __restrict__ flag,
    float* __restrict__ cw, unsigned long long* __restrict__ counters, uint64_t* __restrict__ gkeys,
    unsigned* __restrict__ cc, const i32x4_t* __restrict__ xrow, const int* __restrict__ hsq,
    const i32x4_t* __restrict__ crow, const int* __restrict__ chsq, int nw8) {
#pragma clang fp contract(off)
    extern __shared__ float rows_all[];   // [4 waves][2 * RC_CELL][ld]: p rows, then q rows
    __shared__ uint64_t gbuf_all[4][GBUF];
    const unsigned long long nc = min(counters[0], cap);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ld = cell_rows_ld(d);
    float* rows = rows_all + (size_t)wave * (2 * RC_CELL) * ld;
    uint64_t* gbuf = gbuf_all[wave];
    const double band = 1e-5 * (double)eps;
    const float two_eps = 2.0f * eps;
    // this wave's range of candidates
    const unsigned long long nw = (unsigned long long)gridDim.x * 4;
    constexpr int RC = I8 ? 64 : RC_CELL;   // candidates per round
    const unsigned long long per = ((nc + nw - 1) / nw + RC - 1) / RC * RC;
    const unsigned long long beg = ((unsigned long long)blockIdx.x * 4 + wave) * per;
    const unsigned long long end = min(nc, beg + per);
    unsigned long long tests = 0;
    int gn = 0;                              // keys in gbuf
    int run_cell = -1;                       // per-cell counts of the current row cell
    unsigned run_cand = 0, run_edges = 0;
    auto flush_ghosts = [&]() {
        unsigned long long o = 0;
        if (lane == 0 && gn) o = atomicAdd(&counters[5], (unsigned long long)gn);
        o = __shfl(o, 0, 64);
        for (int i = lane; i < gn; i += 64) gkeys[o + i] = gbuf[i];
        gn = 0;
        __builtin_amdgcn_wave_barrier();
    };
    for (unsigned long long c0 = beg; c0 < end; c0 += RC) {
        const bool live = lane < RC && c0 + lane < end;
        int p = -1, q = -1, sp = 0, sq = 0;
        if (live) {
            const int2 sl = cand[c0 + lane];
            sp = sl.x;
            sq = sl.y;
            p = slot[sp];
            q = slot[sq];
        }
        if (I8) {
        } else if (VEC4 && d <= 128) {
            // row r = 2 i + (lane >> 5), float4 column lane & 31
            const int half = lane >> 5, k4 = lane & 31, nk4 = d >> 2;
            float4 v[RC_CELL];
#pragma unroll
            for (int i = 0; i < RC_CELL; ++i) {
                const int r = 2 * i + half;
                const int src = r < RC_CELL ? r : r - RC_CELL;
                const int id = __shfl(r < RC_CELL ? p : q, src, 64);
                v[i] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (id >= 0 && k4 < nk4) v[i] = *reinterpret_cast<const float4*>(X + (size_t)id * d + 4 * k4);
            }
#pragma unroll
            for (int i = 0; i < RC_CELL; ++i) {
                const int r = 2 * i + half;
                if (k4 < nk4) *reinterpret_cast<float4*>(rows + r * ld + 4 * k4) = v[i];
            }
        } else {
            for (int r = 0; r < 2 * RC_CELL; ++r) {
                const int src = r < RC_CELL ? r : r - RC_CELL;
                const int id = __shfl(r < RC_CELL ? p : q, src, 64);
                if (id < 0) continue;
                const float* row = X + (size_t)id * d;
                for (int k = lane; k < d; k += 64) rows[r * ld + k] = row[k];
            }
        }
        __builtin_amdgcn_wave_barrier();   // this wave's staged rows (LDS is in order within a wave)
        // three sequential sums per candidate i, one lane each, one
        // instruction stream: lane i the edge distance (row p from global
        // memory against row q in LDS), lane 16 + i center a against q, lane
        // 32 + i center b against p (computed for every candidate across
        // cells, needed only for its edges)
        const int ci = lane & (RC_CELL - 1), role = lane / RC_CELL;
        const int pi = __shfl(p, ci, 64), qi = __shfl(q, ci, 64);
        const int a = live ? cell[p] : -1, b = live ? cell[q] : -1;
        const int ai = __shfl(a, ci, 64), bi = __shfl(b, ci, 64);
        (void)qi;
        float acc = 0.0f;
        if (I8) {
            if (live) acc = (float)i8_sq_dist(xrow + (size_t)sp * nw8, xrow + (size_t)sq * nw8, hsq[sp] + hsq[sq], nw8);
        } else if (pi >= 0 && role < 3 && (role == 0 || ai != bi)) {
            const float* u = role == 0 ? X + (size_t)pi * d : C + (size_t)(role == 1 ? ai : bi) * d;
            const float* y = rows + (role == 2 ? ci : RC_CELL + ci) * ld;
            if (VEC4) {
                for (int k = 0; k < d; k += 4) {
                    const float4 uv = *reinterpret_cast<const float4*>(u + k);
                    const float4 yv = *reinterpret_cast<const float4*>(y + k);
                    float t;
                    t = uv.x - yv.x; acc = acc + t * t;
                    t = uv.y - yv.y; acc = acc + t * t;
                    t = uv.z - yv.z; acc = acc + t * t;
                    t = uv.w - yv.w; acc = acc + t * t;
                }
            } else {
                for (int k = 0; k < d; ++k) {
                    const float t = u[k] - y[k];
                    acc = acc + t * t;
                }
            }
        }
        float sa = 0.0f, sb = 0.0f;
        if (!I8) {
            sa = __shfl(acc, RC_CELL + ci, 64);
            sb = __shfl(acc, 2 * RC_CELL + ci, 64);
        }
        bool edge = false, tested = false, g0 = false, g1 = false;
        if (live) {
            const float* x = rows + lane * ld;              // p
            const float* y = rows + (RC_CELL + lane) * ld;  // q
            edge = acc <= T;
            const float w = sqrtf(acc);
            if (fabs((double)w - (double)eps) <= band) {
                atomicAdd(&counters[2], 1ull);
                double s2 = 0.0;
                if (I8) {
                    s2 = (double)acc;   // the exact squared distance
                } else {
                    for (int k = 0; k < d; ++k) {
                        const double t = (double)x[k] - (double)y[k];
                        s2 += t * t;
                    }
                }
                if ((sqrt(s2) <= (double)eps) != edge) atomicAdd(&counters[3], 1ull);
            }
            if (edge && a != b) {
                tested = true;
                const float rq = delta[q] + two_eps, rp = delta[p] + two_eps;
                if (I8) {
                    // L2(c_a, q) <= L2(c_a, p) + L2(p, q) = delta_p + w up to
                    // rounding: where that bound, widened by 3 kappa (kappa =
                    // 4 (d + 8) 2^-24 covers the relative error of an FP32 sum
                    // of d squares and its square root, for every value here)
                    // and 1e-20 (underflow), is within the reference's
                    // threshold, its test passes and the sum is not made
                    const double wid = 1.0 + 12.0 * (double)(d + 8) * 5.9604644775390625e-08;
                    const double w64 = (double)w;
                    g0 = ((double)delta[p] + w64) * wid + 1e-20 <= (double)rq;
                    g1 = ((double)delta[q] + w64) * wid + 1e-20 <= (double)rp;
                    if (!g0)
                        g0 = sqrtf((float)i8_sq_dist(crow + (size_t)a * nw8, xrow + (size_t)sq * nw8,
                                                     chsq[a] + hsq[sq], nw8)) <= rq;
                    if (!g1)
                        g1 = sqrtf((float)i8_sq_dist(crow + (size_t)b * nw8, xrow + (size_t)sp * nw8,
                                                     chsq[b] + hsq[sp], nw8)) <= rp;
                } else {
                    g0 = sqrtf(sa) <= rq;
                    g1 = sqrtf(sb) <= rp;
                }
                if (!g0 && !g1) edge = false;
            }
            ids[c0 + lane] = make_int2(min(p, q), max(p, q));
            flag[c0 + lane] = edge ? 1 : 0;
            cw[c0 + lane] = w;
        }
        __builtin_amdgcn_wave_barrier();   // every lane has read the staged rows
        tests += 2ull * (unsigned long long)__popcll(__ballot(tested));
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bool gs = s ? g1 : g0;
            const unsigned long long mg = __ballot(gs);
            if (mg) {
                const int k = __popcll(mg);
                if (gn + k > GBUF) flush_ghosts();
                const int o = gn + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mg >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((unsigned)mg, 0u));
                if (gs) gbuf[o] = ((uint64_t)(uint32_t)(s ? b : a) << 32) | (uint32_t)(s ? p : q);
                gn += k;
                __builtin_amdgcn_wave_barrier();
            }
        }
        // per-cell counts: the lanes of each distinct row cell in turn
        bool pending = live;
        for (;;) {
            const unsigned long long pm = __ballot(pending);
            if (!pm) break;
            const int a0 = __shfl(a, __builtin_ctzll(pm), 64);
            const bool mine = pending && a == a0;
            const unsigned nm = (unsigned)__popcll(__ballot(mine)), ne = (unsigned)__popcll(__ballot(mine && edge));
            if (a0 != run_cell) {
                if (lane == 0 && run_cell >= 0) {
                    atomicAdd(&cc[2 * run_cell], run_cand);
                    if (run_edges) atomicAdd(&cc[2 * run_cell + 1], run_edges);
                }
                run_cell = a0;
                run_cand = run_edges = 0;
            }
            run_cand += nm;
            run_edges += ne;
            pending = pending && !mine;
        }
    }
    flush_ghosts();
    if (lane == 0) {
        if (tests) atomicAdd(&counters[4], tests);
        if (run_cell >= 0) {
            atomicAdd(&cc[2 * run_cell], run_cand);
            if (run_edges) atomicAdd(&cc[2 * run_cell + 1], run_edges);
        }
    }
}

// Keeps the first of each run of equal keys (keys sorted).
__global__ void k_unique_flags(const uint64_t* __restrict__ keys, int64_t n, uint8_t* __restrict__ flag) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) flag[e] = (e == 0 || keys[e] != keys[e - 1]) ? 1 : 0;
}

// Cost-model features of cell a:
//   f0 = log2(1 + n) / 16                   members
//   f1 = log2(1 + n/2 + paired) / 24         column points its tiles pair with
//                                            (every member of adjacent later cells)
//   f2 = log2(1 + R / eps) / 8               radius against eps
//   f3 = log2(1 + n * k) / 24                within-cell neighbour estimate: k =
//                                            mean count of members within eps of
//                                            the first min(32, n) members (exact FP32)
constexpr int FS = 32;   // sampled members per cell

// Counts of f3: one workgroup per chunk (cell, first member) of up to 256
// members of a cell, one thread per member against the cell's first
// min(FS, n) members (the samples), sixteen at a time (squared FP32 distances
// summed over the dimensions in order, <= T), the counts added to nb[cell]
// per wave.  The samples' addresses are uniform over the workgroup, so their
// values arrive through scalar loads as instruction operands.
template <bool VEC4>
__global__ __launch_bounds__(256) void k_cell_nb(const float* __restrict__ X, int d,
                                                 const int2* __restrict__ chunks, const int* __restrict__ cids,
                                                 const int* __restrict__ cptr, const int* __restrict__ cnt, float T,
                                                 unsigned* __restrict__ nb) {
#pragma clang fp contract(off)
    const int2 ch = chunks[blockIdx.x];
    const int c = __builtin_amdgcn_readfirstlane(ch.x);
    const int nc = __builtin_amdgcn_readfirstlane(cnt[c]), base = __builtin_amdgcn_readfirstlane(cptr[c]);
    const int S = min(FS, nc);
    const int i = __builtin_amdgcn_readfirstlane(ch.y) + threadIdx.x;
    unsigned cntv = 0;
    if (i < nc) {
        const float* q = X + (size_t)cids[base + i] * d;
        for (int s0 = 0; s0 < S; s0 += 16) {
            const float* smp[16];
#pragma unroll
            for (int j = 0; j < 16; ++j)
                smp[j] = X + (size_t)__builtin_amdgcn_readfirstlane(cids[base + min(s0 + j, S - 1)]) * d;
            float acc[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
            if constexpr (VEC4) {
                for (int k = 0; k < d; k += 4) {
                    const float4 qv = *(const float4*)(q + k);
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const float4 sv = *(const float4*)(smp[j] + k);
                        float t = sv.x - qv.x;
                        acc[j] = acc[j] + t * t;
                        t = sv.y - qv.y;
                        acc[j] = acc[j] + t * t;
                        t = sv.z - qv.z;
                        acc[j] = acc[j] + t * t;
                        t = sv.w - qv.w;
                        acc[j] = acc[j] + t * t;
                    }
                }
            } else {
                for (int k = 0; k < d; ++k) {
                    const float qv = q[k];
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const float t = smp[j][k] - qv;
                        acc[j] = acc[j] + t * t;
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < 16; ++j) cntv += (s0 + j < S && acc[j] <= T) ? 1u : 0u;
        }
    }
    for (int o = 32; o > 0; o >>= 1) cntv += __shfl_xor(cntv, o, 64);
    if ((threadIdx.x & 63) == 0 && cntv) atomicAdd(&nb[c], cntv);
}

// The features of cell a (one wave per cell): f1's paired points from the
// adjacency row, f3 from the counts of k_cell_nb.
__global__ __launch_bounds__(256) void k_cell_feat(const unsigned* __restrict__ adj, int words,
                                                   const int* __restrict__ cnt, const unsigned* __restrict__ rad,
                                                   const unsigned* __restrict__ nbc, float eps, int m,
                                                   float* __restrict__ feat) {
    const int a = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (a >= m) return;
    long long col = 0;
    const unsigned* row = adj + (size_t)a * words;
    for (int w = lane; w < words; w += 64) {
        unsigned bits = row[w];
        if (w == (a >> 5)) bits &= ~0u << (a & 31);      // later cells only
        else if (w < (a >> 5)) bits = 0u;
        while (bits) {
            const int b = w * 32 + __builtin_ctz(bits);
            bits &= bits - 1;
            if (b > a && b < m) col += cnt[b];
        }
    }
    for (int o = 32; o > 0; o >>= 1) col += __shfl_xor(col, o, 64);
    if (lane == 0) {
        const int n = cnt[a];
        const int S = min(FS, n);
        const double e = fmax((double)eps, 1e-30);
        const double rr = n ? (double)ord2f_v(rad[a]) : 0.0;
        const double kk = S ? (double)nbc[a] / S : 0.0;
        feat[a * 4 + 0] = (float)(log2(1.0 + n) / 16.0);
        feat[a * 4 + 1] = (float)(log2(1.0 + 0.5 * n + (double)col) / 24.0);
        feat[a * 4 + 2] = (float)(log2(1.0 + fmin(rr / e, 1e30)) / 8.0);
        feat[a * 4 + 3] = (float)(log2(1.0 + n * kk) / 24.0);
    }
}

int sort_keys_u64(uint64_t* keys, int64_t n, unsigned end_bit) {
    if (n <= 1) return RNNG_OK;
    LocalBuf alt, tmp;
    RNNG_TRY(alt.reserve((size_t)n * sizeof(uint64_t)));
    size_t tb = 0;
    RNNG_HIP_TRY(rocprim::radix_sort_keys(nullptr, tb, keys, alt.as<uint64_t>(), (size_t)n, 0, end_bit, stream()));
    RNNG_TRY(tmp.reserve(tb));
    RNNG_HIP_TRY(rocprim::radix_sort_keys(tmp.ptr, tb, keys, alt.as<uint64_t>(), (size_t)n, 0, end_bit, stream()));
    RNNG_HIP_TRY(hipMemcpyAsync(keys, alt.ptr, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, stream()));
    return RNNG_OK;
}

unsigned bits_for(int64_t v) {
    unsigned b = 1;
    while (b < 63 && (int64_t(1) << b) <= v) ++b;
    return b;
}

// Host prefix sums of per-cell counts (int32 CSR pointers).
std::vector<int> prefix(const std::vector<int>& c) {
    std::vector<int> p(c.size() + 1, 0);
    for (size_t i = 0; i < c.size(); ++i) p[i + 1] = p[i] + c[i];
    return p;
}

// Host -> device copy that has completed when it returns (the source may be a
// temporary).
template <class T>
int upload(DevBuf& buf, const std::vector<T>& v) {
    RNNG_TRY(buf.reserve(std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) {
        RNNG_HIP_TRY(hipMemcpyAsync(buf.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    }
    return RNNG_OK;
}

template <class T>
int download(std::vector<T>& v, const DevBuf& buf, size_t n) {
    v.resize(n);
    if (n) RNNG_HIP_TRY(hipMemcpyAsync(v.data(), buf.ptr, n * sizeof(T), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

// Groups the points by cell (coalesce order) once cell/delta are set.
int build_cells(rnng_voronoi_s* v) {
    const int n = (int)v->ds->n, m = (int)v->m;
    LocalBuf keys, cnt;
    RNNG_TRY(keys.reserve((size_t)std::max(n, 1) * sizeof(uint64_t)));
    RNNG_TRY(cnt.reserve((size_t)m * sizeof(int)));
    RNNG_TRY(launch_fill_u32(cnt.as<unsigned>(), m, 0u));
    if (n > 0) {
        k_cell_keys<<<blocks_for(n, 256), 256, 0, stream()>>>(v->cell.as<int>(), n, keys.as<uint64_t>(),
                                                               cnt.as<int>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    RNNG_TRY(sort_keys_u64(keys.as<uint64_t>(), n, 32 + bits_for(m)));
    RNNG_TRY(v->cids.reserve((size_t)std::max(n, 1) * sizeof(int)));
    if (n > 0) {
        k_low_words<<<blocks_for(n, 256), 256, 0, stream()>>>(keys.as<uint64_t>(), n, v->cids.as<int>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    RNNG_TRY(download(v->cnt, cnt, (size_t)m));
    RNNG_TRY(upload(v->cptr, prefix(v->cnt)));
    k_center_swap<<<blocks_for(m, 256), 256, 0, stream()>>>(v->ctr.as<int>(), v->cptr.as<int>(), m,
                                                            v->cids.as<int>());
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

int gather_centers(rnng_voronoi_s* v) {
    const int m = (int)v->m, d = v->ds->d;
    RNNG_TRY(v->C.reserve((size_t)m * d * sizeof(float)));
    k_gather_rows<<<blocks_for((int64_t)m * d, 256), 256, 0, stream()>>>(v->ds->X, d, v->ctr.as<int>(), m,
                                                                         v->C.as<float>());
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

int assign(rnng_voronoi_s* v) {
    const int n = (int)v->ds->n, m = (int)v->m, d = v->ds->d;
    RNNG_TRY(v->cell.reserve((size_t)std::max(n, 1) * sizeof(int)));
    RNNG_TRY(v->delta.reserve((size_t)std::max(n, 1) * sizeof(float)));
    if (n > 0) {
        k_assign<<<blocks_for(n, AB), 256, 0, stream()>>>(v->ds->X, n, v->C.as<float>(), m, d, v->cell.as<int>(),
                                                          v->delta.as<float>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    return RNNG_OK;
}

int fps(rnng_voronoi_s* v, int first) {
    const int n = (int)v->ds->n, m = (int)v->m, d = v->ds->d;
    RNNG_TRY(v->cell.reserve((size_t)n * sizeof(int)));
    RNNG_TRY(v->delta.reserve((size_t)n * sizeof(float)));
    const int nb = (int)blocks_for(n, 256);
    LocalBuf part;
    RNNG_TRY(part.reserve((size_t)nb * sizeof(unsigned long long)));
    RNNG_HIP_TRY(hipMemcpyAsync(v->ctr.ptr, &first, sizeof(int), hipMemcpyHostToDevice, stream()));
    const size_t lds = (size_t)(d + (d + 1) / 2) * sizeof(float);
    LocalBuf dc, xt, xb, ex;
    RNNG_TRY(v->C.reserve((size_t)m * d * sizeof(float)));
    RNNG_TRY(dc.reserve((size_t)m * sizeof(float)));
    // dimension-major FP32 and packed BF16 copies of the points for the
    // per-landmark sweeps, and each point's BF16 rounding error
    const int64_t ldt = ((int64_t)n + 63) / 64 * 64;
    RNNG_TRY(xt.reserve((size_t)ldt * d * sizeof(float)));
    RNNG_TRY(xb.reserve((size_t)ldt * ((d + 1) / 2) * sizeof(unsigned)));
    RNNG_TRY(ex.reserve((size_t)std::max(n, 1) * sizeof(float)));
    k_transpose<<<dim3(blocks_for(n, 32), blocks_for(d, 32)), 256, 0, stream()>>>(v->ds->X, n, d, xt.as<float>(),
                                                                                xb.as<unsigned>(), ldt);
    k_bf16_error<<<blocks_for(n, 256), 256, 0, stream()>>>(v->ds->X, n, d, ex.as<float>());
    for (int k = 0; k < m; ++k) {
        k_fps_center<<<blocks_for(std::max(k, 1), 256), 256, 0, stream()>>>(v->ds->X, d, v->ctr.as<int>(), k,
                                                                            v->C.as<float>(), dc.as<float>());
        k_fps_update<<<nb, 256, lds, stream()>>>(v->ds->X, xt.as<float>(), xb.as<unsigned>(), ex.as<float>(), ldt,
                                                 n, d, v->ctr.as<int>(), k,
                                                 v->delta.as<float>(), v->cell.as<int>(), dc.as<float>(),
                                                 part.as<unsigned long long>());
        if (k + 1 < m)
            k_fps_select<<<1, 1024, 0, stream()>>>(part.as<unsigned long long>(), nb, v->ctr.as<int>(), k);
    }
    RNNG_HIP_TRY(hipGetLastError());
    v->centers.resize(m);
    RNNG_HIP_TRY(hipMemcpyAsync(v->centers.data(), v->ctr.ptr, (size_t)m * sizeof(int), hipMemcpyDeviceToHost,
                                stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

int check_voronoi(const rnng_voronoi_s* v) {
    if (!v || !v->ds) {
        set_error("null voronoi handle");
        return RNNG_ERR_ARG;
    }
    return RNNG_OK;
}

// Autotuned filter configurations: (variant, XCD ordering).  Variant 17 = the
// direct-to-LDS ring with two stages of 8 column tiles shared by 12 waves (24
// row tiles per work item), 13 and 14 = two stages of 2 or 3 tiles shared by 4
// waves, 12 = three stages of 2 tiles (all 2 row tiles per wave); on the
// 16-row fragment layout their 16x16 forms 18 (of 17) and 19 (of 13).
constexpr int kTuneVariants[4] = {17, 13, 14, 12};
constexpr int kTuneVariants16[2] = {19, 18};
constexpr int kPadRows = 1536;   // multiple of every row block (256, 512, 768 rows)

// int8 kernel of the cell path: the eight-wave shapes (codes 5-9) cover
// 512-row blocks, whose column-tile unions over short cell runs cost more than
// their longer groups save (1M x 128 integer points in 4,096 cells: 43.0 ms
// against 34.7 ms for code 3), so those codes map to the four-wave code 3, or
// on the 16-row layout its 16x16 form, code 10.
int cell_i8_code(bool l16) {
    if (l16) return 10;
    const int c = filter_i8_code();
    return c > 4 ? 3 : c;
}
// the eight-wave int8 shape the tuner also times on runs
int cell_i8_wide(bool l16) { return l16 ? 9 : 7; }

int tuning_enabled() {
    static int on = -1;
    if (on < 0) {
        on = 1;
        if (const char* e = getenv("RNNG_CELL_TUNE")) on = atoi(e) ? 1 : 0;
    }
    return on;
}

int graph_enabled() {
    static int on = -1;
    if (on < 0) {
        on = 1;
        if (const char* e = getenv("RNNG_CELL_GRAPH")) on = atoi(e) ? 1 : 0;
    }
    return on;
}

// Work list of the cell-ordered self join: for every row block, the column
// tiles of its own cells from the block's first tile on and of every cell
// adjacent to one of its cells, merged into runs and cut at chunk
// boundaries; chunk-major order.  Returns the pairs the tiles cover.
int64_t cell_work(const std::vector<int>& rs, const std::vector<unsigned>& adj, int words, int m, int bt,
                  int n_rowblocks, std::vector<int4>& items) {
    // tile -> cell
    const int ntiles = rs[m] / 32;
    std::vector<int> tcell(std::max(1, ntiles));
    for (int b = 0; b < m; ++b)
        for (int t = rs[b] / 32; t < rs[b + 1] / 32; ++t) tcell[t] = b;
    std::vector<unsigned> need(words);
    std::vector<std::vector<int4>> chunks((ntiles + kFilterChunk - 1) / kFilterChunk + 1);
    int64_t covered = 0;
    for (int rb = 0; rb < n_rowblocks; ++rb) {
        const int t0 = rb * bt, t1 = std::min(ntiles, t0 + bt);
        if (t0 >= ntiles) break;
        std::fill(need.begin(), need.end(), 0u);
        for (int a = tcell[t0]; a <= tcell[t1 - 1]; ++a) {
            need[a >> 5] |= 1u << (a & 31);
            const unsigned* row = adj.data() + (size_t)a * words;
            for (int w = 0; w < words; ++w) need[w] |= row[w];
        }
        // runs of needed cells from the block's first cell on
        auto next = [&](int x, bool set) {   // first cell >= x whose bit is `set` (m if none)
            while (x < m) {
                unsigned w = need[x >> 5] ^ (set ? 0u : ~0u);
                w &= ~0u << (x & 31);
                if (w) return std::min(m, (x & ~31) + __builtin_ctz(w));
                x = (x & ~31) + 32;
            }
            return m;
        };
        int b = next(tcell[t0], true);
        while (b < m) {
            const int e = next(b, false) - 1;
            int lo = std::max(rs[b] / 32, t0), hi = rs[e + 1] / 32;
            while (lo < hi) {
                const int cut = std::min(hi, (lo / kFilterChunk + 1) * kFilterChunk);
                chunks[lo / kFilterChunk].push_back(make_int4(rb, lo, cut, 0));
                covered += (int64_t)(cut - lo) * 32 * (t1 - t0) * 32;
                lo = cut;
            }
            b = next(e + 1, true);
        }
    }
    items.clear();
    for (auto& c : chunks) items.insert(items.end(), c.begin(), c.end());
    return covered;
}

// Cells along a nearest-neighbour chain of their centers (greedy from cell
// 0, each step to the closest unvisited center), so that cells next to each
// other in the row layout are near each other: a row block spanning two
// cells then needs the column tiles of two overlapping adjacency sets.
int center_chain(rnng_voronoi_s* v, int m, int d) {
    v->chain.assign(1, 0);
    if (m <= 1) return RNNG_OK;
    RNNG_TRY(v->cdist.reserve((size_t)m * m * sizeof(float)));
    const dim3 cgrid((unsigned)((m + CPB - 1) / CPB), (unsigned)((m + CPB - 1) / CPB));
    k_center_pairs<false><<<cgrid, 256, 0, stream()>>>(v->C.as<float>(), m, d, nullptr, nullptr, 0.0f, 0, nullptr,
                                                       v->cdist.as<float>());
    RNNG_HIP_TRY(hipGetLastError());
    std::vector<float> D;
    RNNG_TRY(download(D, v->cdist, (size_t)m * m));
    std::vector<char> seen((size_t)m, 0);
    seen[0] = 1;
    int cur = 0;
    for (int k = 1; k < m; ++k) {
        const float* row = D.data() + (size_t)cur * m;
        int best = -1;
        float bd = INFINITY;
        for (int b = 0; b < m; ++b)
            if (!seen[(size_t)b] && (best < 0 || row[b] < bd)) {
                best = b;
                bd = row[b];
            }
        seen[(size_t)best] = 1;
        v->chain.push_back(best);
        cur = best;
    }
    return RNNG_OK;
}

// The row layout of one cell order: positions in order, each cell's members
// padded to 32 slots.
struct CellLayout {
    std::vector<int> perm;      // position -> cell
    std::vector<int> pos;       // cell -> position
    std::vector<int> rs;        // position prefix of padded sizes (m + 1)
    std::vector<int> start;     // cell -> first slot
    std::vector<int> tcell;     // tile -> position
    std::vector<unsigned> adj;  // adjacency between positions (host work lists only)
};

// With adj (cell adjacency), also relabels it to positions for the host
// work lists.
void make_layout(const std::vector<int>& perm, const std::vector<int>& cnt, const std::vector<unsigned>* adj,
                 int m, int words, CellLayout& L) {
    L.perm = perm;
    L.pos.assign((size_t)m, 0);
    for (int k = 0; k < m; ++k) L.pos[(size_t)perm[(size_t)k]] = k;
    std::vector<int> R((size_t)m);
    for (int k = 0; k < m; ++k) R[(size_t)k] = (cnt[(size_t)perm[(size_t)k]] + 31) & ~31;
    L.rs = prefix(R);
    L.start.resize((size_t)m);
    for (int c = 0; c < m; ++c) L.start[(size_t)c] = L.rs[(size_t)L.pos[(size_t)c]];
    L.tcell.assign((size_t)std::max(1, L.rs[(size_t)m] / 32), 0);
    for (int k = 0; k < m; ++k)
        for (int t = L.rs[(size_t)k] / 32; t < L.rs[(size_t)k + 1] / 32; ++t) L.tcell[(size_t)t] = k;
    L.adj.clear();
    if (!adj) return;
    L.adj.assign((size_t)m * words, 0u);
    for (int k = 0; k < m; ++k) {
        const unsigned* row = adj->data() + (size_t)perm[(size_t)k] * words;
        unsigned* out = L.adj.data() + (size_t)k * words;
        for (int w = 0; w < words; ++w)
            for (unsigned bits = row[w]; bits; bits &= bits - 1) {
                const int c = 32 * w + __builtin_ctz(bits);
                const int j = L.pos[(size_t)c];
                out[j >> 5] |= 1u << (j & 31);
            }
    }
}

// Tile-list work of the cell-ordered self join: for every row block, the
// column tiles of the same needed cells as cell_work (its own cells from the
// block's first tile on and every cell adjacent to one of them) concatenated
// in ctl, cut into items of at most `seg` tiles: (row block, first, end)
// index ctl.  Items in row-block order.  Returns the pairs the tiles cover.
// (The host form of the device builder below, kept as its reference.)
int64_t cell_work_tiles(const std::vector<int>& rs, const std::vector<unsigned>& adj, int words, int m, int bt,
                        int n_rowblocks, int seg, std::vector<int4>& items, std::vector<int>& ctl) {
    seg = std::min(std::max(seg, 1), kMaxListItem);   // the kernel holds an item's tiles in registers
    const int ntiles = rs[m] / 32;
    std::vector<int> tcell(std::max(1, ntiles));
    for (int b = 0; b < m; ++b)
        for (int t = rs[b] / 32; t < rs[b + 1] / 32; ++t) tcell[t] = b;
    std::vector<unsigned> need(words);
    items.clear();
    ctl.clear();
    int64_t covered = 0;
    for (int rb = 0; rb < n_rowblocks; ++rb) {
        const int t0 = rb * bt, t1 = std::min(ntiles, t0 + bt);
        if (t0 >= ntiles) break;
        std::fill(need.begin(), need.end(), 0u);
        for (int a = tcell[t0]; a <= tcell[t1 - 1]; ++a) {
            need[a >> 5] |= 1u << (a & 31);
            const unsigned* row = adj.data() + (size_t)a * words;
            for (int w = 0; w < words; ++w) need[w] |= row[w];
        }
        const int p0 = (int)ctl.size();
        for (int w = tcell[t0] >> 5; w < words; ++w) {
            unsigned bits = need[w];
            if (w == (tcell[t0] >> 5)) bits &= ~0u << (tcell[t0] & 31);
            for (; bits; bits &= bits - 1) {
                const int b = 32 * w + __builtin_ctz(bits);
                if (b >= m) break;
                for (int t = std::max(rs[b] / 32, t0); t < rs[b + 1] / 32; ++t) ctl.push_back(t);
            }
        }
        const int p1 = (int)ctl.size();
        covered += (int64_t)(p1 - p0) * 32 * (t1 - t0) * 32;
        for (int p = p0; p < p1; p += seg) items.push_back(make_int4(rb, p, std::min(p1, p + seg), 0));
    }
    return covered;
}

// Cost of a tile-list item of `tiles` column tiles whose row block has
// rb_tiles tiles and rb_cand predicted candidates (float arithmetic, the same
// on the host and the device).
__host__ __device__ inline float item_cost(int tiles, float rb_cand, int rb_tiles) {
    return rb_tiles > 0 ? (float)tiles + rb_cand * (float)tiles / (float)rb_tiles / 64.0f : (float)tiles;
}

// Items of a persistent queue, heaviest first: cost = column tiles + the
// item's share (by tiles) of its row block's predicted candidates / 64 (a
// tile's candidate path costs about as much as the tile itself at 64
// candidates).  rb_cand[rb]: the cost model's candidate estimate of row block
// rb (zeros without a model: longest items first).
void order_by_cost(std::vector<int4>& items, const std::vector<float>& rb_cand) {
    std::vector<int> rb_tiles(rb_cand.size(), 0);
    for (const int4& w : items)
        if ((size_t)w.x < rb_tiles.size()) rb_tiles[(size_t)w.x] += w.z - w.y;
    std::vector<std::pair<float, int>> key(items.size());
    for (size_t i = 0; i < items.size(); ++i) {
        const int4 w = items[i];
        key[i] = {-item_cost(w.z - w.y, (size_t)w.x < rb_cand.size() ? rb_cand[(size_t)w.x] : 0.0f,
                             (size_t)w.x < rb_tiles.size() ? rb_tiles[(size_t)w.x] : 0),
                  (int)i};
    }
    std::stable_sort(key.begin(), key.end());
    std::vector<int4> out(items.size());
    for (size_t i = 0; i < items.size(); ++i) out[i] = items[(size_t)key[i].second];
    items.swap(out);
}

// ---- the cell filter's work lists, built on the device ----
// The same lists as cell_work (runs) and cell_work_tiles (tile lists), from
// the adjacency on the device and the layout's arrays: perm (position ->
// cell), pos (cell -> position), rs (padded row starts by position, m + 1),
// tcell (tile -> position).

// Needed positions of row block rb: the cells of its tiles and every cell
// adjacent to one of them, from the block's first cell on (position space).
__global__ __launch_bounds__(256) void k_rb_need(const unsigned* __restrict__ adj, int words, int m,
                                                 const int* __restrict__ perm, const int* __restrict__ pos,
                                                 const int* __restrict__ tcell, int ntiles, int bt,
                                                 unsigned* __restrict__ need) {
    extern __shared__ unsigned nbits[];   // [words] by cell, [words] by position
    unsigned* ucell = nbits;
    unsigned* upos = nbits + words;
    const int rb = blockIdx.x;
    const int t0 = rb * bt, t1 = min(ntiles, t0 + bt);
    const int k0 = tcell[t0], k1 = tcell[t1 - 1];
    for (int w = threadIdx.x; w < words; w += 256) {
        unsigned u = 0u;
        for (int k = k0; k <= k1; ++k) u |= adj[(size_t)perm[k] * words + w];
        ucell[w] = u;
        upos[w] = 0u;
    }
    __syncthreads();
    for (int w = threadIdx.x; w < words; w += 256)
        for (unsigned bits = ucell[w]; bits; bits &= bits - 1) {
            const int c = 32 * w + __builtin_ctz(bits);
            if (c < m) {
                const int k = pos[c];
                atomicOr(&upos[k >> 5], 1u << (k & 31));
            }
        }
    for (int k = k0 + (int)threadIdx.x; k <= k1; k += 256) atomicOr(&upos[k >> 5], 1u << (k & 31));
    __syncthreads();
    for (int w = threadIdx.x; w < words; w += 256) {
        unsigned u = upos[w];
        if (w < (k0 >> 5)) u = 0u;
        else if (w == (k0 >> 5)) u &= ~0u << (k0 & 31);
        need[(size_t)rb * words + w] = u;
    }
}

// The pieces of a row block's work in column chunk c (kFilterChunk tiles):
// its runs of consecutive needed positions as tile ranges [lo, hi), clipped
// to the chunk and to the block's first tile t0.
template <class F>
__device__ __forceinline__ void chunk_pieces(const unsigned* __restrict__ need, const int* __restrict__ rs,
                                             const int* __restrict__ tcell, int ntiles, int t0, int c, F&& emit) {
    const int lo_t = c * kFilterChunk, hi_t = min(ntiles, lo_t + kFilterChunk);
    const int tl = max(lo_t, t0);
    if (tl >= hi_t) return;
    const int kS = tcell[tl], kE = tcell[hi_t - 1];
    int lo = -1, hi = -1;
    for (int k = kS; k <= kE; ++k) {
        if ((need[k >> 5] >> (k & 31)) & 1u) {
            if (lo < 0) lo = hi = max(rs[k] >> 5, tl);
            hi = max(hi, min(rs[k + 1] >> 5, hi_t));
        } else {
            if (lo >= 0 && hi > lo) emit(lo, hi);
            lo = -1;
        }
    }
    if (lo >= 0 && hi > lo) emit(lo, hi);
}

// Pair i of (chunk, row block): runs (tiles_form 0, i = c * nrb + rb) count
// their pieces, tile lists (i = rb * nch + c) their tiles; counts[P] = 0 for
// the scan.  tot[0] += the pairs the pieces cover.
__global__ __launch_bounds__(256) void k_piece_counts(const unsigned* __restrict__ need, int words,
                                                      const int* __restrict__ rs, const int* __restrict__ tcell,
                                                      int ntiles, int bt, int nrb, int nch, int tiles_form,
                                                      int64_t* __restrict__ counts,
                                                      unsigned long long* __restrict__ tot) {
    __shared__ unsigned long long part[4];
    const int64_t P = (int64_t)nrb * nch;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    unsigned long long cov = 0;
    if (i < P) {
        const int rb = tiles_form ? (int)(i / nch) : (int)(i % nrb);
        const int c = tiles_form ? (int)(i % nch) : (int)(i / nrb);
        const int t0 = rb * bt, rows = (min(ntiles, t0 + bt) - t0) * 32;
        int64_t pieces = 0, tiles = 0;
        chunk_pieces(need + (size_t)rb * words, rs, tcell, ntiles, t0, c, [&](int lo, int hi) {
            ++pieces;
            tiles += hi - lo;
        });
        counts[i] = tiles_form ? tiles : pieces;
        cov = (unsigned long long)tiles * 32ull * (unsigned long long)rows;
    } else if (i == P) {
        counts[i] = 0;
    }
    for (int o = 32; o > 0; o >>= 1) cov += __shfl_xor(cov, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cov;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long b = part[0] + part[1] + part[2] + part[3];
        if (b) atomicAdd(&tot[0], b);
    }
}

// Writes the pieces of pair i at off[i]: runs as items (rb, lo, hi, 0), tile
// lists as their tiles in ctl.
__global__ __launch_bounds__(256) void k_piece_write(const unsigned* __restrict__ need, int words,
                                                     const int* __restrict__ rs, const int* __restrict__ tcell,
                                                     int ntiles, int bt, int nrb, int nch, int tiles_form,
                                                     const int64_t* __restrict__ off, int4* __restrict__ items,
                                                     int* __restrict__ ctl) {
    const int64_t P = (int64_t)nrb * nch;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= P) return;
    const int rb = tiles_form ? (int)(i / nch) : (int)(i % nrb);
    const int c = tiles_form ? (int)(i % nch) : (int)(i / nrb);
    int64_t o = off[i];
    chunk_pieces(need + (size_t)rb * words, rs, tcell, ntiles, rb * bt, c, [&](int lo, int hi) {
        if (tiles_form) {
            for (int t = lo; t < hi; ++t) ctl[o++] = t;
        } else {
            items[o++] = make_int4(rb, lo, hi, 0);
        }
    });
}

// Tile lists: items per row block (its tiles cut into kMaxListItem), icnt[nrb] = 0.
__global__ void k_list_item_counts(const int64_t* __restrict__ off, int nrb, int nch, int64_t* __restrict__ icnt) {
    const int rb = blockIdx.x * blockDim.x + threadIdx.x;
    if (rb > nrb) return;
    if (rb == nrb) {
        icnt[rb] = 0;
        return;
    }
    const int64_t T = off[(int64_t)(rb + 1) * nch] - off[(int64_t)rb * nch];
    icnt[rb] = (T + kMaxListItem - 1) / kMaxListItem;
}

// Tile lists: the items of row block rb at ioff[rb]; with keys, the sort key
// of the persistent queue's order (cost descending, then list order).
__global__ void k_list_items(const int64_t* __restrict__ off, const int64_t* __restrict__ ioff, int nrb, int nch,
                             const float* __restrict__ rbc, int4* __restrict__ items, uint64_t* __restrict__ keys) {
    const int rb = blockIdx.x * blockDim.x + threadIdx.x;
    if (rb >= nrb) return;
    const int64_t b0 = off[(int64_t)rb * nch];
    const int T = (int)(off[(int64_t)(rb + 1) * nch] - b0);
    int64_t it = ioff[rb];
    for (int p = 0; p < T; p += kMaxListItem, ++it) {
        const int e = min(T, p + kMaxListItem);
        items[it] = make_int4(rb, (int)(b0 + p), (int)(b0 + e), 0);
        if (keys) {
            const float cost = item_cost(e - p, rbc ? rbc[rb] : 0.0f, T);
            keys[it] = ((uint64_t)~__float_as_uint(cost) << 32) | (uint32_t)it;
        }
    }
}

__global__ void k_gather_items(const int4* __restrict__ src, const uint64_t* __restrict__ keys, int64_t n,
                               int4* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[(uint32_t)keys[i]];
}

// The launch list: every stride-th of the n items, in XCD order when xcd
// (workgroup b runs on XCD b % 8 and takes item (b % 8) * per + b / 8 of the
// selection, per = ceil(M / 8)); n_out slots, the ones past the selection
// empty items.
__global__ void k_items_select(const int4* __restrict__ src, int64_t n, int stride, int xcd, int64_t n_out,
                               int4* __restrict__ dst) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_out) return;
    const int64_t M = (n + stride - 1) / stride;
    const int64_t sel = xcd ? (b % 8) * ((M + 7) / 8) + b / 8 : b;
    dst[b] = sel < M ? src[sel * stride] : make_int4(0, 0, 0, 0);
}

int64_t launch_list_size(int64_t n, int stride, int xcd) {
    const int64_t M = (n + stride - 1) / stride;
    return xcd ? 8 * ((M + 7) / 8) : M;
}

// Builds the launch list of schedule sched (0 runs, 1 tile lists, 2 tile
// lists by cost) for row blocks of bt tiles into out (n_out items; the tile
// lists' tiles into v->ctl): every stride-th item, in XCD order if xcd (runs
// and plain tile lists).  rbc: per row block predicted candidates (sched 2).
// *covered = the pairs of every item (before the stride).
int device_work_list(rnng_voronoi_s* v, const unsigned* adj, int m, int words, int ntiles, int bt, int sched,
                     int xcd, int stride, const std::vector<float>& rbc, DevBuf& out, int64_t* n_out,
                     int64_t* covered) {
    const int nrb = (ntiles + bt - 1) / bt;
    const int nch = (ntiles + kFilterChunk - 1) / kFilterChunk;
    const int64_t P = (int64_t)nrb * nch;
    const int tiles_form = sched > 0 ? 1 : 0;
    RNNG_TRY(v->wl_need.reserve((size_t)nrb * words * sizeof(unsigned)));
    RNNG_TRY(v->wl_cnt.reserve((size_t)(P + 1) * sizeof(int64_t)));
    RNNG_TRY(v->wl_off.reserve((size_t)(P + 1) * sizeof(int64_t)));
    RNNG_TRY(v->wl_tot.reserve(4 * sizeof(int64_t)));
    k_rb_need<<<nrb, 256, (size_t)2 * words * sizeof(unsigned), stream()>>>(
        adj, words, m, v->lay_perm.as<int>(), v->lay_pos.as<int>(), v->lay_tcell.as<int>(),
        ntiles, bt, v->wl_need.as<unsigned>());
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_TRY(launch_fill_u64(v->wl_tot.as<unsigned long long>(), 1, 0ull));
    k_piece_counts<<<(unsigned)((P + 1 + 255) / 256), 256, 0, stream()>>>(
        v->wl_need.as<unsigned>(), words, v->lay_rs.as<int>(), v->lay_tcell.as<int>(), ntiles, bt, nrb, nch,
        tiles_form, v->wl_cnt.as<int64_t>(), v->wl_tot.as<unsigned long long>());
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_TRY(exclusive_scan_i64(v->wl_cnt.as<int64_t>(), v->wl_off.as<int64_t>(), P + 1));
    if (tiles_form) {
        RNNG_TRY(v->wl_icnt.reserve((size_t)(nrb + 1) * sizeof(int64_t)));
        RNNG_TRY(v->wl_ioff.reserve((size_t)(nrb + 1) * sizeof(int64_t)));
        k_list_item_counts<<<(nrb + 1 + 255) / 256, 256, 0, stream()>>>(v->wl_off.as<int64_t>(), nrb, nch,
                                                                          v->wl_icnt.as<int64_t>());
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(exclusive_scan_i64(v->wl_icnt.as<int64_t>(), v->wl_ioff.as<int64_t>(), nrb + 1));
    }
    // totals: pieces or tiles, tile-list items, covered pairs
    int64_t tot[3] = {0, 0, 0};
    RNNG_HIP_TRY(hipMemcpyAsync(&tot[0], v->wl_off.as<int64_t>() + P, sizeof(int64_t), hipMemcpyDeviceToHost,
                                stream()));
    if (tiles_form)
        RNNG_HIP_TRY(hipMemcpyAsync(&tot[1], v->wl_ioff.as<int64_t>() + nrb, sizeof(int64_t), hipMemcpyDeviceToHost,
                                    stream()));
    RNNG_HIP_TRY(hipMemcpyAsync(&tot[2], v->wl_tot.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    *covered = tot[2];
    if (tiles_form && tot[0] > INT32_MAX) {
        set_error("cell tile lists: more than 2^31 tiles");
        return RNNG_ERR_UNSUPPORTED;
    }
    const int64_t n_items = tiles_form ? tot[1] : tot[0];
    RNNG_TRY(v->wl_stage.reserve((size_t)std::max<int64_t>(1, n_items) * sizeof(int4)));
    if (tiles_form) RNNG_TRY(v->ctl.reserve((size_t)std::max<int64_t>(1, tot[0]) * sizeof(int)));
    k_piece_write<<<(unsigned)((P + 255) / 256), 256, 0, stream()>>>(
        v->wl_need.as<unsigned>(), words, v->lay_rs.as<int>(), v->lay_tcell.as<int>(), ntiles, bt, nrb, nch,
        tiles_form, v->wl_off.as<int64_t>(), v->wl_stage.as<int4>(), v->ctl.as<int>());
    RNNG_HIP_TRY(hipGetLastError());
    const DevBuf* src = &v->wl_stage;
    if (tiles_form) {
        uint64_t* keys = nullptr;
        if (sched == 2) {
            RNNG_TRY(v->wl_keys.reserve((size_t)std::max<int64_t>(1, n_items) * sizeof(uint64_t)));
            RNNG_TRY(upload(v->wl_rbc, rbc));
            keys = v->wl_keys.as<uint64_t>();
        }
        k_list_items<<<(nrb + 255) / 256, 256, 0, stream()>>>(v->wl_off.as<int64_t>(), v->wl_ioff.as<int64_t>(), nrb,
                                                               nch, rbc.empty() ? nullptr : v->wl_rbc.as<float>(),
                                                               v->wl_stage.as<int4>(), keys);
        RNNG_HIP_TRY(hipGetLastError());
        if (sched == 2 && n_items > 0) {
            RNNG_TRY(sort_keys_u64(keys, n_items, 64));
            RNNG_TRY(v->wl_stage2.reserve((size_t)n_items * sizeof(int4)));
            k_gather_items<<<(unsigned)((n_items + 255) / 256), 256, 0, stream()>>>(
                v->wl_stage.as<int4>(), keys, n_items, v->wl_stage2.as<int4>());
            RNNG_HIP_TRY(hipGetLastError());
            src = &v->wl_stage2;
        }
    }
    const int64_t no = launch_list_size(n_items, stride, sched == 2 ? 0 : xcd);
    RNNG_TRY(out.reserve((size_t)std::max<int64_t>(1, no) * sizeof(int4)));
    if (no > 0) {
        k_items_select<<<(unsigned)((no + 255) / 256), 256, 0, stream()>>>(src->as<int4>(), n_items, stride,
                                                                          sched == 2 ? 0 : xcd, no, out.as<int4>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    if (!tiles_form) RNNG_TRY(v->ctl.reserve(sizeof(int)));
    *n_out = no;
    return RNNG_OK;
}

int voronoi_graph(rnng_voronoi_s* v, float eps, rnng_graph_s* g, rnng_stats* st) {
    const int n = (int)v->ds->n, m = (int)v->m, d = v->ds->d;
    const int KS = mfma_ksteps(d);
    if (KS > kMaxKS) {
        set_error("MFMA mode supports d <= 144");
        return RNNG_ERR_UNSUPPORTED;
    }
    const float T = sq_threshold_host(eps);
    const float eps_band = (float)((double)eps * (1.0 + 1e-5));
    float Tb = sq_threshold_host(std::nextafter(eps_band, INFINITY));
    if (Tb < T) Tb = T;

    Events ev;
    RNNG_TRY(ev.create());
    RNNG_HIP_TRY(hipEventRecord(ev.e[0], stream()));
    // coordinates the cell filter (and the cost model's features) can use
    RNNG_TRY(mfma_input_check(v->prep, v->ds->X, n, d));
    // ---- cell adjacency ----
    const int words = (m + 31) / 32;
    LocalBuf& rad = v->rad;
    LocalBuf& cnt_dev = v->cnt_dev;
    LocalBuf& adj_dev = v->adj_dev;
    RNNG_TRY(adj_dev.reserve((size_t)m * words * sizeof(unsigned)));
    if (!v->rad_ready) {   // cell radii and sizes: fixed per handle
        RNNG_TRY(rad.reserve((size_t)m * sizeof(unsigned)));
        RNNG_TRY(launch_fill_u32(rad.as<unsigned>(), m, 0u));
        k_cell_radius<<<blocks_for(n, 256), 256, 0, stream()>>>(v->cell.as<int>(), v->delta.as<float>(), n,
                                                                rad.as<unsigned>());
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(upload(cnt_dev, v->cnt));
        v->rad_ready = true;
    }
    {
        const dim3 cgrid((unsigned)((m + CPB - 1) / CPB), (unsigned)((m + CPB - 1) / CPB));
        k_center_pairs<true><<<cgrid, 256, 0, stream()>>>(v->C.as<float>(), m, d, cnt_dev.as<int>(),
                                                          rad.as<unsigned>(), eps, words, adj_dev.as<unsigned>(),
                                                          nullptr);
    }
    RNNG_HIP_TRY(hipGetLastError());
    if (words > kMaxCellWords) {
        set_error("Voronoi graph: at most 32 * 20000 cells");
        return RNNG_ERR_UNSUPPORTED;
    }
    LocalBuf& slot = v->slot;
    LocalBuf& d_items = v->items;
    // ---- cost model: per-cell features -> predicted buffer sizes ----
    RNNG_TRY(v->feat.reserve((size_t)m * 4 * sizeof(float)));
    RNNG_TRY(v->nbc.reserve((size_t)m * sizeof(unsigned)));
    RNNG_TRY(launch_fill_u32(v->nbc.as<unsigned>(), m, 0u));
    const bool xv4 = d % 4 == 0 && ((uintptr_t)v->ds->X % 16) == 0;
    {
        // chunks of up to 256 members of a cell (the cells are fixed per handle)
        if (v->n_nbchunks == 0) {
            std::vector<int2> ch;
            for (int a = 0; a < m; ++a)
                for (int o = 0; o < v->cnt[(size_t)a]; o += 256) ch.push_back(make_int2(a, o));
            RNNG_TRY(upload(v->nbchunks, ch));
            v->n_nbchunks = (int64_t)ch.size();
        }
        if (v->n_nbchunks > 0) {
            auto kern = xv4 ? k_cell_nb<true> : k_cell_nb<false>;
            kern<<<(unsigned)v->n_nbchunks, 256, 0, stream()>>>(v->ds->X, d, v->nbchunks.as<int2>(),
                                                                v->cids.as<int>(), v->cptr.as<int>(),
                                                                cnt_dev.as<int>(), T, v->nbc.as<unsigned>());
        }
    }
    RNNG_HIP_TRY(hipGetLastError());
    k_cell_feat<<<(m + 3) / 4, 256, 0, stream()>>>(adj_dev.as<unsigned>(), words, cnt_dev.as<int>(),
                                                   rad.as<unsigned>(), v->nbc.as<unsigned>(), eps, m,
                                                   v->feat.as<float>());
    RNNG_HIP_TRY(hipGetLastError());
    const float* model = cost_model_device();
    double pred_cand = 0.0, pred_edges = 0.0;
    std::vector<double> pc((size_t)m, 0.0);   // predicted candidates per cell (0 without a model)
    if (model) {
        LocalBuf& yd = v->yd;
        RNNG_TRY(yd.reserve((size_t)m * 2 * sizeof(float)));
        RNNG_TRY(mlp_infer_device(model, v->feat.as<float>(), m, yd.as<float>()));
        std::vector<float> y;
        RNNG_TRY(download(y, yd, (size_t)m * 2));
        for (int a = 0; a < m; ++a) {
            pc[(size_t)a] = std::exp2(24.0 * std::min(std::max((double)y[2 * a], 0.0), 1.5)) - 1.0;
            pred_cand += pc[(size_t)a];
            pred_edges += std::exp2(24.0 * std::min(std::max((double)y[2 * a + 1], 0.0), 1.5)) - 1.0;
        }
        // a non-finite output (NaN passes the clamps) sizes nothing: count-then-fill
        if (!std::isfinite(pred_cand) || !std::isfinite(pred_edges)) {
            model = nullptr;
            pred_cand = pred_edges = 0.0;
            std::fill(pc.begin(), pc.end(), 0.0);
        }
    }

    RNNG_HIP_TRY(hipEventRecord(ev.e[1], stream()));
    // ---- cell order: the row layout, its slot map and the fragments ----
    MfmaPrep& P = v->prep;
    CellLayout L;
    int n_pad = 0, layout_order = -1;
    // blocks of the current layout (m cells, or the one block of the flat
    // order) and their adjacency
    int lm = m, lwords = words;
    const unsigned* ladj = adj_dev.as<unsigned>();
    auto apply_layout = [&](int order) -> int {
        if (order == layout_order) return RNNG_OK;
        if (order == 2) {
            // flat: every pair of the points in their own order (one block
            // adjacent to itself)
            make_layout(std::vector<int>{0}, std::vector<int>{n}, nullptr, 1, 1, L);
            lm = 1;
            lwords = 1;
            RNNG_TRY(upload(v->flat_adj, std::vector<unsigned>{1u}));
            ladj = v->flat_adj.as<unsigned>();
        } else {
            std::vector<int> perm((size_t)m);
            if (order == 1) {
                if ((int)v->chain.size() != m) RNNG_TRY(center_chain(v, m, d));
                perm = v->chain;
            } else {
                std::iota(perm.begin(), perm.end(), 0);
            }
            make_layout(perm, v->cnt, nullptr, m, words, L);
            lm = m;
            lwords = words;
            ladj = adj_dev.as<unsigned>();
        }
        n_pad = std::max(kPadRows, (L.rs[lm] + kPadRows - 1) / kPadRows * kPadRows);
        RNNG_TRY(upload(v->lay_perm, L.perm));
        RNNG_TRY(upload(v->lay_pos, L.pos));
        RNNG_TRY(upload(v->lay_rs, L.rs));
        RNNG_TRY(upload(v->lay_tcell, L.tcell));
        RNNG_TRY(slot.reserve((size_t)n_pad * sizeof(int)));
        if (order == 2) {
            k_slot_identity<<<blocks_for(n_pad, 256), 256, 0, stream()>>>(n, n_pad, slot.as<int>());
        } else {
            RNNG_TRY(upload(v->start, L.start));
            RNNG_TRY(launch_fill_u32(slot.as<unsigned>(), n_pad, 0xffffffffu));
            k_slot_map_at<<<m, 256, 0, stream()>>>(v->cptr.as<int>(), v->cids.as<int>(), v->start.as<int>(),
                                                   slot.as<int>());
        }
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(mfma_prepare_mapped(P, v->ds->X, n, d, Tb, slot.as<int>(), n_pad, slot.as<int>(), n_pad));
        if (P.i8) {
            const int m_pad = (m + 31) / 32 * 32;
            std::vector<int> cm((size_t)m_pad, -1);
            std::copy(v->centers.begin(), v->centers.end(), cm.begin());
            RNNG_TRY(upload(v->cmap, cm));
            RNNG_TRY(v->crow.reserve((size_t)m_pad * P.KS8 * 32));
            RNNG_TRY(v->chsq.reserve((size_t)m_pad * sizeof(int)));
            RNNG_TRY(mfma_i8_rows(P, v->ds->X, n, d, v->cmap.as<int>(), m_pad, v->chsq.as<int>(),
                                  v->crow.as<i32x4_t>()));
        }
        layout_order = order;
        return RNNG_OK;
    };
    RNNG_TRY(apply_layout(v->order >= 0 ? v->order : 0));
    RNNG_HIP_TRY(hipEventRecord(ev.e[2], stream()));
    const auto host0 = std::chrono::steady_clock::now();

    LocalBuf& counters = v->counters;
    RNNG_TRY(counters.reserve(8 * sizeof(unsigned long long)));
    RNNG_TRY(v->queue.reserve(sizeof(unsigned)));
    auto filter_args = [&](const DevBuf& work, size_t n_items, int2* cand_ptr, unsigned long long cap_) {
        FilterArgs fa;
        fa.xa = P.xa.as<bf16x8_t>();
        fa.yb = P.yb.as<bf16x8_t>();
        fa.xn = P.xn.as<float>();
        fa.ycn = P.ycn.as<float>();
        fa.work = work.as<int4>();
        fa.n_items = (unsigned)n_items;
        fa.n_rows = n_pad;
        fa.n_cols = n_pad;
        fa.self_join = 1;
        fa.prm = P.params.as<MfmaParams>();
        fa.cand = cand_ptr;
        fa.cap = cap_;
        fa.counter = counters.as<unsigned long long>();
        return fa;
    };
    // the filter over a work list: the exact int8 kernel when the prep built
    // int8 forms, else the BF16 variant var; runs of column tiles (sched 0)
    // or tile lists ctl (1; 2 on a persistent queue)
    auto run_filter = [&](int sched_, const DevBuf& work, size_t n_items, const DevBuf& ctl_, int2* cand_ptr,
                          unsigned long long cap_, int var) -> int {
        if (sched_ > 0 || sched_ == -2) {
            TileFilterArgs ta;
            ta.dry_run = sched_ == -2;
            ta.i8 = P.i8;
            ta.xa = P.xa.ptr;
            ta.yb = P.yb.ptr;
            ta.rterm = P.xn.ptr;
            ta.cterm = P.ycn.ptr;
            ta.work = work.as<int4>();
            ta.n_items = (unsigned)n_items;
            ta.ctl = ctl_.as<int>();
            ta.persistent = sched_ == 2 || sched_ == -2;
            ta.queue = v->queue.as<unsigned>();
            ta.n_rows = n_pad;
            ta.n_cols = n_pad;
            ta.self_join = 1;
            ta.cand = cand_ptr;
            ta.cap = cap_;
            ta.counter = counters.as<unsigned long long>();
            ta.ks = P.i8 ? P.KS8 : KS;
            ta.variant = var;
            ta.l16 = P.l16;
            return filter_dispatch_tiles(ta);
        }
        if (P.i8) {
            I8Args ia;
            ia.xa = P.xa.as<i32x4_t>();
            ia.yb = P.yb.as<i32x4_t>();
            ia.rterm = P.xn.as<int>();
            ia.cterm = P.ycn.as<int>();
            ia.work = work.as<int4>();
            ia.n_items = (unsigned)n_items;
            ia.n_rows = n_pad;
            ia.n_cols = n_pad;
            ia.self_join = 1;
            ia.cand = cand_ptr;
            ia.cap = cap_;
            ia.counter = counters.as<unsigned long long>();
            ia.l16 = P.l16;
            return filter_dispatch_i8(ia, P.KS8, var > 0 ? var : cell_i8_code(P.l16));
        }
        FilterArgs fa = filter_args(work, n_items, cand_ptr, cap_);
        fa.l16 = P.l16;
        return filter_dispatch(fa, KS, var);
    };
    // predicted candidates of every row block (the persistent queue's order)
    auto rb_cands = [&](int bt) {
        std::vector<float> out((size_t)((L.rs[lm] + bt * 32 - 1) / (bt * 32)), 0.0f);
        std::vector<double> acc(out.size(), 0.0);
        for (int k = 0; k < lm && lm == m; ++k) {
            const int c = L.perm[(size_t)k];
            const int tiles = (L.rs[k + 1] - L.rs[k]) / 32;
            for (int t = L.rs[k] / 32; t < L.rs[k + 1] / 32; ++t) acc[(size_t)(t / bt)] += pc[(size_t)c] / tiles;
        }
        for (size_t i = 0; i < out.size(); ++i) out[i] = (float)acc[i];
        return out;
    };
    // the launch list of a schedule on the current layout, built on the device
    auto work_list = [&](int sched_, int var, int xcd, int stride, DevBuf& out, int64_t* n_out) -> int64_t {
        const int bt = P.i8 ? filter_i8_row_tiles(var > 0 ? var : cell_i8_code(P.l16)) : filter_block_tiles(var);
        int64_t cov = 0;
        const std::vector<float> rbc = sched_ == 2 ? rb_cands(bt) : std::vector<float>();
        const int rc =
            device_work_list(v, ladj, lm, lwords, L.rs[lm] / 32, bt, sched_, xcd, stride, rbc, out, n_out, &cov);
        return rc != RNNG_OK ? -1 : cov;
    };
    // ---- autotuner: cell order x work form x filter shape, on every 8th item ----
    // (again when the kernels in use changed the fragment layout)
    if (v->tuned_variant >= 0 && v->tuned_l16 != P.l16) v->tuned_variant = -1;
    if (v->tuned_variant < 0) {
        v->tuned_l16 = P.l16;
        // the default variant in the layout the prep took (its 32x32 form for
        // an odd k-step count)
        (void)l16_variant(filter_default_variant(), KS, &v->tuned_variant);
        v->tuned_i8 = 0;
        v->tuned_xcd = 1;
        v->order = v->forced_order >= 0 ? v->forced_order : 0;
        v->sched = v->forced_sched >= 0 ? v->forced_sched : 0;
        const bool forced = v->forced_order >= 0 || v->forced_sched >= 0;
        if (!forced && (v->autotune >= 0 ? v->autotune : tuning_enabled())) {
            struct Cfg {
                int order, sched, var, xcd;
            };
            std::vector<Cfg> cfgs;
            for (int order = 0; order < 2; ++order) {
                if (P.i8) {
                    // the four-wave kernel (256-row blocks) on every form, the
                    // eight-wave one (code 7 or 9, 512-row blocks) on runs
                    cfgs.push_back({order, 0, 0, 1});
                    cfgs.push_back({order, 0, cell_i8_wide(P.l16), 1});
                    cfgs.push_back({order, 1, 0, 1});
                    cfgs.push_back({order, 2, 0, 0});
                } else if (P.l16) {
                    for (int var : kTuneVariants16)
                        for (int xcd = 0; xcd < 2; ++xcd) cfgs.push_back({order, 0, var, xcd});
                    for (int var : kTuneVariants16) {
                        cfgs.push_back({order, 1, var, 1});
                        cfgs.push_back({order, 2, var, 0});
                    }
                } else {
                    for (int var : kTuneVariants)
                        for (int xcd = 0; xcd < 2; ++xcd) cfgs.push_back({order, 0, var, xcd});
                    for (int var : {17, 13}) {
                        cfgs.push_back({order, 1, var, 1});
                        cfgs.push_back({order, 2, var, 0});
                    }
                }
            }
            // flat order (every pair, no cell padding): runs only
            if (P.i8) {
                cfgs.push_back({2, 0, 0, 1});
                cfgs.push_back({2, 0, cell_i8_wide(P.l16), 1});
            } else {
                for (int var : P.l16 ? std::vector<int>{19, 18} : std::vector<int>{17, 13})
                    cfgs.push_back({2, 0, var, 1});
            }
            v->sched_cfg.clear();
            v->sched_ms.clear();
            double best = 1e30;
            for (size_t k = 0; k < cfgs.size(); ++k) {
                const Cfg c = cfgs[k];
                RNNG_TRY(apply_layout(c.order));
                int64_t n_sub = 0;
                if (work_list(c.sched, c.var, c.xcd, 8, v->tune_items, &n_sub) < 0) return RNNG_ERR_HIP;
                double t = 0.0;
                for (int rep = 0; rep < 2; ++rep) {   // first run warms caches and code
                    RNNG_TRY(launch_fill_u64(counters.as<unsigned long long>(), 8, 0ull));
                    RNNG_HIP_TRY(hipEventRecord(ev.e[3], stream()));
                    RNNG_TRY(run_filter(c.sched, v->tune_items, (size_t)n_sub, v->ctl, nullptr, 0ull, c.var));
                    RNNG_HIP_TRY(hipEventRecord(ev.e[4], stream()));
                    RNNG_HIP_TRY(hipEventSynchronize(ev.e[4]));
                    t = ev.ms(3, 4);
                }
                v->sched_cfg.push_back(c.var << 16 | c.order << 8 | c.sched << 4 | (c.xcd & 1));
                v->sched_ms.push_back(t);
                if (!P.i8 && c.order == 0 && c.sched == 0 && k < 8) v->tune_ms[k] = t;
                if (t < best) {
                    best = t;
                    v->order = c.order;
                    v->sched = c.sched;
                    v->tuned_xcd = c.xcd;
                    if (P.i8) v->tuned_i8 = c.var;
                    else v->tuned_variant = c.var;
                }
            }
        }
        // tile lists run the ring shapes 17 and 13 (18 and 19 on the 16-row layout)
        if (!P.i8 && v->sched > 0) {
            const int a = P.l16 ? 18 : 17, b = P.l16 ? 19 : 13;
            if (v->tuned_variant != a && v->tuned_variant != b) v->tuned_variant = a;
        }
        RNNG_TRY(apply_layout(v->order));
    }
    const int variant = P.i8 ? v->tuned_i8 : v->tuned_variant;
    const int sched = v->sched;
    int64_t n_launch = 0;
    const int64_t covered = work_list(sched, variant, v->tuned_xcd, 1, d_items, &n_launch);
    if (covered < 0) return RNNG_ERR_HIP;
    const size_t n_items = (size_t)n_launch;

    LocalBuf& cand = v->cand;
    LocalBuf& cid = v->cid;
    LocalBuf& cflag = v->cflag;
    LocalBuf& cw = v->cw;
    // With the model: one pass into buffers sized by its estimate (+50 %), a
    // second only if that was short.  Without: count-then-fill, a counting
    // pass (capacity 0) and a filling pass at the exact size.
    // (never above the pairs the work list covers, which bound the candidates)
    unsigned long long cap =
        model ? (unsigned long long)std::min(pred_cand * 1.5 + 4096.0, (double)std::max<int64_t>(covered, 1)) : 0ull;
    unsigned long long ncand = 0;
    int passes = 0;
    bool graphed = false;
    // Buffers of a fill pass of capacity c (grown only, so their addresses stay
    // put across builds of one handle and a captured pass can be replayed).
    auto reserve_fill = [&](unsigned long long c) -> int {
        const size_t c1 = (size_t)std::max<unsigned long long>(1, c);
        RNNG_TRY(cand.reserve(c1 * sizeof(int2)));
        RNNG_TRY(cid.reserve(c1 * sizeof(int2)));
        RNNG_TRY(cflag.reserve(c1));
        RNNG_TRY(cw.reserve(c1 * sizeof(float)));
        RNNG_TRY(v->gkeys.reserve(2 * c1 * sizeof(uint64_t)));
        RNNG_TRY(v->counts.reserve((size_t)m * 2 * sizeof(unsigned)));
        return RNNG_OK;
    };
    // The fill pass: filter into capacity c, then one kernel for candidate ids,
    // FP32 recheck, reference semantics and per-cell counts, reading the
    // candidate count from the device (its grid is sized by c), so the pass
    // can be captured.
    const bool vec4 = d % 4 == 0 && ((uintptr_t)v->ds->X % 16) == 0 && ((uintptr_t)v->C.ptr % 16) == 0;
    auto fill_pass = [&](unsigned long long c) -> int {
        RNNG_TRY(launch_fill_u64(counters.as<unsigned long long>(), 8, 0ull));
        RNNG_TRY(run_filter(sched, d_items, n_items, v->ctl, cand.as<int2>(), c, variant));
        RNNG_TRY(launch_fill_u32(v->counts.as<unsigned>(), 2 * m, 0u));
        const unsigned rg =
            (unsigned)std::max<unsigned long long>(1, std::min<unsigned long long>(2048, (c + 63) / 64));
        const bool i8 = P.i8;
        const size_t lds = i8 ? 0 : (size_t)4 * 2 * RC_CELL * cell_rows_ld(d) * sizeof(float);
        auto kern = i8 ? k_cell_recheck<false, true> : (vec4 ? k_cell_recheck<true, false> : k_cell_recheck<false, false>);
        kern<<<rg, 256, lds, stream()>>>(v->ds->X, v->C.as<float>(), d, cand.as<int2>(), c, slot.as<int>(),
                                         v->cell.as<int>(), v->delta.as<float>(), T, eps, cid.as<int2>(),
                                         cflag.as<uint8_t>(), cw.as<float>(), counters.as<unsigned long long>(),
                                         v->gkeys.as<uint64_t>(), v->counts.as<unsigned>(),
                                         i8 ? P.xrow.as<i32x4_t>() : nullptr, i8 ? P.hsq.as<int>() : nullptr,
                                         i8 ? v->crow.as<i32x4_t>() : nullptr, i8 ? v->chsq.as<int>() : nullptr,
                                         i8 ? 2 * P.KS8 : 0);
        RNNG_HIP_TRY(hipGetLastError());
        return RNNG_OK;
    };
    auto read_count = [&]() -> int {
        RNNG_HIP_TRY(hipMemcpyAsync(&ncand, counters.ptr, sizeof(ncand), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        return RNNG_OK;
    };
    // buffers the estimate cannot get: the count-then-fill passes below
    if (model && cap > 0 && reserve_fill(cap) != RNNG_OK) {
        (void)hipGetLastError();
        cap = 0;
    }
    // the persistent grid's size comes from an occupancy query, made here
    // rather than inside a stream capture
    if (sched == 2) RNNG_TRY(run_filter(-2, d_items, n_items, v->ctl, nullptr, 0ull, variant));
    const double host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - host0).count();
    if (model && cap > 0) {
        // one pass sized by the model; captured once and replayed while the
        // buffers, the work list and eps are unchanged
        uint32_t eb, tb;
        std::memcpy(&eb, &eps, 4);
        std::memcpy(&tb, &T, 4);
        const std::vector<uint64_t> key = {
            (uint64_t)(uintptr_t)cand.ptr, (uint64_t)(uintptr_t)cid.ptr,  (uint64_t)(uintptr_t)cflag.ptr,
            (uint64_t)(uintptr_t)cw.ptr,   (uint64_t)(uintptr_t)v->gkeys.ptr, (uint64_t)(uintptr_t)v->counts.ptr,
            (uint64_t)(uintptr_t)d_items.ptr, (uint64_t)n_items, (uint64_t)(uintptr_t)P.xa.ptr,
            (uint64_t)(uintptr_t)P.yb.ptr, (uint64_t)(uintptr_t)P.xn.ptr, (uint64_t)(uintptr_t)P.ycn.ptr,
            (uint64_t)(uintptr_t)P.params.ptr, (uint64_t)(uintptr_t)slot.ptr,
            (uint64_t)(uintptr_t)counters.ptr, (uint64_t)(uintptr_t)v->ds->X, (uint64_t)(uintptr_t)v->C.ptr,
            (uint64_t)(uintptr_t)v->cell.ptr, (uint64_t)(uintptr_t)v->delta.ptr, (uint64_t)(uintptr_t)v->ctl.ptr,
            (uint64_t)(uintptr_t)v->queue.ptr, (uint64_t)sched, (uint64_t)v->order,
            cap, (uint64_t)variant,
            (uint64_t)n_pad, (uint64_t)eb, (uint64_t)tb, (uint64_t)m, (uint64_t)d,
            (uint64_t)(uintptr_t)P.xrow.ptr, (uint64_t)(uintptr_t)P.hsq.ptr, (uint64_t)(uintptr_t)v->crow.ptr,
            (uint64_t)(uintptr_t)v->chsq.ptr, (uint64_t)P.i8, (uint64_t)P.KS8};
        if (v->use_graph >= 0 ? v->use_graph : graph_enabled()) {
            if (!v->fill_exec || v->fill_key != key) {
                if (v->fill_exec) {
                    (void)hipGraphExecDestroy(v->fill_exec);
                    v->fill_exec = nullptr;
                }
                RNNG_HIP_TRY(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
                const int fs = fill_pass(cap);
                hipGraph_t graph = nullptr;
                const hipError_t ec = hipStreamEndCapture(stream(), &graph);
                if (fs != RNNG_OK) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return fs;
                }
                RNNG_HIP_TRY(ec);
                const hipError_t ei = hipGraphInstantiate(&v->fill_exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                RNNG_HIP_TRY(ei);
                v->fill_key = key;
            }
            RNNG_HIP_TRY(hipEventRecord(ev.e[3], stream()));
            RNNG_HIP_TRY(hipGraphLaunch(v->fill_exec, stream()));
            graphed = true;
        } else {
            RNNG_HIP_TRY(hipEventRecord(ev.e[3], stream()));
            RNNG_TRY(fill_pass(cap));
        }
        RNNG_HIP_TRY(hipEventRecord(ev.e[4], stream()));
        ++passes;
        RNNG_TRY(read_count());
        if (ncand > cap && st) st->refills++;
    } else {
        // count-then-fill: a counting pass (capacity 0, filter only)
        RNNG_TRY(launch_fill_u64(counters.as<unsigned long long>(), 8, 0ull));
        RNNG_HIP_TRY(hipEventRecord(ev.e[3], stream()));
        RNNG_TRY(run_filter(sched, d_items, n_items, v->ctl, nullptr, 0ull, variant));
        RNNG_HIP_TRY(hipEventRecord(ev.e[4], stream()));
        ++passes;
        RNNG_TRY(read_count());
        cap = 0;
    }
    if (ncand > cap || cap == 0) {
        // fill at the exact count (after the counting pass or a short estimate)
        cap = ncand;
        RNNG_TRY(reserve_fill(cap));
        RNNG_HIP_TRY(hipEventRecord(ev.e[3], stream()));
        RNNG_TRY(fill_pass(cap));
        RNNG_HIP_TRY(hipEventRecord(ev.e[4], stream()));
        ++passes;
        RNNG_TRY(read_count());
    }
    // the candidate count bounds the edge count (the model's edge estimate is
    // reported, not trusted with memory safety)
    if ((int64_t)ncand > g->cap) RNNG_TRY(edge_buffers(g, (int64_t)ncand));
    int64_t ne = 0;
    RNNG_TRY(compact_edges(cid.as<int2>(), cw.as<float>(), cflag.as<uint8_t>(), (int64_t)ncand, g->ij.as<int2>(),
                           g->w.as<float>(), &ne));
    g->n_edges = ne;
    unsigned long long cnt[8];
    RNNG_HIP_TRY(hipMemcpyAsync(cnt, counters.ptr, sizeof(cnt), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    // ---- ghosts: distinct (cell, point) pairs ----
    const int64_t nhits = (int64_t)cnt[5];
    int64_t ng = 0;
    if (nhits > 0) {
        // keys (cell << 32 | point), cell < m: the bits above 32 + bits_for(m) are zero
        RNNG_TRY(sort_keys_u64(v->gkeys.as<uint64_t>(), nhits, 32 + bits_for(m)));
        LocalBuf uflag, uniq, nsel;
        RNNG_TRY(uflag.reserve((size_t)nhits));
        RNNG_TRY(uniq.reserve((size_t)nhits * sizeof(uint64_t)));
        RNNG_TRY(nsel.reserve(sizeof(unsigned long long)));
        k_unique_flags<<<blocks_for(nhits, 256), 256, 0, stream()>>>(v->gkeys.as<uint64_t>(), nhits,
                                                                      uflag.as<uint8_t>());
        RNNG_HIP_TRY(hipGetLastError());
        size_t tb = 0;
        RNNG_HIP_TRY(rocprim::select(nullptr, tb, v->gkeys.as<uint64_t>(), uflag.as<uint8_t>(), uniq.as<uint64_t>(),
                                     nsel.as<unsigned long long>(), (size_t)nhits, stream()));
        LocalBuf tmp;
        RNNG_TRY(tmp.reserve(tb));
        RNNG_HIP_TRY(rocprim::select(tmp.ptr, tb, v->gkeys.as<uint64_t>(), uflag.as<uint8_t>(), uniq.as<uint64_t>(),
                                     nsel.as<unsigned long long>(), (size_t)nhits, stream()));
        unsigned long long h = 0;
        RNNG_HIP_TRY(hipMemcpyAsync(&h, nsel.ptr, sizeof(h), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        ng = (int64_t)h;
        RNNG_HIP_TRY(hipMemcpyAsync(v->gkeys.ptr, uniq.ptr, (size_t)ng * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                                    stream()));
    }
    v->n_ghosts = ng;
    RNNG_HIP_TRY(hipEventRecord(ev.e[5], stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    if (st) {
        st->pair_evals += covered;
        st->candidates += (int64_t)ncand;
        st->int8_filter += P.i8 ? 1 : 0;
        st->edges += ne;
        st->band_pairs += (int64_t)cnt[2];
        st->band_flips += (int64_t)cnt[3];
        st->ghost_tests += (int64_t)cnt[4];
        st->ghost_candidates += nhits;
        st->ghosts += ng;
        st->ghost_ms += ev.ms(0, 1);
        st->filter_passes += passes;
        st->predicted_candidates += (int64_t)pred_cand;
        st->predicted_edges += (int64_t)pred_edges;
        st->prep_ms += ev.ms(1, 2);
        st->filter_ms += ev.ms(3, 4);   // the fill pass (filter through per-cell counts)
        st->recheck_ms += ev.ms(4, 5);  // compaction and ghost list
        st->total_ms += ev.ms(0, 2) + ev.ms(3, 5);
        st->host_ms += host_ms;
        (void)graphed;
    }
    return RNNG_OK;
}

}  // namespace
}  // namespace rnng

using namespace rnng;

extern "C" {

int rnng_selection_sample(int64_t range, int64_t size, int64_t* sample, int32_t seed) {
    if (range < 0 || size < 0 || size > range || (size > 0 && !sample)) {
        set_error("selection_sample: need 0 <= size <= range and an output buffer");
        return RNNG_ERR_ARG;
    }
    if (seed < 0) {
        std::random_device rd;
        seed = (int32_t)rd();
    }
    // Knuth's Algorithm S with the same engine and distribution as the
    // reference (libstdc++ minstd_rand0 and generate_canonical<double>).
    std::default_random_engine gen(seed);
    std::uniform_real_distribution<double> U(0, std::nextafter(1.0, std::numeric_limits<double>::max()));
    int64_t pos = 0;
    for (int64_t i = 0; i < range && pos < size; ++i)
        if ((double)(range - i) * U(gen) < (double)(size - pos)) sample[pos++] = i;
    return RNNG_OK;
}

int rnng_voronoi_assign_f32(const float* points, int64_t n, const float* centers, int64_t m, int32_t d,
                            int64_t* cell, float* dist) {
    if (!points || !centers || !cell || !dist || n <= 0 || m <= 0 || d <= 0 || n >= INT32_MAX ||
        m >= INT32_MAX) {
        set_error("voronoi_assign: bad arguments");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    LocalBuf X, C, c32, dd;
    RNNG_TRY(X.reserve((size_t)n * d * sizeof(float)));
    RNNG_TRY(C.reserve((size_t)m * d * sizeof(float)));
    RNNG_TRY(c32.reserve((size_t)n * sizeof(int)));
    RNNG_TRY(dd.reserve((size_t)n * sizeof(float)));
    RNNG_HIP_TRY(hipMemcpyAsync(X.ptr, points, (size_t)n * d * sizeof(float), hipMemcpyHostToDevice, stream()));
    RNNG_HIP_TRY(hipMemcpyAsync(C.ptr, centers, (size_t)m * d * sizeof(float), hipMemcpyHostToDevice, stream()));
    k_assign<<<blocks_for(n, AB), 256, 0, stream()>>>(X.as<float>(), (int)n, C.as<float>(), (int)m, d, c32.as<int>(),
                                                      dd.as<float>());
    RNNG_HIP_TRY(hipGetLastError());
    std::vector<int> h((size_t)n);
    RNNG_HIP_TRY(hipMemcpyAsync(h.data(), c32.ptr, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipMemcpyAsync(dist, dd.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    for (int64_t i = 0; i < n; ++i) cell[i] = h[(size_t)i];
    return RNNG_OK;
}

static int voronoi_new(rnng_dataset_t ds, int64_t m, rnng_voronoi_s** out) {
    if (!ds || !out || m <= 0 || m > ds->n || ds->n >= INT32_MAX) {
        set_error("voronoi: need a dataset and 0 < m <= n < 2^31");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    auto* v = new rnng_voronoi_s();
    v->ds = ds;
    v->m = m;
    *out = v;
    return RNNG_OK;
}

int rnng_voronoi_create(rnng_dataset_t ds, const int64_t* centers, int64_t m, rnng_voronoi_t* out) {
    if (!centers) {
        set_error("voronoi_create: null centers");
        return RNNG_ERR_ARG;
    }
    rnng_voronoi_s* v = nullptr;
    RNNG_TRY(voronoi_new(ds, m, &v));
    v->centers.resize(m);
    for (int64_t c = 0; c < m; ++c) {
        if (centers[c] < 0 || centers[c] >= ds->n) {
            delete v;
            set_error("voronoi_create: center id out of range");
            return RNNG_ERR_ARG;
        }
        v->centers[c] = (int)centers[c];
    }
    int s = upload(v->ctr, v->centers);
    if (s == RNNG_OK) s = gather_centers(v);
    if (s == RNNG_OK) s = assign(v);
    if (s == RNNG_OK) s = build_cells(v);
    if (s != RNNG_OK) {
        delete v;
        return s;
    }
    *out = v;
    return RNNG_OK;
}

int rnng_voronoi_create_fps(rnng_dataset_t ds, int64_t m, int64_t first, rnng_voronoi_t* out) {
    rnng_voronoi_s* v = nullptr;
    RNNG_TRY(voronoi_new(ds, m, &v));
    if (first < 0 || first >= ds->n) {
        delete v;
        set_error("voronoi_create_fps: first point out of range");
        return RNNG_ERR_ARG;
    }
    int s = v->ctr.reserve((size_t)m * sizeof(int));
    if (s == RNNG_OK) s = fps(v, (int)first);
    if (s == RNNG_OK) s = gather_centers(v);
    if (s == RNNG_OK) s = build_cells(v);
    if (s != RNNG_OK) {
        delete v;
        return s;
    }
    *out = v;
    return RNNG_OK;
}

int rnng_voronoi_centers(rnng_voronoi_t v, int64_t* ids) {
    RNNG_TRY(check_voronoi(v));
    if (!ids) {
        set_error("null output");
        return RNNG_ERR_ARG;
    }
    for (int64_t c = 0; c < v->m; ++c) ids[c] = v->centers[(size_t)c];
    return RNNG_OK;
}

int rnng_voronoi_assignment(rnng_voronoi_t v, int64_t* cell, float* dist) {
    RNNG_TRY(check_voronoi(v));
    const int64_t n = v->ds->n;
    if (cell) {
        std::vector<int> h;
        RNNG_TRY(download(h, v->cell, (size_t)n));
        for (int64_t i = 0; i < n; ++i) cell[i] = h[(size_t)i];
    }
    if (dist) {
        RNNG_HIP_TRY(hipMemcpyAsync(dist, v->delta.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    }
    return RNNG_OK;
}

int rnng_voronoi_cells(rnng_voronoi_t v, int64_t* cell_ptr, int64_t* cell_ids) {
    RNNG_TRY(check_voronoi(v));
    if (cell_ptr) {
        const std::vector<int> p = prefix(v->cnt);
        for (size_t c = 0; c < p.size(); ++c) cell_ptr[c] = p[c];
    }
    if (cell_ids) {
        std::vector<int> h;
        RNNG_TRY(download(h, v->cids, (size_t)v->ds->n));
        for (size_t i = 0; i < h.size(); ++i) cell_ids[i] = h[i];
    }
    return RNNG_OK;
}

int rnng_voronoi_graph_build(rnng_voronoi_t v, float eps, int32_t mode, rnng_graph_t* out, rnng_stats* st) {
    RNNG_TRY(check_voronoi(v));
    if (!out || !(eps >= 0.0f) || !std::isfinite(eps)) {
        set_error("voronoi_graph_build: bad arguments");
        return RNNG_ERR_ARG;
    }
    if (mode != RNNG_MODE_MFMA) {
        set_error("voronoi_graph_build: the batched cell query runs in MFMA mode");
        return RNNG_ERR_UNSUPPORTED;
    }
    const auto t0 = std::chrono::steady_clock::now();
    auto* g = new rnng_graph_s();
    const int s = voronoi_graph(v, eps, g, st);
    if (st) st->call_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (s != RNNG_OK) {
        g->ij.release();
        g->w.release();
        delete g;
        return s;
    }
    *out = g;
    return RNNG_OK;
}

int rnng_voronoi_cell_stats(rnng_voronoi_t v, float* features, float* counts) {
    RNNG_TRY(check_voronoi(v));
    const size_t m = (size_t)v->m;
    if (!v->feat.ptr || !v->counts.ptr) {
        set_error("voronoi_cell_stats: no graph build yet");
        return RNNG_ERR_ARG;
    }
    if (features) {
        RNNG_HIP_TRY(hipMemcpyAsync(features, v->feat.ptr, m * 4 * sizeof(float), hipMemcpyDeviceToHost, stream()));
    }
    std::vector<unsigned> c;
    RNNG_TRY(download(c, v->counts, m * 2));
    if (counts)
        for (size_t i = 0; i < 2 * m; ++i) counts[i] = (float)(std::log2(1.0 + (double)c[i]) / 24.0);
    return RNNG_OK;
}

int rnng_voronoi_set_options(rnng_voronoi_t v, int32_t use_graph, int32_t autotune) {
    RNNG_TRY(check_voronoi(v));
    v->use_graph = use_graph < 0 ? -1 : (use_graph ? 1 : 0);
    v->autotune = autotune < 0 ? -1 : (autotune ? 1 : 0);
    if (v->use_graph == 0 && v->fill_exec) {
        (void)hipGraphExecDestroy(v->fill_exec);
        v->fill_exec = nullptr;
        v->fill_key.clear();
    }
    if (v->autotune >= 0) {   // choose again at the next build
        v->tuned_variant = -1;
        v->tuned_i8 = 0;
    }
    return RNNG_OK;
}

namespace {

// The work lists of one layout by the host builder (device 0) or the device
// builder (device 1, stride 1, no XCD order).
int debug_cell_work(const int32_t* cnt, int32_t m, const uint32_t* adj, const int32_t* perm, int32_t bt,
                    int32_t sched, const float* rb_cand, int32_t device, int32_t* start, int32_t* items,
                    int64_t cap_items, int32_t* ctl, int64_t cap_ctl, int64_t* n_items, int64_t* n_ctl,
                    int64_t* covered) {
    if (!cnt || !adj || !perm || m <= 0 || bt <= 0 || sched < 0 || sched > 2 || !n_items || !n_ctl || !covered) {
        set_error("rnng_debug_cell_work: bad arguments");
        return RNNG_ERR_ARG;
    }
    const int words = (m + 31) / 32;
    std::vector<int> cn(cnt, cnt + m), pm(perm, perm + m);
    std::vector<unsigned> ad(adj, adj + (size_t)m * words);
    CellLayout L;
    make_layout(pm, cn, &ad, m, words, L);
    const int nrb = (L.rs[m] + bt * 32 - 1) / (bt * 32);
    std::vector<float> rbc;
    if (sched == 2) rbc = rb_cand ? std::vector<float>(rb_cand, rb_cand + nrb) : std::vector<float>((size_t)nrb, 0.0f);
    std::vector<int4> it;
    std::vector<int> cl;
    int64_t cov = 0;
    if (device) {
        if (L.rs[m] == 0 || words > kMaxCellWords) {
            set_error("rnng_debug_cell_work: no tiles, or too many cells");
            return RNNG_ERR_ARG;
        }
        RNNG_TRY(ensure_init());
        rnng_voronoi_s tmp;
        RNNG_TRY(upload(tmp.adj_dev, ad));
        RNNG_TRY(upload(tmp.lay_perm, L.perm));
        RNNG_TRY(upload(tmp.lay_pos, L.pos));
        RNNG_TRY(upload(tmp.lay_rs, L.rs));
        RNNG_TRY(upload(tmp.lay_tcell, L.tcell));
        LocalBuf out;
        int64_t no = 0;
        RNNG_TRY(device_work_list(&tmp, tmp.adj_dev.as<unsigned>(), m, words, L.rs[m] / 32, bt, sched, 0, 1, rbc,
                                  out, &no, &cov));
        RNNG_TRY(download(it, out, (size_t)no));
        if (sched > 0) {
            int64_t nt = 0;
            for (const int4& w : it) nt = std::max<int64_t>(nt, w.z);
            RNNG_TRY(download(cl, tmp.ctl, (size_t)nt));
        }
        const int rc = hipStreamSynchronize(stream()) == hipSuccess ? RNNG_OK : RNNG_ERR_HIP;
        tmp.adj_dev.release();
        RNNG_TRY(rc);
    } else {
        if (sched == 0) cov = cell_work(L.rs, L.adj, words, m, bt, nrb, it);
        else cov = cell_work_tiles(L.rs, L.adj, words, m, bt, nrb, kMaxListItem, it, cl);
        if (sched == 2) order_by_cost(it, rbc);
    }
    *n_items = (int64_t)it.size();
    *n_ctl = (int64_t)cl.size();
    *covered = cov;
    if (start)
        for (int c = 0; c < m; ++c) start[c] = L.start[(size_t)c];
    if ((int64_t)it.size() > cap_items || (int64_t)cl.size() > cap_ctl) {
        set_error("rnng_debug_cell_work: output capacity too small");
        return RNNG_ERR_CAPACITY;
    }
    for (size_t i = 0; i < it.size(); ++i) {
        items[4 * i] = it[i].x;
        items[4 * i + 1] = it[i].y;
        items[4 * i + 2] = it[i].z;
        items[4 * i + 3] = it[i].w;
    }
    for (size_t i = 0; i < cl.size(); ++i) ctl[i] = cl[i];
    return RNNG_OK;
}

}  // namespace

int rnng_debug_cell_work(const int32_t* cnt, int32_t m, const uint32_t* adj, const int32_t* perm, int32_t bt,
                         int32_t sched, const float* rb_cand, int32_t* start, int32_t* items, int64_t cap_items,
                         int32_t* ctl, int64_t cap_ctl, int64_t* n_items, int64_t* n_ctl, int64_t* covered) {
    return debug_cell_work(cnt, m, adj, perm, bt, sched, rb_cand, 0, start, items, cap_items, ctl, cap_ctl, n_items,
                           n_ctl, covered);
}

int rnng_debug_cell_work_device(const int32_t* cnt, int32_t m, const uint32_t* adj, const int32_t* perm, int32_t bt,
                                int32_t sched, const float* rb_cand, int32_t* start, int32_t* items,
                                int64_t cap_items, int32_t* ctl, int64_t cap_ctl, int64_t* n_items, int64_t* n_ctl,
                                int64_t* covered) {
    return debug_cell_work(cnt, m, adj, perm, bt, sched, rb_cand, 1, start, items, cap_items, ctl, cap_ctl, n_items,
                           n_ctl, covered);
}

int rnng_voronoi_set_schedule(rnng_voronoi_t v, int32_t order, int32_t sched) {
    RNNG_TRY(check_voronoi(v));
    if (order > 2 || sched > 2 || (order == 2 && sched > 0)) {
        set_error("rnng_voronoi_set_schedule: order 0..2, sched 0..2, the flat order 2 with runs (sched 0) only "
                  "(-1: the autotuner's)");
        return RNNG_ERR_ARG;
    }
    v->forced_order = order < 0 ? -1 : order;
    v->forced_sched = sched < 0 ? -1 : sched;
    v->tuned_variant = -1;   // lay out and choose again at the next build
    v->tuned_i8 = 0;
    v->order = -1;
    v->sched = -1;
    return RNNG_OK;
}

int rnng_voronoi_schedule(rnng_voronoi_t v, int32_t* order, int32_t* sched, int32_t* configs, double* config_ms,
                          int32_t cap, int32_t* n_configs) {
    RNNG_TRY(check_voronoi(v));
    if (order) *order = v->order;
    if (sched) *sched = v->sched;
    const int nc = (int)v->sched_cfg.size();
    if (n_configs) *n_configs = nc;
    for (int k = 0; k < std::min(nc, (int)cap); ++k) {
        if (configs) configs[k] = v->sched_cfg[(size_t)k];
        if (config_ms) config_ms[k] = v->sched_ms[(size_t)k];
    }
    return RNNG_OK;
}

int rnng_voronoi_tuning(rnng_voronoi_t v, int32_t* variant, int32_t* xcd_order, double* config_ms) {
    RNNG_TRY(check_voronoi(v));
    if (variant) *variant = v->tuned_variant;
    if (xcd_order) *xcd_order = v->tuned_xcd;
    if (config_ms)
        for (int k = 0; k < 8; ++k) config_ms[k] = v->tune_ms[k];
    return RNNG_OK;
}

int rnng_voronoi_num_ghosts(rnng_voronoi_t v, int64_t* n) {
    RNNG_TRY(check_voronoi(v));
    if (!n) {
        set_error("null output");
        return RNNG_ERR_ARG;
    }
    *n = v->n_ghosts;
    return RNNG_OK;
}

int rnng_voronoi_ghosts(rnng_voronoi_t v, int64_t* pairs, int64_t cap) {
    RNNG_TRY(check_voronoi(v));
    if (cap < v->n_ghosts || (v->n_ghosts > 0 && !pairs)) {
        set_error("voronoi_ghosts: output holds fewer pairs than the ghost count");
        return RNNG_ERR_CAPACITY;
    }
    std::vector<uint64_t> k;
    RNNG_TRY(download(k, v->gkeys, (size_t)v->n_ghosts));
    for (size_t e = 0; e < k.size(); ++e) {
        pairs[2 * e] = (int64_t)(k[e] >> 32);
        pairs[2 * e + 1] = (int64_t)(k[e] & 0xffffffffull);
    }
    return RNNG_OK;
}

int rnng_voronoi_destroy(rnng_voronoi_t v) {
    if (!v) return RNNG_OK;
    for (DevBuf* b : {&v->ctr, &v->C, &v->cell, &v->delta, &v->cptr, &v->cids, &v->gkeys, &v->feat, &v->counts})
        b->release();
    delete v;
    return RNNG_OK;
}

}  // extern "C"
