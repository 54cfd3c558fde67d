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
//   cell query  : member x member and member x ghost pairs within eps of every
//                 cell (drivers/landmark.cpp:203-210), then the symmetric,
//                 deduplicated edge set (include/graph.hpp:59-116)
//
// The ghost center test is pruned per cell with the triangle inequality:
// L2(c_b, q) >= L2(c_b, c_a) - delta_q for q in cell a, so a center b can pass
// only if L2(c_a, c_b) <= 2 R_a + 2 eps (R_a = max delta in cell a).  The bound
// carries a 1e-3 relative margin, far above the FP32 distance error, so the
// pruning never removes a center that the exact test would accept.  The member
// existence test of a candidate ghost is a by-product of the cell query: a
// candidate ghost is a ghost exactly when it has an edge to a member.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "mfma.hpp"

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
    // candidate ghosts of the last graph build, sorted by (cell, point)
    rnng::DevBuf gkeys;                   // u64 (cell << 32 | point)
    rnng::DevBuf greal;                   // uint8 per candidate ghost: is a ghost
    int64_t n_gcand = 0;
    int64_t n_ghosts = 0;
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
constexpr int FPS_K = 16;   // dims per LDS round

__global__ __launch_bounds__(256) void k_fps_update(const float* __restrict__ X, int n, int d,
                                                    int* __restrict__ ctr, int k, float* __restrict__ mind,
                                                    int* __restrict__ cell,
                                                    unsigned long long* __restrict__ part) {
#pragma clang fp contract(off)
    extern __shared__ float fps_lds[];   // [d] center, then [4 waves][FPS_K][65]
    float* cs = fps_lds;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    float* tile = fps_lds + ((d + 3) & ~3) + wave * FPS_K * 65;
    const int c = ctr[k];
    for (int t = tid; t < d; t += 256) cs[t] = X[(size_t)c * d + t];
    __syncthreads();
    const int p0 = blockIdx.x * 256 + wave * 64;
    const int i = p0 + lane;
    float acc = 0.0f;
    for (int k0 = 0; k0 < d; k0 += FPS_K) {
        // 64 rows x FPS_K dims, 16 consecutive lanes per row segment
#pragma unroll
        for (int j = 0; j < FPS_K; ++j) {
            const int e = lane + 64 * j;
            const int r = e / FPS_K, kk = e % FPS_K;
            const int row = p0 + r, kd = k0 + kk;
            tile[kk * 65 + r] = (row < n && kd < d) ? X[(size_t)row * d + kd] : 0.0f;
        }
        // rows are written and read by lanes of this wave only
        __builtin_amdgcn_wave_barrier();
        const int kend = min(FPS_K, d - k0);
        for (int kk = 0; kk < kend; ++kk) {
            const float t = tile[kk * 65 + lane] - cs[k0 + kk];
            acc = acc + t * t;
        }
        __builtin_amdgcn_wave_barrier();
    }
    unsigned long long key = 0ull;
    if (i < n) {
        const float dd = sqrtf(acc);
        if (k == 0) cell[i] = 0;
        float md = k == 0 ? FLT_MAX : mind[i];
        if (dd <= md) {
            md = dd;
            cell[i] = k;
        }
        mind[i] = md;
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

// Centers b that can hold ghosts of cell a: b != a, b non-empty and
// L2(c_a, c_b) <= (2 R_a + 2 eps)(1 + 1e-3).  One workgroup per cell a;
// FILL = 0 counts, FILL = 1 writes the list at lptr[a].
template <int FILL>
__global__ __launch_bounds__(256) void k_center_lists(const float* __restrict__ C, int m, int d,
                                                      const int* __restrict__ cnt,
                                                      const unsigned* __restrict__ rad, float eps,
                                                      int* __restrict__ lcnt, const int* __restrict__ lptr,
                                                      int* __restrict__ list) {
    extern __shared__ float ca[];
    const int a = blockIdx.x;
    if (cnt[a] == 0) {
        if (!FILL && threadIdx.x == 0) lcnt[a] = 0;
        return;
    }
    for (int t = threadIdx.x; t < d; t += 256) ca[t] = C[(size_t)a * d + t];
    __syncthreads();
    const double bound = (2.0 * (double)ord2f_v(rad[a]) + 2.0 * (double)eps) * (1.0 + 1e-3) + 1e-30;
    __shared__ int wcount;
    if (threadIdx.x == 0) wcount = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int base = FILL ? lptr[a] : 0;
    for (int b0 = 0; b0 < m; b0 += 256) {
        const int b = b0 + threadIdx.x;
        bool keep = false;
        if (b < m && b != a && cnt[b] > 0) {
            const float* cb = C + (size_t)b * d;
            float acc = 0.0f;
            for (int k = 0; k < d; ++k) {
                const float t = ca[k] - cb[k];
                acc = acc + t * t;
            }
            keep = (double)sqrtf(acc) <= bound;
        }
        const unsigned long long bal = __ballot(keep);
        int off = 0;
        if (lane == 0 && bal) off = atomicAdd(&wcount, __popcll(bal));
        off = __shfl(off, 0, 64);
        if (FILL && keep)
            list[base + off + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u))] = b;
    }
    __syncthreads();
    if (!FILL && threadIdx.x == 0) lcnt[a] = wcount;
}

// Ghost center test.  Work item (cell a, 128-member tile, 128-center tile of
// a's list): the exact test L2(c_b, q) <= fl(delta_q + 2 eps) for every pair;
// passing pairs are written as keys (b << 32 | q).
__global__ __launch_bounds__(256) void k_ghost_test(const float* __restrict__ X, const float* __restrict__ C,
                                                    int d, const int4* __restrict__ items,
                                                    const int* __restrict__ cptr, const int* __restrict__ cids,
                                                    const int* __restrict__ lptr, const int* __restrict__ lcnt,
                                                    const int* __restrict__ list, const float* __restrict__ delta,
                                                    float eps, uint64_t* __restrict__ out, unsigned long long cap,
                                                    unsigned long long* __restrict__ counter) {
    __shared__ float As[AK][AB + 4];
    __shared__ float Bs[AK][AB + 4];
    const int4 wi = items[blockIdx.x];
    const int a = wi.x;
    const int m0 = cptr[a] + wi.y * AB, m1 = cptr[a + 1];
    const int l0 = lptr[a] + wi.z * AB, l1 = lptr[a] + lcnt[a];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    float acc[8][8];
    fp32_tile(
        d, [&](int i) -> const float* { return m0 + i < m1 ? X + (size_t)cids[m0 + i] * d : nullptr; },
        [&](int j) -> const float* { return l0 + j < l1 ? C + (size_t)list[l0 + j] * d : nullptr; }, As, Bs,
        acc);
    const float two_eps = 2.0f * eps;
    unsigned cnt = 0;
    int qid[8];
    float reach[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int mi = m0 + micro(ty, r);
        qid[r] = mi < m1 ? cids[mi] : -1;
        reach[r] = qid[r] >= 0 ? delta[qid[r]] + two_eps : -1.0f;
    }
    int bid[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int lj = l0 + micro(tx, c);
        bid[c] = lj < l1 ? list[lj] : -1;
    }
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) cnt += (qid[r] >= 0 && bid[c] >= 0 && sqrtf(acc[r][c]) <= reach[r]) ? 1u : 0u;
    if (!__any(cnt != 0)) return;
    unsigned long long pos = wave_reserve(cnt, counter);
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (qid[r] >= 0 && bid[c] >= 0 && sqrtf(acc[r][c]) <= reach[r]) {
                if (pos < cap) out[pos] = ((uint64_t)(uint32_t)bid[c] << 32) | (uint32_t)qid[r];
                ++pos;
            }
}

// Per-cell slot maps of the batched query: rows = members padded to 32,
// columns = members then candidate ghosts, padded to 32 (-1 = padding).
__global__ __launch_bounds__(256) void k_slot_maps(const int* __restrict__ cptr, const int* __restrict__ cids,
                                                   const int* __restrict__ gptr, const uint64_t* __restrict__ gkeys,
                                                   const int* __restrict__ rs, const int* __restrict__ cs,
                                                   int* __restrict__ rowmap, int* __restrict__ colmap) {
    const int b = blockIdx.x;
    const int nb = cptr[b + 1] - cptr[b], gb = gptr[b + 1] - gptr[b];
    const int nr = rs[b + 1] - rs[b], nc = cs[b + 1] - cs[b];
    for (int k = threadIdx.x; k < nr; k += 256) rowmap[rs[b] + k] = k < nb ? cids[cptr[b] + k] : -1;
    for (int k = threadIdx.x; k < nc; k += 256) {
        int v = -1;
        if (k < nb) v = cids[cptr[b] + k];
        else if (k < nb + gb) v = (int)(gkeys[gptr[b] + k - nb] & 0xffffffffull);
        colmap[cs[b] + k] = v;
    }
}

// ------------------------------------------------------------------------
// Batched cell query (BF16 MFMA filter)
// ------------------------------------------------------------------------
// Persistent: each wave takes work items gw, gw + W, ... of a list sorted by
// decreasing cost.  Item = (first row tile, first column tile, end column
// tile, cell); the wave keeps RT row tiles' fragments in registers and
// streams the column tiles from L2 with a one-tile register prefetch.  A
// candidate (row slot r, column slot c) of cell b needs row r a member
// (r - rs_b < n_b) and column c either a later member (r - rs_b < c - cs_b <
// n_b) or a candidate ghost (n_b <= c - cs_b < n_b + g_b).  Candidates are
// staged per wave in LDS and written with one reservation per flush.
constexpr int WST = 512;   // staged candidates per wave

struct CellInfo {
    int rs, cs, nb, dom;   // first row slot, first column slot, members, members + candidate ghosts
};

template <int KS, int RT>
__global__ __launch_bounds__(256) void k_cell_filter(const bf16x8_t* __restrict__ xa,
                                                     const bf16x8_t* __restrict__ yb,
                                                     const int4* __restrict__ items, int n_items,
                                                     const CellInfo* __restrict__ info, int2* __restrict__ cand,
                                                     unsigned long long cap,
                                                     unsigned long long* __restrict__ counter) {
    __shared__ int2 stage[4][WST];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = gridDim.x * 4;
    int2* st = stage[wave];
    unsigned scnt = 0;   // wave-uniform
    auto flush = [&]() {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(counter, (unsigned long long)scnt);
        base = __shfl(base, 0, 64);
        for (unsigned i = lane; i < scnt; i += 64)
            if (base + i < cap) cand[base + i] = st[i];
        scnt = 0;
    };
    for (int it = blockIdx.x * 4 + wave; it < n_items; it += nw) {
        const int4 wi = items[it];
        const CellInfo ci = info[wi.w];
        bf16x8_t a[RT][KS];
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int s = 0; s < KS; ++s) a[r][s] = xa[((size_t)(wi.x + r) * KS + s) * 64 + lane];
        bf16x8_t b[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) b[s] = yb[((size_t)wi.y * KS + s) * 64 + lane];
        for (int ct = wi.y; ct < wi.z; ++ct) {
            const int ctn = ct + 1 < wi.z ? ct + 1 : ct;
            bf16x8_t pre[KS];
#pragma unroll
            for (int s = 0; s < KS; ++s) pre[s] = yb[((size_t)ctn * KS + s) * 64 + lane];
            f32x16_t acc[RT];
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int q = 0; q < 16; ++q) acc[r][q] = 0.0f;
#pragma unroll
            for (int s = 0; s < KS; ++s)
#pragma unroll
                for (int r = 0; r < RT; ++r)
                    acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[r][s], b[s], acc[r], 0, 0, 0);
            unsigned o = 0;
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int q = 0; q < 16; ++q) o |= __float_as_uint(acc[r][q]);
            if (__any((int)o < 0)) {
                unsigned long long mask = 0;
#pragma unroll
                for (int r = 0; r < RT; ++r)
#pragma unroll
                    for (int q = 0; q < 16; ++q)
                        mask |= (unsigned long long)(__float_as_uint(acc[r][q]) >> 31) << (r * 16 + q);
                const int col = ct * 32 + (lane & 31);
                const int cl = col - ci.cs;
                unsigned long long any = mask;
#pragma unroll
                for (int o2 = 1; o2 < 64; o2 <<= 1) any |= __shfl_xor(any, o2, 64);
                {
                    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)any);
                    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(any >> 32));
                    any = ((unsigned long long)hi << 32) | lo;
                }
#pragma unroll 1
                while (any) {
                    const int rq = __builtin_ctzll(any);
                    any &= any - 1;
                    const int q = rq & 15;
                    const int row = (wi.x + (rq >> 4)) * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                    const int rl = row - ci.rs;
                    const bool ok = ((mask >> rq) & 1ull) && rl < ci.nb &&
                                    ((cl > rl && cl < ci.nb) || (cl >= ci.nb && cl < ci.dom));
                    const unsigned long long bal = __ballot(ok);
                    if (bal) {
                        const unsigned pos =
                            scnt + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
                        if (ok) st[pos] = make_int2(row, col);
                        scnt += (unsigned)__popcll(bal);
                        if (scnt > WST - 64) flush();
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < KS; ++s) b[s] = pre[s];
        }
    }
    if (scnt) flush();
}

// Candidate slots -> point ids (rowmap, colmap).
__global__ void k_cand_ids(const int2* __restrict__ cand, const unsigned long long* __restrict__ counter,
                           unsigned long long cap, const int* __restrict__ rowmap,
                           const int* __restrict__ colmap, int2* __restrict__ ids) {
    const unsigned long long nc = min(counter[0], cap);
    for (unsigned long long c = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (unsigned long long)gridDim.x * blockDim.x) {
        const int2 s = cand[c];
        ids[c] = make_int2(rowmap[s.x], colmap[s.y]);
    }
}

// After the recheck: marks candidate ghosts with an edge to a member, counts
// candidates and edges per cell, and writes each edge as (min, max) key.
__global__ void k_cell_edges(const int2* __restrict__ cand, const int2* __restrict__ ids,
                             const uint8_t* __restrict__ flag, const unsigned long long* __restrict__ counter,
                             unsigned long long cap, const int* __restrict__ colcell,
                             const CellInfo* __restrict__ info, uint8_t* __restrict__ colreal,
                             int* __restrict__ ccand, int* __restrict__ cedge) {
    const unsigned long long nc = min(counter[0], cap);
    for (unsigned long long c = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (unsigned long long)gridDim.x * blockDim.x) {
        const int col = cand[c].y;
        const int b = colcell[col >> 5];
        atomicAdd(&ccand[b], 1);
        if (flag[c]) {
            atomicAdd(&cedge[b], 1);
            if (col - info[b].cs >= info[b].nb) colreal[col] = 1;
        }
    }
}

__global__ void k_edge_keys(const int2* __restrict__ ij, int64_t n, uint64_t* __restrict__ keys) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) {
        const int i = ij[e].x, j = ij[e].y;
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        keys[e] = ((uint64_t)(uint32_t)lo << 32) | (uint32_t)hi;
    }
}

// Keeps the first of each run of equal keys (keys sorted).
__global__ void k_unique_flags(const uint64_t* __restrict__ keys, int64_t n, uint8_t* __restrict__ flag) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) flag[e] = (e == 0 || keys[e] != keys[e - 1]) ? 1 : 0;
}

__global__ void k_keys_to_ij(const uint64_t* __restrict__ keys, int64_t n, int2* __restrict__ ij) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) ij[e] = make_int2((int)(keys[e] >> 32), (int)(keys[e] & 0xffffffffull));
}

// Real-ghost flag per candidate ghost k of cell b: column slot cs_b + n_b + k.
__global__ void k_ghost_flags(const uint64_t* __restrict__ gkeys, int64_t ng, const int* __restrict__ gptr,
                              const CellInfo* __restrict__ info, const uint8_t* __restrict__ colreal,
                              uint8_t* __restrict__ greal) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < ng) {
        const int b = (int)(gkeys[e] >> 32);
        greal[e] = colreal[info[b].cs + info[b].nb + (int)(e - gptr[b])];
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
    const size_t lds = (size_t)(((d + 3) & ~3) + 4 * FPS_K * 65) * sizeof(float);
    for (int k = 0; k < m; ++k) {
        k_fps_update<<<nb, 256, lds, stream()>>>(v->ds->X, n, d, v->ctr.as<int>(), k, v->delta.as<float>(),
                                                 v->cell.as<int>(), part.as<unsigned long long>());
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

// Ghost candidates: (cell, point) keys sorted, in v->gkeys; per-cell counts in gcnt.
int ghost_candidates(rnng_voronoi_s* v, float eps, std::vector<int>& gcnt, rnng_stats* st, double* ms) {
    const int n = (int)v->ds->n, m = (int)v->m, d = v->ds->d;
    Events ev;
    RNNG_TRY(ev.create());
    RNNG_HIP_TRY(hipEventRecord(ev.e[0], stream()));
    LocalBuf rad, lcnt, lptr, list, items, counter;
    RNNG_TRY(rad.reserve((size_t)m * sizeof(unsigned)));
    RNNG_TRY(lcnt.reserve((size_t)m * sizeof(int)));
    RNNG_TRY(launch_fill_u32(rad.as<unsigned>(), m, 0u));
    k_cell_radius<<<blocks_for(n, 256), 256, 0, stream()>>>(v->cell.as<int>(), v->delta.as<float>(), n,
                                                            rad.as<unsigned>());
    LocalBuf cnt_dev;
    RNNG_TRY(upload(cnt_dev, v->cnt));
    const size_t clds = (size_t)d * sizeof(float);
    k_center_lists<0><<<m, 256, clds, stream()>>>(v->C.as<float>(), m, d, cnt_dev.as<int>(), rad.as<unsigned>(),
                                                  eps, lcnt.as<int>(), nullptr, nullptr);
    RNNG_HIP_TRY(hipGetLastError());
    std::vector<int> lc;
    RNNG_TRY(download(lc, lcnt, (size_t)m));
    const std::vector<int> lp = prefix(lc);
    RNNG_TRY(upload(lptr, lp));
    RNNG_TRY(list.reserve((size_t)std::max(1, lp[m]) * sizeof(int)));
    k_center_lists<1><<<m, 256, clds, stream()>>>(v->C.as<float>(), m, d, cnt_dev.as<int>(), rad.as<unsigned>(),
                                                  eps, lcnt.as<int>(), lptr.as<int>(), list.as<int>());
    RNNG_HIP_TRY(hipGetLastError());
    std::vector<int4> it;
    int64_t tests = 0;
    for (int a = 0; a < m; ++a) {
        const int na = v->cnt[a];
        if (na == 0 || lc[a] == 0) continue;
        tests += (int64_t)na * lc[a];
        for (int t = 0; t * AB < na; ++t)
            for (int u = 0; u * AB < lc[a]; ++u) it.push_back(make_int4(a, t, u, 0));
    }
    RNNG_TRY(upload(items, it));
    RNNG_TRY(counter.reserve(sizeof(unsigned long long)));
    unsigned long long cap = std::max<unsigned long long>(1 << 16, (unsigned long long)n);
    int64_t ng = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        RNNG_TRY(v->gkeys.reserve(cap * sizeof(uint64_t)));
        RNNG_TRY(launch_fill_u64(counter.as<unsigned long long>(), 1, 0ull));
        if (!it.empty()) {
            k_ghost_test<<<(unsigned)it.size(), 256, 0, stream()>>>(
                v->ds->X, v->C.as<float>(), d, items.as<int4>(), v->cptr.as<int>(), v->cids.as<int>(),
                lptr.as<int>(), lcnt.as<int>(), list.as<int>(), v->delta.as<float>(), eps, v->gkeys.as<uint64_t>(),
                cap, counter.as<unsigned long long>());
            RNNG_HIP_TRY(hipGetLastError());
        }
        unsigned long long got = 0;
        RNNG_HIP_TRY(hipMemcpyAsync(&got, counter.ptr, sizeof(got), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        if (got <= cap) {
            ng = (int64_t)got;
            break;
        }
        if (st) st->refills++;
        cap = got + got / 8 + 1024;
        if (attempt == 1) {
            set_error("ghost buffer refill did not converge");
            return RNNG_ERR_HIP;
        }
    }
    RNNG_HIP_TRY(hipEventRecord(ev.e[1], stream()));
    RNNG_TRY(sort_keys_u64(v->gkeys.as<uint64_t>(), ng, 64));
    // per-cell candidate-ghost counts from the sorted keys (host side, m small)
    std::vector<uint64_t> hk;
    RNNG_TRY(download(hk, v->gkeys, (size_t)ng));
    gcnt.assign(m, 0);
    for (uint64_t k : hk) gcnt[(int)(k >> 32)]++;
    v->n_gcand = ng;
    RNNG_HIP_TRY(hipEventRecord(ev.e[2], stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    if (st) {
        st->ghost_tests += tests;
        st->ghost_candidates += ng;
    }
    *ms = ev.ms(0, 2);
    return RNNG_OK;
}

// Cell-query tuning: row tiles per item and the longest column run of an item.
struct CellTune {
    int rt = 1;
    int ct_max = 64;
};

template <int KS>
int launch_cell_filter(int rt, const DevBuf& xa, const DevBuf& yb, const DevBuf& items, int n_items,
                       const DevBuf& info, DevBuf& cand, unsigned long long cap, DevBuf& counter) {
    const unsigned grid = (unsigned)std::max(1, std::min((n_items + 3) / 4, 256 * 8));
    if (rt == 2)
        k_cell_filter<KS, 2><<<grid, 256, 0, stream()>>>(xa.as<bf16x8_t>(), yb.as<bf16x8_t>(), items.as<int4>(),
                                                         n_items, info.as<CellInfo>(), cand.as<int2>(), cap,
                                                         counter.as<unsigned long long>());
    else
        k_cell_filter<KS, 1><<<grid, 256, 0, stream()>>>(xa.as<bf16x8_t>(), yb.as<bf16x8_t>(), items.as<int4>(),
                                                         n_items, info.as<CellInfo>(), cand.as<int2>(), cap,
                                                         counter.as<unsigned long long>());
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

int dispatch_cell_filter(int KS, int rt, const DevBuf& xa, const DevBuf& yb, const DevBuf& items, int n_items,
                         const DevBuf& info, DevBuf& cand, unsigned long long cap, DevBuf& counter) {
    switch (KS) {
#define RNNG_KS(k) \
    case k: return launch_cell_filter<k>(rt, xa, yb, items, n_items, info, cand, cap, counter);
        RNNG_KS(1) RNNG_KS(2) RNNG_KS(3) RNNG_KS(4) RNNG_KS(5) RNNG_KS(6) RNNG_KS(7) RNNG_KS(8) RNNG_KS(9)
#undef RNNG_KS
        default:
            set_error("MFMA mode supports d <= 135");
            return RNNG_ERR_UNSUPPORTED;
    }
}

int voronoi_graph(rnng_voronoi_s* v, float eps, rnng_graph_s* g, rnng_stats* st) {
    const int n = (int)v->ds->n, m = (int)v->m, d = v->ds->d;
    const int KS = mfma_ksteps(d);
    if (KS > kMaxKS) {
        set_error("MFMA mode supports d <= 135");
        return RNNG_ERR_UNSUPPORTED;
    }
    const float T = sq_threshold_host(eps);
    const float eps_band = (float)((double)eps * (1.0 + 1e-5));
    float Tb = sq_threshold_host(std::nextafter(eps_band, INFINITY));
    if (Tb < T) Tb = T;

    std::vector<int> gcnt;
    double ghost_ms = 0.0;
    RNNG_TRY(ghost_candidates(v, eps, gcnt, st, &ghost_ms));

    // ---- slot layout and work items ----
    const CellTune tune;
    std::vector<int> R(m), Cc(m);
    for (int b = 0; b < m; ++b) {
        R[b] = (v->cnt[b] + 31) & ~31;
        Cc[b] = (v->cnt[b] + gcnt[b] + 31) & ~31;
    }
    const std::vector<int> rs = prefix(R), cs = prefix(Cc);
    const std::vector<int> gp = prefix(gcnt);
    std::vector<CellInfo> info(m);
    std::vector<int> colcell(std::max(1, cs[m] / 32));
    int64_t pairs = 0;
    struct Item {
        int4 w;
        int64_t cost;
    };
    std::vector<Item> items;
    for (int b = 0; b < m; ++b) {
        const int nb = v->cnt[b], gb = gcnt[b];
        info[b] = CellInfo{rs[b], cs[b], nb, nb + gb};
        for (int t = cs[b] / 32; t < cs[b + 1] / 32; ++t) colcell[t] = b;
        pairs += (int64_t)nb * (nb - 1) / 2 + (int64_t)nb * gb;
        if (nb == 0 || (nb == 1 && gb == 0)) continue;
        const int ntile_r = R[b] / 32, c_end = (cs[b] + nb + gb + 31) / 32;
        for (int t = 0; t < ntile_r; t += tune.rt) {
            const int c_begin = cs[b] / 32 + t;   // columns left of the diagonal tile hold no pairs
            for (int c0 = c_begin; c0 < c_end; c0 += tune.ct_max) {
                const int c1 = std::min(c_end, c0 + tune.ct_max);
                items.push_back(Item{make_int4(rs[b] / 32 + t, c0, c1, b), (int64_t)(c1 - c0) * tune.rt});
            }
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.cost > y.cost; });
    std::vector<int4> witems(items.size());
    for (size_t i = 0; i < items.size(); ++i) witems[i] = items[i].w;

    const int n_rowpad = rs[m] + 32 * tune.rt;   // a trailing item may read rt - 1 tiles past the last cell
    const int n_colpad = std::max(cs[m], 32);
    LocalBuf d_rs, d_cs, d_gptr, d_rowmap, d_colmap, d_info, d_items, d_colcell;
    RNNG_TRY(upload(d_rs, rs));
    RNNG_TRY(upload(d_cs, cs));
    RNNG_TRY(upload(d_gptr, gp));
    RNNG_TRY(upload(d_info, info));
    RNNG_TRY(upload(d_items, witems));
    RNNG_TRY(upload(d_colcell, colcell));
    RNNG_TRY(d_rowmap.reserve((size_t)n_rowpad * sizeof(int)));
    RNNG_TRY(d_colmap.reserve((size_t)n_colpad * sizeof(int)));
    RNNG_TRY(launch_fill_u32(d_rowmap.as<unsigned>(), n_rowpad, 0xffffffffu));
    RNNG_TRY(launch_fill_u32(d_colmap.as<unsigned>(), n_colpad, 0xffffffffu));
    k_slot_maps<<<m, 256, 0, stream()>>>(v->cptr.as<int>(), v->cids.as<int>(), d_gptr.as<int>(),
                                         v->gkeys.as<uint64_t>(), d_rs.as<int>(), d_cs.as<int>(), d_rowmap.as<int>(),
                                         d_colmap.as<int>());
    RNNG_HIP_TRY(hipGetLastError());

    Events ev;
    RNNG_TRY(ev.create());
    RNNG_HIP_TRY(hipEventRecord(ev.e[0], stream()));
    MfmaPrep P;
    RNNG_TRY(mfma_prepare_mapped(P, v->ds->X, n, d, Tb, d_rowmap.as<int>(), n_rowpad, d_colmap.as<int>(),
                                 n_colpad));
    RNNG_HIP_TRY(hipEventRecord(ev.e[1], stream()));

    LocalBuf counters, cand, cid, cflag, cw, colreal, ccand, cedge;
    RNNG_TRY(counters.reserve(8 * sizeof(unsigned long long)));
    unsigned long long cap = std::max<unsigned long long>(1 << 20, 16ull * (unsigned long long)n);
    unsigned long long ncand = 0;
    for (int attempt = 0; attempt < 3; ++attempt) {
        RNNG_TRY(cand.reserve(cap * sizeof(int2)));
        RNNG_TRY(launch_fill_u64(counters.as<unsigned long long>(), 8, 0ull));
        RNNG_HIP_TRY(hipEventRecord(ev.e[2], stream()));
        RNNG_TRY(dispatch_cell_filter(KS, tune.rt, P.xa, P.yb, d_items, (int)witems.size(), d_info, cand, cap,
                                      counters));
        RNNG_HIP_TRY(hipEventRecord(ev.e[3], stream()));
        RNNG_HIP_TRY(hipMemcpyAsync(&ncand, counters.ptr, sizeof(ncand), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        if (ncand <= cap) break;
        if (st) st->refills++;
        cap = ncand + ncand / 8 + 1024;
        if (attempt == 2) {
            set_error("candidate buffer refill did not converge");
            return RNNG_ERR_HIP;
        }
    }
    const size_t nc1 = std::max<size_t>(1, ncand);
    RNNG_TRY(cid.reserve(nc1 * sizeof(int2)));
    RNNG_TRY(cflag.reserve(nc1));
    RNNG_TRY(cw.reserve(nc1 * sizeof(float)));
    RNNG_TRY(colreal.reserve((size_t)n_colpad));
    RNNG_TRY(ccand.reserve((size_t)m * sizeof(int)));
    RNNG_TRY(cedge.reserve((size_t)m * sizeof(int)));
    RNNG_HIP_TRY(hipEventRecord(ev.e[4], stream()));
    const unsigned egrid = (unsigned)std::max<unsigned long long>(1, std::min<unsigned long long>(4096, (ncand + 255) / 256));
    k_cand_ids<<<egrid, 256, 0, stream()>>>(cand.as<int2>(), counters.as<unsigned long long>(), cap,
                                            d_rowmap.as<int>(), d_colmap.as<int>(), cid.as<int2>());
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_TRY(launch_recheck(v->ds->X, v->ds->X, d, cid.as<int2>(), cap, ncand, T, eps, cflag.as<uint8_t>(),
                            cw.as<float>(), counters.as<unsigned long long>()));
    RNNG_TRY(launch_fill_u32(ccand.as<unsigned>(), m, 0u));
    RNNG_TRY(launch_fill_u32(cedge.as<unsigned>(), m, 0u));
    {
        // colreal as bytes: clear with a u32 fill over the rounded-up size
        RNNG_TRY(colreal.reserve((size_t)((n_colpad + 3) & ~3)));
        RNNG_TRY(launch_fill_u32(colreal.as<unsigned>(), (n_colpad + 3) / 4, 0u));
    }
    k_cell_edges<<<egrid, 256, 0, stream()>>>(cand.as<int2>(), cid.as<int2>(), cflag.as<uint8_t>(),
                                              counters.as<unsigned long long>(), cap, d_colcell.as<int>(),
                                              d_info.as<CellInfo>(), colreal.as<uint8_t>(), ccand.as<int>(),
                                              cedge.as<int>());
    RNNG_HIP_TRY(hipGetLastError());

    // ---- edges: compact, canonical (min, max), sort, unique ----
    if ((int64_t)ncand > g->cap) RNNG_TRY(edge_buffers(g, (int64_t)ncand));
    int64_t ne = 0;
    RNNG_TRY(compact_edges(cid.as<int2>(), cw.as<float>(), cflag.as<uint8_t>(), (int64_t)ncand, g->ij.as<int2>(),
                           g->w.as<float>(), &ne));
    int64_t nu = 0;
    if (ne > 0) {
        LocalBuf keys, uflag, tmp_ij;
        RNNG_TRY(keys.reserve((size_t)ne * sizeof(uint64_t)));
        RNNG_TRY(uflag.reserve((size_t)ne));
        RNNG_TRY(tmp_ij.reserve((size_t)ne * sizeof(int2)));
        k_edge_keys<<<blocks_for(ne, 256), 256, 0, stream()>>>(g->ij.as<int2>(), ne, keys.as<uint64_t>());
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(sort_pairs_u64_f32(keys.as<uint64_t>(), g->w.as<float>(), ne));
        k_unique_flags<<<blocks_for(ne, 256), 256, 0, stream()>>>(keys.as<uint64_t>(), ne, uflag.as<uint8_t>());
        k_keys_to_ij<<<blocks_for(ne, 256), 256, 0, stream()>>>(keys.as<uint64_t>(), ne, tmp_ij.as<int2>());
        RNNG_HIP_TRY(hipGetLastError());
        LocalBuf tmp_w;
        RNNG_TRY(tmp_w.reserve((size_t)ne * sizeof(float)));
        RNNG_HIP_TRY(hipMemcpyAsync(tmp_w.ptr, g->w.ptr, (size_t)ne * sizeof(float), hipMemcpyDeviceToDevice,
                                    stream()));
        RNNG_TRY(compact_edges(tmp_ij.as<int2>(), tmp_w.as<float>(), uflag.as<uint8_t>(), ne, g->ij.as<int2>(),
                               g->w.as<float>(), &nu));
    }
    g->n_edges = nu;

    // ---- ghosts: candidates with an edge to a member ----
    RNNG_TRY(v->greal.reserve((size_t)std::max<int64_t>(1, v->n_gcand)));
    if (v->n_gcand > 0) {
        k_ghost_flags<<<blocks_for(v->n_gcand, 256), 256, 0, stream()>>>(
            v->gkeys.as<uint64_t>(), v->n_gcand, d_gptr.as<int>(), d_info.as<CellInfo>(), colreal.as<uint8_t>(),
            v->greal.as<uint8_t>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    std::vector<uint8_t> gr;
    RNNG_TRY(download(gr, v->greal, (size_t)v->n_gcand));
    v->n_ghosts = 0;
    for (uint8_t x : gr) v->n_ghosts += x;
    RNNG_HIP_TRY(hipEventRecord(ev.e[5], stream()));
    unsigned long long cnt[8];
    RNNG_HIP_TRY(hipMemcpyAsync(cnt, counters.ptr, sizeof(cnt), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    if (st) {
        st->pair_evals += pairs;
        st->candidates += (int64_t)ncand;
        st->edges += nu;
        st->band_pairs += (int64_t)cnt[2];
        st->band_flips += (int64_t)cnt[3];
        st->ghosts += v->n_ghosts;
        st->ghost_ms += ghost_ms;
        st->prep_ms += ev.ms(0, 1);
        st->filter_ms += ev.ms(2, 3);
        st->recheck_ms += ev.ms(4, 5);
        st->total_ms += ghost_ms + ev.ms(0, 1) + ev.ms(2, 3) + ev.ms(4, 5);
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
    auto* g = new rnng_graph_s();
    const int s = voronoi_graph(v, eps, g, st);
    if (s != RNNG_OK) {
        g->ij.release();
        g->w.release();
        delete g;
        return s;
    }
    *out = g;
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
    std::vector<uint8_t> r;
    RNNG_TRY(download(k, v->gkeys, (size_t)v->n_gcand));
    RNNG_TRY(download(r, v->greal, (size_t)v->n_gcand));
    int64_t o = 0;
    for (size_t e = 0; e < k.size(); ++e)
        if (r[e]) {
            pairs[2 * o] = (int64_t)(k[e] >> 32);
            pairs[2 * o + 1] = (int64_t)(k[e] & 0xffffffffull);
            ++o;
        }
    return RNNG_OK;
}

int rnng_voronoi_destroy(rnng_voronoi_t v) {
    if (!v) return RNNG_OK;
    for (DevBuf* b : {&v->ctr, &v->C, &v->cell, &v->delta, &v->cptr, &v->cids, &v->gkeys, &v->greal}) b->release();
    delete v;
    return RNNG_OK;
}

}  // extern "C"
