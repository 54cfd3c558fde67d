// Uniform-grid epsilon self join for points in at most three dimensions, for
// gfx950.
//
// Reference semantics (gabe-raulet/RNNG): the edge set is that of the
// brute-force radius query (include/search.hpp:389-407) with the L2 functor of
// drivers/rips.cpp:326-342 (sequential FP32 sum of fl(fl(p_k - q_k)^2), not
// fused, then sqrt; edge iff dist <= eps).  For low-dimensional inputs (the
// Rips driver's 3-D point clouds) almost every pair is far apart, so instead of
// the all-pairs tiles the points are bucketed into a dense grid of cubes of
// side h >= eps (1 + 2^-9): a pair within eps differs by at most one cell per
// coordinate, so only the 13 forward neighbour cells and the rest of a point's
// own cell are scanned.  Every scanned pair is decided by the same FP32
// expression and T(eps) threshold as the plain kernel, so the edge set is the
// plain kernel's.
//
// Margin: an FP32 distance <= eps bounds the true coordinate difference by
// eps (1 + 1e-6) (three terms, relative error below 5 ulp; h >= 2^-60 covers
// underflow), while cell coordinates are computed in FP64 from the bounding
// box, so two such points never land two cells apart.
//
// Pipeline: bounding box of the finite points -> cell key per point (FP64
// floor) -> radix sort by key -> dense cell start table -> per-point count of
// edges over five contiguous ranges of the sorted order -> exclusive scan ->
// fill.  Deterministic, no atomics on the edge list.  Points with a non-finite
// coordinate have no edge when eps is finite and are left out of the grid.
#include "internal.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace rnng {
namespace {

__device__ __forceinline__ unsigned ord_of(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
float float_of_ord(unsigned u) {
    const unsigned v = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &v, sizeof(f));
    return f;
}

// stat[0..2] = ordered min per dim, stat[3..5] = ordered max, stat[6] = finite count
__global__ __launch_bounds__(256) void k_grid_bbox(const float* __restrict__ X, int64_t n, int d,
                                                   unsigned* __restrict__ stat) {
    unsigned mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    unsigned fin = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float x[3];
        bool ok = true;
        for (int k = 0; k < d; ++k) {
            x[k] = X[i * d + k];
            ok = ok && isfinite(x[k]);
        }
        if (!ok) continue;
        ++fin;
        for (int k = 0; k < d; ++k) {
            mn[k] = min(mn[k], ord_of(x[k]));
            mx[k] = max(mx[k], ord_of(x[k]));
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = min(mn[k], (unsigned)__shfl_down((int)mn[k], off, 64));
            mx[k] = max(mx[k], (unsigned)__shfl_down((int)mx[k], off, 64));
        }
        fin += (unsigned)__shfl_down((int)fin, off, 64);
    }
    // the block's four waves meet in LDS: one set of atomics per block
    __shared__ unsigned part[4][7];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int k = 0; k < 3; ++k) {
            part[wv][k] = mn[k];
            part[wv][3 + k] = mx[k];
        }
        part[wv][6] = fin;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 4; ++q)
            for (int k = 0; k < 3; ++k) {
                mn[k] = min(mn[k], part[q][k]);
                mx[k] = max(mx[k], part[q][3 + k]);
            }
        for (int q = 1; q < 4; ++q) fin += part[q][6];
        for (int k = 0; k < d; ++k) {
            atomicMin(&stat[k], mn[k]);
            atomicMax(&stat[3 + k], mx[k]);
        }
        atomicAdd(&stat[6], fin);
    }
}

struct GridGeom {
    double lo[3];
    double h;
    int g[3];        // cells per dimension (1 beyond d)
    unsigned cells;  // g[0] g[1] g[2]
};

// key = linear cell index (x fastest); non-finite points get key = cells.
__global__ __launch_bounds__(256) void k_grid_keys(const float* __restrict__ X, int64_t n, int d, GridGeom gg,
                                                   uint32_t* __restrict__ key, int* __restrict__ id) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c[3] = {0, 0, 0};
    bool ok = true;
    for (int k = 0; k < d; ++k) {
        const float x = X[i * d + k];
        ok = ok && isfinite(x);
        if (ok) {
            const double t = floor(((double)x - gg.lo[k]) / gg.h);
            c[k] = (int)fmin(fmax(t, 0.0), (double)(gg.g[k] - 1));
        }
    }
    key[i] = ok ? (uint32_t)(((unsigned)c[2] * (unsigned)gg.g[1] + (unsigned)c[1]) * (unsigned)gg.g[0] + (unsigned)c[0])
                : gg.cells;
    id[i] = (int)i;
}

// Sorted points padded to float4 (dims >= d are zero and add exact zeros).
__global__ void k_grid_gather(const float* __restrict__ X, int d, const int* __restrict__ id, int64_t n,
                              float4* __restrict__ P) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float* x = X + (int64_t)id[p] * d;
    P[p] = make_float4(x[0], d > 1 ? x[1] : 0.0f, d > 2 ? x[2] : 0.0f, 0.0f);
}

// cstart[c] = first sorted position with key >= c, for c = 0..cells.
__global__ void k_grid_starts(const uint32_t* __restrict__ key, int64_t n, uint32_t cells,
                              int* __restrict__ cstart) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > n) return;
    const int64_t cur = p < n ? (int64_t)min(key[p], cells) : (int64_t)cells;
    const int64_t prev = p > 0 ? (int64_t)min(key[p - 1], cells) : -1;
    for (int64_t c = prev + 1; c <= cur; ++c) cstart[c] = (int)p;
}

// Edges of sorted point p with every later point of its own cell and every
// point of the 13 forward neighbour cells, which form five contiguous ranges
// of the sorted order: (own cell rest + x+1), (y+1 row), and three z+1 rows.
// FILL = false counts (and adds evaluated pairs to evals), FILL = true writes
// at offset[p].
template <bool FILL>
__global__ __launch_bounds__(256) void k_grid_join(
    const float4* __restrict__ P, const int* __restrict__ id, const uint32_t* __restrict__ key,
    const int* __restrict__ cstart, int nf, int d, int g0, int g1, int g2, float T,
    int64_t* __restrict__ count, unsigned long long* __restrict__ evals, const int64_t* __restrict__ offset,
    int2* __restrict__ out_ij, float* __restrict__ out_w) {
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long ev = 0;
    if (p < nf) {
        const float4 a = P[p];
        const int ia = id[p];
        const unsigned c = key[p];
        const int cx = (int)(c % (unsigned)g0);
        const int cy = (int)((c / (unsigned)g0) % (unsigned)g1);
        const int cz = (int)(c / ((unsigned)g0 * (unsigned)g1));
        const int xl = cx > 0 ? 1 : 0, xh = cx + 1 < g0 ? 1 : 0;
        int lo[5], hi[5];
        int nr = 0;
        lo[nr] = p + 1;
        hi[nr++] = cstart[c + 1 + xh];
        if (cy + 1 < g1) {
            const unsigned r = c + (unsigned)g0;
            lo[nr] = cstart[r - xl];
            hi[nr++] = cstart[r + xh + 1];
        }
        if (cz + 1 < g2) {
            for (int dy = -1; dy <= 1; ++dy) {
                if (cy + dy < 0 || cy + dy >= g1) continue;
                const unsigned r = c + (unsigned)g0 * (unsigned)g1 + (unsigned)(dy * g0);
                lo[nr] = cstart[r - xl];
                hi[nr++] = cstart[r + xh + 1];
            }
        }
        int64_t cnt = 0, out = FILL ? offset[p] : 0;
        for (int r = 0; r < nr; ++r) {
            ev += (unsigned long long)max(0, hi[r] - lo[r]);
            for (int q = lo[r]; q < hi[r]; ++q) {
                const float4 b = P[q];
                float s = 0.0f, t;
                t = a.x - b.x;
                s = s + t * t;
                if (d > 1) {
                    t = a.y - b.y;
                    s = s + t * t;
                }
                if (d > 2) {
                    t = a.z - b.z;
                    s = s + t * t;
                }
                if (s <= T) {
                    if (FILL) {
                        const int ib = id[q];
                        out_ij[out] = make_int2(min(ia, ib), max(ia, ib));
                        out_w[out] = sqrtf(s);
                        ++out;
                    } else {
                        ++cnt;
                    }
                }
            }
        }
        if (!FILL) count[p] = cnt;
    }
    if (!FILL) {
        for (int off = 32; off > 0; off >>= 1) ev += __shfl_down(ev, off, 64);
        if ((threadIdx.x & 63) == 0 && ev) atomicAdd(evals, ev);
    }
}

__global__ void k_grid_clear(unsigned* stat) {
    const int k = threadIdx.x;
    if (k < 3) stat[k] = 0xffffffffu;
    else if (k < 7) stat[k] = 0u;
}

__global__ void k_grid_clear_u64(unsigned long long* p) { *p = 0ull; }

}  // namespace

int grid_graph_build(const rnng_dataset_s* A, float eps, rnng_graph_s* g, rnng_stats* st) {
    const int64_t n = A->n;
    const int d = A->d;
    if (d > 3) {
        set_error("grid mode supports d <= 3");
        return RNNG_ERR_UNSUPPORTED;
    }
    if (n > INT32_MAX - 1) {
        set_error("grid mode supports fewer than 2^31 - 1 points");
        return RNNG_ERR_UNSUPPORTED;
    }
    const float T = sq_threshold_host(eps);
    Events ev;
    RNNG_TRY(ev.create());
    RNNG_HIP_TRY(hipEventRecord(ev.e[0], stream()));
    if (!(T >= 0.0f) || n < 2) {   // negative or NaN eps: no pair qualifies
        RNNG_TRY(edge_buffers(g, 1));
        g->n_edges = 0;
        return RNNG_OK;
    }

    LocalBuf stat;
    RNNG_TRY(stat.reserve(8 * sizeof(unsigned)));
    k_grid_clear<<<1, 64, 0, stream()>>>(stat.as<unsigned>());
    k_grid_bbox<<<std::min<unsigned>(blocks_for(n, 256), 256), 256, 0, stream()>>>(A->X, n, d, stat.as<unsigned>());
    RNNG_HIP_TRY(hipGetLastError());
    unsigned hs[8];
    RNNG_HIP_TRY(hipMemcpyAsync(hs, stat.ptr, sizeof(hs), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    const int64_t nf = hs[6];

    // Cell side: at least eps (1 + 2^-9) and 2^-60, grown until the dense
    // grid has at most max(2^20, 2 n) cells.
    GridGeom gg{};
    double ext[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 3; ++k) {
        gg.lo[k] = 0.0;
        if (k < d && nf > 0) {
            gg.lo[k] = (double)float_of_ord(hs[k]);
            ext[k] = (double)float_of_ord(hs[3 + k]) - gg.lo[k];
        }
    }
    double h = std::max((double)eps * (1.0 + 1.0 / 512.0), std::ldexp(1.0, -60));
    if (!std::isfinite(h)) h = DBL_MAX;
    const double limit = (double)std::max<int64_t>(1 << 20, 2 * n);
    for (;;) {
        double cells = 1.0;
        for (int k = 0; k < 3; ++k) {
            const double gk = std::floor(ext[k] / h) + 1.0;
            gg.g[k] = (int)std::min(gk, 2147483647.0);
            cells *= gk;
        }
        if (cells <= limit) {
            gg.cells = (unsigned)cells;
            break;
        }
        h *= std::max(1.01, std::pow(cells / limit, 1.0 / d));
    }
    gg.h = h;
    unsigned end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) <= (unsigned long long)gg.cells) ++end_bit;

    LocalBuf key, id, P, cstart, cnt, off;
    RNNG_TRY(key.reserve((size_t)n * sizeof(uint32_t)));
    RNNG_TRY(id.reserve((size_t)n * sizeof(int)));
    RNNG_TRY(P.reserve((size_t)n * sizeof(float4)));
    RNNG_TRY(cstart.reserve(((size_t)gg.cells + 1) * sizeof(int)));
    RNNG_TRY(cnt.reserve((size_t)std::max<int64_t>(nf, 1) * sizeof(int64_t)));
    RNNG_TRY(off.reserve((size_t)std::max<int64_t>(nf, 1) * sizeof(int64_t)));
    k_grid_keys<<<blocks_for(n, 256), 256, 0, stream()>>>(A->X, n, d, gg, key.as<uint32_t>(), id.as<int>());
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_TRY(sort_pairs_u32_i32(key.as<uint32_t>(), id.as<int>(), n, end_bit));
    k_grid_gather<<<blocks_for(n, 256), 256, 0, stream()>>>(A->X, d, id.as<int>(), n, P.as<float4>());
    k_grid_starts<<<blocks_for(n + 1, 256), 256, 0, stream()>>>(key.as<uint32_t>(), n, gg.cells,
                                                               cstart.as<int>());
    RNNG_HIP_TRY(hipGetLastError());

    int64_t ne = 0;
    unsigned long long evals = 0;
    if (nf > 0) {
        LocalBuf evc;
        RNNG_TRY(evc.reserve(sizeof(unsigned long long)));
        k_grid_clear_u64<<<1, 1, 0, stream()>>>(evc.as<unsigned long long>());
        k_grid_join<false><<<blocks_for(nf, 256), 256, 0, stream()>>>(
            P.as<float4>(), id.as<int>(), key.as<uint32_t>(), cstart.as<int>(), (int)nf, d, gg.g[0], gg.g[1],
            gg.g[2], T, cnt.as<int64_t>(), evc.as<unsigned long long>(), nullptr, nullptr, nullptr);
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(exclusive_scan_i64(cnt.as<int64_t>(), off.as<int64_t>(), nf));
        int64_t lo = 0, lc = 0;
        RNNG_HIP_TRY(hipMemcpyAsync(&lo, off.as<int64_t>() + (nf - 1), sizeof(lo), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipMemcpyAsync(&lc, cnt.as<int64_t>() + (nf - 1), sizeof(lc), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipMemcpyAsync(&evals, evc.ptr, sizeof(evals), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        ne = lo + lc;
    }
    if (ne > g->cap || g->cap == 0) RNNG_TRY(edge_buffers(g, std::max<int64_t>(ne, 1)));
    if (ne > 0) {
        k_grid_join<true><<<blocks_for(nf, 256), 256, 0, stream()>>>(
            P.as<float4>(), id.as<int>(), key.as<uint32_t>(), cstart.as<int>(), (int)nf, d, gg.g[0], gg.g[1],
            gg.g[2], T, nullptr, nullptr, off.as<int64_t>(), g->ij.as<int2>(), g->w.as<float>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    RNNG_HIP_TRY(hipEventRecord(ev.e[1], stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    g->n_edges = ne;
    if (st) {
        st->pair_evals += (int64_t)evals;
        st->edges += ne;
        st->filter_ms += ev.ms(0, 1);
        st->total_ms += ev.ms(0, 1);
    }
    return RNNG_OK;
}

}  // namespace rnng
