// Vietoris-Rips filtration (vertices, edges, triangles, tetrahedra) of the
// L2 epsilon graph, for gfx950.
//
// Reference semantics (gabe-raulet/RNNG):
//   VoronoiComplex::build_filtration  include/point.hpp:1134-1232
//     - cliques of the epsilon graph with at most maxdim+1 vertices
//       (bron_kerbosch, :1234-1267), each enumerated once, vertices ascending
//     - value: 0 for a vertex, distance() for an edge (:1182-1183), the max
//       over facet values for higher simplices (:1194-1208)
//   Simplex id  include/simplex.hpp:5-19: sum_i C(v_i, i+1) | (p << 60)
//   merge_and_write_filtration  include/simplex.hpp:130-196: sort by
//       (value, id), unique by id
// The union over Voronoi cells with ghosts is the Rips complex of the whole
// point set (every clique lies in the complex of the cell of any of its
// vertices), so one GPU builds it directly from the global epsilon graph.
//
// Pipeline:
//  1. epsilon graph (uniform grid for d <= 3, MFMA filter for d >= 16, the
//     plain kernel otherwise);
//  2. lower CSR: row = larger vertex, columns = smaller neighbours ascending.
//     Its edge order is (larger, smaller) lexicographic, which is the order of
//     the edge ids C(j, 2) + i ("colex");
//  3. value groups: edges sorted by (weight, id) with one stable 32-bit sort;
//     every distinct value gets a dense rank (group), vertices are group 0;
//  4. cliques in colex order: triangles (u, v, w) through lower edge (v, w)
//     from the merge of the two lower lists, tetrahedra (u, v, w, x) through
//     (w, x) from a merge for v and a three-way merge for u.  Edges are visited
//     in colex order and each edge's cliques come out ascending, so every
//     clique array is already sorted by id.  A clique's group is the largest
//     group of its edges (its value is the largest edge weight);
//  5. order: one stable sort of each clique array by group (ceil(log2 G)
//     bits) turns id order into (value, id) order; per-group counts and one
//     scan place vertices, edges, triangles and tetrahedra of each group
//     (p << 60 in the id puts them in that order inside a group).
// Count, scan, fill throughout: deterministic, no atomics.
#include "internal.hpp"

#include <algorithm>
#include <vector>

struct rnng_filtration_s {
    rnng::DevBuf value;   // float
    rnng::DevBuf id;      // int64
    int64_t n = 0;
    int64_t counts[4] = {0, 0, 0, 0};   // simplices per dimension
};

namespace rnng {
namespace {

__device__ __forceinline__ int64_t binom2(int64_t v) { return v * (v - 1) / 2; }
__device__ __forceinline__ int64_t binom3(int64_t v) { return (v * (v - 1) / 2) * (v - 2) / 3; }
// C(v, 4) = C(v, 3) (v - 3) / 4, exact while C(v, 3) (v - 3) < 2^63
__device__ __forceinline__ int64_t binom4(int64_t v) { return binom3(v) * (v - 3) / 4; }

// C(v, k) for the simplex dimensions of a filtration (k = 1..4).
__device__ __forceinline__ int64_t binomk(int64_t v, int k) {
    return k == 1 ? v : k == 2 ? binom2(v) : k == 3 ? binom3(v) : binom4(v);
}

// Relabelling of a filtration built on n points: each simplex's vertices
// (decoded from its id, include/simplex.hpp:43-79: the largest l < n with
// C(l, i + 1) <= the remaining uid, for i = dim .. 0) become map[v], and the
// id is rebuilt from them (include/simplex.hpp:5-19); sflag = OR of vflag over
// the simplex's local vertices.
__global__ void k_relabel(int64_t* __restrict__ id, int64_t ns, const int64_t* __restrict__ map, int64_t n,
                          const uint8_t* __restrict__ vflag, uint8_t* __restrict__ sflag) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int64_t x = id[s];
    const int p = (int)((uint64_t)x >> 60) & 15;
    int64_t uid = x & 0x0FFFFFFFFFFFFFFFll;
    int64_t g[4];
    uint8_t f = 0;
    for (int i = p; i >= 0; --i) {
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = lo + ((hi - lo) >> 1);
            if (binomk(mid, i + 1) > uid) hi = mid;
            else lo = mid + 1;
        }
        const int64_t l = hi - 1;
        uid -= binomk(l, i + 1);
        g[i] = map[l];
        if (vflag) f |= vflag[l];
    }
    // ascending global vertices (at most four: insertion sort)
    for (int i = 1; i <= p; ++i)
        for (int j = i; j > 0 && g[j - 1] > g[j]; --j) {
            const int64_t t = g[j];
            g[j] = g[j - 1];
            g[j - 1] = t;
        }
    int64_t nu = 0;
    for (int i = 0; i <= p; ++i) nu += binomk(g[i], i + 1);
    id[s] = nu | ((int64_t)p << 60);
    if (sflag) sflag[s] = f;
}

// Largest vertex count whose 4-clique ids the reference computes without
// overflow: its Binom multiplies v (v-1) (v-2) (v-3) in int64
// (include/utils.hpp:126-127), which stays below 2^63 for v <= 55110.
constexpr int64_t kMaxTetraVertices = 55111;

constexpr int64_t kP1 = (int64_t)1 << 60, kP2 = (int64_t)2 << 60, kP3 = (int64_t)3 << 60;

// key = larger vertex << 32 | smaller vertex
__global__ void k_lower_keys(const int2* __restrict__ ij, uint64_t* __restrict__ keys, int64_t E) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) {
        const int2 p = ij[e];
        keys[e] = ((uint64_t)(uint32_t)max(p.x, p.y) << 32) | (uint32_t)min(p.x, p.y);
    }
}

// rowptr[r] = first edge of row r, col = smaller vertex, row = larger vertex.
__global__ void k_lower_csr(const uint64_t* __restrict__ keys, int64_t E, int n, int64_t* __restrict__ rowptr,
                            int* __restrict__ col, int* __restrict__ row) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e > E) return;
    const int r = e < E ? (int)(keys[e] >> 32) : n;
    const int pr = e > 0 ? (int)(keys[e - 1] >> 32) : -1;
    for (int x = pr + 1; x <= r; ++x) rowptr[x] = e;
    if (e < E) {
        col[e] = (int)(keys[e] & 0xffffffffu);
        row[e] = r;
    }
}

// Weights are non-negative, so their bit patterns order like the values.
__global__ void k_weight_keys(const float* __restrict__ w, int64_t E, uint32_t* __restrict__ bits,
                              uint32_t* __restrict__ perm) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) {
        bits[e] = __float_as_uint(w[e]);
        perm[e] = (uint32_t)e;
    }
}

__global__ void k_group_flags(const uint32_t* __restrict__ bits, int64_t E, int64_t* __restrict__ flag) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < E) flag[k] = (k == 0 || bits[k] != bits[k - 1]) ? 1 : 0;
}

// Dense value rank of every edge: group 0 is value 0 (the vertices, and the
// edges of weight 0 if any); vrS in sorted order, vrE by colex edge index,
// cv = group << 32 | smaller vertex by colex edge index, gval[group] = the
// group's value.
__global__ void k_edge_groups(const uint32_t* __restrict__ bits, const uint32_t* __restrict__ perm,
                              const int64_t* __restrict__ flag, const int64_t* __restrict__ excl, int64_t E,
                              const int* __restrict__ col, uint32_t* __restrict__ vrS, uint32_t* __restrict__ vrE,
                              uint64_t* __restrict__ cv, float* __restrict__ gval) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= E) return;
    const int64_t shift = bits[0] != 0u ? 1 : 0;
    const uint32_t g = (uint32_t)(excl[k] + flag[k] - 1 + shift);
    const uint32_t e = perm[k];
    vrS[k] = g;
    vrE[e] = g;
    cv[e] = ((uint64_t)g << 32) | (uint32_t)col[e];
    if (flag[k]) gval[g] = __uint_as_float(bits[k]);
    if (k == 0 && shift) gval[0] = 0.0f;
}

// Triangles (u, v, w), u < v < w, through lower edge e = (v, w): the common
// smaller neighbours u < v of w and v.  Eight lanes per edge: the vertex keys
// of the longer of the two sorted lists (lower(w) below v, lower(v)) are first
// copied into the group's LDS slot (up to LCAP, coalesced; a longer list is
// searched in global memory), then the shorter one is read in rounds of eight
// elements, each lane binary-searches its element in the longer list, and a
// round's matches are compacted in lane order, so each edge's triangles come
// out ascending in u and the group's stores are contiguous.  cv packs
// (group << 32 | vertex), so a probe that matches also yields the edge's
// value group.  FILL = false counts, FILL = true writes group and id at
// offset[e].
template <bool FILL>
__global__ __launch_bounds__(256) void k_triangles(
    const int* __restrict__ row, const int64_t* __restrict__ rowptr, const uint64_t* __restrict__ cv, int64_t E,
    int64_t* __restrict__ count, const int64_t* __restrict__ offset, uint32_t* __restrict__ tg,
    int64_t* __restrict__ tid) {
    constexpr int LCAP = 64;
    __shared__ uint32_t slot_all[32][LCAP];   // vertex keys, one slot per 8-lane group of the workgroup
    const int64_t e = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3;
    const int j = threadIdx.x & 7;
    const int gshift = threadIdx.x & 56;   // first lane of the group in the wave
    uint32_t* slot = slot_all[threadIdx.x >> 3];
    if (e >= E) return;                    // the 8 lanes of a group leave together
    const uint64_t ce = cv[e];
    const int v = (int)(uint32_t)ce, w = row[e];
    const uint32_t gvw = (uint32_t)(ce >> 32);
    const int64_t a0 = rowptr[w], a1 = e, b0 = rowptr[v], b1 = rowptr[v + 1];
    const bool sw = (a1 - a0) > (b1 - b0);
    const int64_t s0 = sw ? b0 : a0, s1 = sw ? b1 : a1, l0 = sw ? a0 : b0, l1 = sw ? a1 : b1;
    const int64_t L = l1 - l0;
    const bool staged = L <= LCAP;
    if (staged && s1 > s0)
        for (int64_t k = j; k < L; k += 8) slot[k] = (uint32_t)cv[l0 + k];
    // the group's lanes read what the others wrote: LDS is in order within a wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int64_t out = FILL ? offset[e] : 0;
    const int64_t base = binom2(v) + binom3(w) + kP2;
    for (int64_t r = s0; r < s1; r += 8) {
        const int64_t k = r + j;
        bool hit = false;
        uint64_t se = 0, le = 0;
        if (k < s1) {
            se = cv[k];
            const uint32_t key = (uint32_t)se;
            int64_t lo = 0, hi = L;
            if (staged) {
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (slot[mid] < key) lo = mid + 1;
                    else hi = mid;
                }
                hit = lo < L && slot[lo] == key;
                // the group of a match from global memory (only the fill needs it)
                if (FILL && hit) le = cv[l0 + lo];
            } else {
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if ((uint32_t)cv[l0 + mid] < key) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < L) le = cv[l0 + lo];
                hit = lo < L && (uint32_t)le == key;
            }
        }
        const uint32_t gm = (uint32_t)(__ballot(hit) >> gshift) & 0xffu;
        if (FILL && hit) {
            const int64_t pos = out + __popc(gm & ((1u << j) - 1u));
            tg[pos] = max(gvw, max((uint32_t)(se >> 32), (uint32_t)(le >> 32)));
            tid[pos] = base + (int64_t)(uint32_t)se;
        }
        out += __popc(gm);
    }
    if (!FILL && j == 0) count[e] = out;
}

// Tetrahedra (u, v, w, x), u < v < w < x, through lower edge e = (w, x): for
// each common smaller neighbour v of x and w (ascending), the common smaller
// neighbours u < v of x, w and v (ascending).
template <bool FILL>
__global__ __launch_bounds__(256) void k_tetrahedra(
    const int* __restrict__ row, const int64_t* __restrict__ rowptr, const int* __restrict__ col,
    const uint32_t* __restrict__ vrE, int64_t E, int64_t* __restrict__ count, const int64_t* __restrict__ offset,
    uint32_t* __restrict__ qg, int64_t* __restrict__ qid) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int x = row[e], w = col[e];
    const int64_t bx = rowptr[x], bw = rowptr[w], ew = rowptr[w + 1];
    int64_t i1 = bx, i2 = bw;
    int64_t c = 0, out = FILL ? offset[e] : 0;
    const uint32_t gwx = FILL ? vrE[e] : 0u;
    const int64_t base = binom3(w) + binom4(x) + kP3;
    while (i1 < e && i2 < ew) {
        const int a = col[i1], b = col[i2];
        if (a < b) {
            ++i1;
            continue;
        }
        if (b < a) {
            ++i2;
            continue;
        }
        const int v = a;
        const uint32_t g3 = FILL ? max(gwx, max(vrE[i1], vrE[i2])) : 0u;
        int64_t j1 = bx, j2 = bw, j3 = rowptr[v];
        const int64_t e3 = rowptr[v + 1];
        while (j1 < i1 && j2 < i2 && j3 < e3) {
            const int x1 = col[j1], x2 = col[j2], x3 = col[j3];
            const int mn = min(x1, min(x2, x3));
            if (x1 == x2 && x2 == x3) {
                if (FILL) {
                    qg[out] = max(g3, max(vrE[j1], max(vrE[j2], vrE[j3])));
                    qid[out] = base + binom2(v) + x1;
                    ++out;
                } else {
                    ++c;
                }
            }
            j1 += x1 == mn;
            j2 += x2 == mn;
            j3 += x3 == mn;
        }
        ++i1;
        ++i2;
    }
    if (!FILL) count[e] = c;
}

// gs[g] = first position of a group-sorted array with group >= g, g = 0..G.
__global__ void k_group_starts(const uint32_t* __restrict__ grp, int64_t L, uint32_t G, int64_t* __restrict__ gs) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > L) return;
    const int64_t cur = k < L ? (int64_t)grp[k] : (int64_t)G;
    const int64_t prev = k > 0 ? (int64_t)grp[k - 1] : -1;
    for (int64_t g = prev + 1; g <= cur; ++g) gs[g] = k;
}

// Simplices of group g: vertices (group 0 only), then edges, triangles and
// tetrahedra of the group.
__global__ void k_group_sizes(const int64_t* __restrict__ gsE, const int64_t* __restrict__ gsT,
                              const int64_t* __restrict__ gsQ, int64_t n, uint32_t G, int64_t* __restrict__ size) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    int64_t s = g == 0 ? n : 0;
    if (gsE) s += gsE[g + 1] - gsE[g];
    if (gsT) s += gsT[g + 1] - gsT[g];
    if (gsQ) s += gsQ[g + 1] - gsQ[g];
    size[g] = s;
}

__global__ void k_place_vertices(float* __restrict__ val, int64_t* __restrict__ id, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        val[i] = 0.0f;
        id[i] = i;
    }
}

__global__ void k_place_edges(const uint32_t* __restrict__ vrS, const uint32_t* __restrict__ perm,
                              const int* __restrict__ row, const int* __restrict__ col, int64_t E,
                              const int64_t* __restrict__ gsE, const int64_t* __restrict__ goff,
                              const float* __restrict__ gval, int64_t n, float* __restrict__ val,
                              int64_t* __restrict__ id) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= E) return;
    const uint32_t g = vrS[k];
    const int64_t pos = goff[g] + (g == 0 ? n : 0) + (k - gsE[g]);
    const uint32_t e = perm[k];
    val[pos] = gval[g];
    id[pos] = (int64_t)col[e] + binom2(row[e]) + kP1;
}

// Triangles (T = gsQ null) or tetrahedra of a group-sorted array.
__global__ void k_place_cliques(const uint32_t* __restrict__ grp, const int64_t* __restrict__ cid, int64_t L,
                                const int64_t* __restrict__ gs, const int64_t* __restrict__ gsE,
                                const int64_t* __restrict__ gsT, const int64_t* __restrict__ goff,
                                const float* __restrict__ gval, int64_t n, float* __restrict__ val,
                                int64_t* __restrict__ id) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= L) return;
    const uint32_t g = grp[k];
    int64_t pos = goff[g] + (g == 0 ? n : 0) + (gsE[g + 1] - gsE[g]) + (k - gs[g]);
    if (gsT) pos += gsT[g + 1] - gsT[g];
    val[pos] = gval[g];
    id[pos] = cid[k];
}

unsigned bits_for(uint64_t v) {   // smallest b with v < 2^b (at least 1)
    unsigned b = 1;
    while (b < 64 && (1ull << b) <= v) ++b;
    return b;
}

int last_total(const int64_t* cnt, const int64_t* off, int64_t L, int64_t* total) {
    int64_t lo = 0, lc = 0;
    RNNG_HIP_TRY(hipMemcpyAsync(&lo, off + (L - 1), sizeof(lo), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipMemcpyAsync(&lc, cnt + (L - 1), sizeof(lc), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    *total = lo + lc;
    return RNNG_OK;
}

int rips_build(const rnng_dataset_s* ds, float eps, int maxdim, rnng_filtration_s* f,
               rnng_stats* st) {
    const int n = (int)ds->n;
    if (maxdim < 0 || maxdim > 3) {
        set_error("maxdim must be 0, 1, 2 or 3");
        return RNNG_ERR_UNSUPPORTED;
    }
    if (maxdim == 3 && n > kMaxTetraVertices) {
        set_error("4-clique ids of more than 55,111 vertices overflow the reference's int64 binomials");
        return RNNG_ERR_UNSUPPORTED;
    }
    // C(v, 3) must fit the 60-bit id field (include/simplex.hpp:18)
    if (maxdim == 2 && n > 1800000) {
        set_error("triangle ids of more than 1.8M vertices exceed 60 bits");
        return RNNG_ERR_UNSUPPORTED;
    }
    Events ev;
    RNNG_TRY(ev.create());
    RNNG_HIP_TRY(hipEventRecord(ev.e[0], stream()));
    rnng_graph_s g;
    struct Release {
        rnng_graph_s& g;
        ~Release() {
            g.ij.release();
            g.w.release();
        }
    } release_g{g};
    int64_t E = 0;
    rnng_stats gst{};
    if (maxdim >= 1 && n >= 2) {
        const int mode = ds->d <= 3 ? RNNG_MODE_GRID : ds->d >= 16 ? RNNG_MODE_MFMA : RNNG_MODE_PLAIN;
        RNNG_TRY(l2_graph_build(ds, ds, true, eps, mode, &g, &gst));
        E = g.n_edges;
    }
    RNNG_HIP_TRY(hipEventRecord(ev.e[1], stream()));
    const int64_t E1 = std::max<int64_t>(E, 1);

    // 2. lower CSR in colex edge order; weights follow their edges
    LocalBuf keys, rowptr, col, row, wc;
    RNNG_TRY(rowptr.reserve(((size_t)n + 1) * sizeof(int64_t)));
    RNNG_TRY(col.reserve((size_t)E1 * sizeof(int)));
    RNNG_TRY(row.reserve((size_t)E1 * sizeof(int)));
    if (E > 0) {
        RNNG_TRY(keys.reserve((size_t)E * sizeof(uint64_t)));
        k_lower_keys<<<blocks_for(E, 256), 256, 0, stream()>>>(g.ij.as<int2>(), keys.as<uint64_t>(), E);
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(sort_pairs_u64_f32_swap(keys, g.w, E, 32 + bits_for((uint64_t)n)));
        k_lower_csr<<<blocks_for(E + 1, 256), 256, 0, stream()>>>(keys.as<uint64_t>(), E, n, rowptr.as<int64_t>(),
                                                                 col.as<int>(), row.as<int>());
        RNNG_HIP_TRY(hipGetLastError());
        keys.release();
    }

    // 3. value groups
    LocalBuf bits, perm, flag, excl, vrS, vrE, cv, gval;
    uint32_t G = 1;
    RNNG_TRY(gval.reserve((size_t)(E + 1) * sizeof(float)));
    if (E > 0) {
        RNNG_TRY(bits.reserve((size_t)E * sizeof(uint32_t)));
        RNNG_TRY(perm.reserve((size_t)E * sizeof(uint32_t)));
        RNNG_TRY(flag.reserve((size_t)E * sizeof(int64_t)));
        RNNG_TRY(excl.reserve((size_t)E * sizeof(int64_t)));
        RNNG_TRY(vrS.reserve((size_t)E * sizeof(uint32_t)));
        RNNG_TRY(vrE.reserve((size_t)E * sizeof(uint32_t)));
        RNNG_TRY(cv.reserve((size_t)E * sizeof(uint64_t)));
        k_weight_keys<<<blocks_for(E, 256), 256, 0, stream()>>>(g.w.as<float>(), E, bits.as<uint32_t>(),
                                                                perm.as<uint32_t>());
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(sort_pairs_u32_u32_swap(bits, perm, E, 32));
        k_group_flags<<<blocks_for(E, 256), 256, 0, stream()>>>(bits.as<uint32_t>(), E, flag.as<int64_t>());
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(exclusive_scan_i64(flag.as<int64_t>(), excl.as<int64_t>(), E));
        k_edge_groups<<<blocks_for(E, 256), 256, 0, stream()>>>(bits.as<uint32_t>(), perm.as<uint32_t>(),
                                                                flag.as<int64_t>(), excl.as<int64_t>(), E,
                                                                col.as<int>(), vrS.as<uint32_t>(),
                                                                vrE.as<uint32_t>(), cv.as<uint64_t>(),
                                                                gval.as<float>());
        RNNG_HIP_TRY(hipGetLastError());
        int64_t groups = 0;
        RNNG_TRY(last_total(flag.as<int64_t>(), excl.as<int64_t>(), E, &groups));
        uint32_t b0 = 0;
        RNNG_HIP_TRY(hipMemcpyAsync(&b0, bits.ptr, sizeof(b0), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        G = (uint32_t)(groups + (b0 != 0u ? 1 : 0));
        flag.release();
        excl.release();
    }
    const unsigned gbits = bits_for(G);

    // 4. cliques in colex (= id) order
    LocalBuf cntT, offT, tg, tid, cntQ, offQ, qg, qid;
    int64_t T = 0, Q = 0;
    if (maxdim >= 2 && E > 0) {
        RNNG_TRY(cntT.reserve((size_t)E * sizeof(int64_t)));
        RNNG_TRY(offT.reserve((size_t)E * sizeof(int64_t)));
        k_triangles<false><<<blocks_for(8 * E, 256), 256, 0, stream()>>>(
            row.as<int>(), rowptr.as<int64_t>(), cv.as<uint64_t>(), E, cntT.as<int64_t>(), nullptr, nullptr,
            nullptr);
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(exclusive_scan_i64(cntT.as<int64_t>(), offT.as<int64_t>(), E));
        RNNG_TRY(last_total(cntT.as<int64_t>(), offT.as<int64_t>(), E, &T));
    }
    if (maxdim >= 3 && E > 0) {
        RNNG_TRY(cntQ.reserve((size_t)E * sizeof(int64_t)));
        RNNG_TRY(offQ.reserve((size_t)E * sizeof(int64_t)));
        k_tetrahedra<false><<<blocks_for(E, 256), 256, 0, stream()>>>(row.as<int>(), rowptr.as<int64_t>(),
                                                                      col.as<int>(), nullptr, E, cntQ.as<int64_t>(),
                                                                      nullptr, nullptr, nullptr);
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(exclusive_scan_i64(cntQ.as<int64_t>(), offQ.as<int64_t>(), E));
        RNNG_TRY(last_total(cntQ.as<int64_t>(), offQ.as<int64_t>(), E, &Q));
    }
    if (T > 0) {
        RNNG_TRY(tg.reserve((size_t)T * sizeof(uint32_t)));
        RNNG_TRY(tid.reserve((size_t)T * sizeof(int64_t)));
        RNNG_HIP_TRY(hipEventRecord(ev.e[4], stream()));
        k_triangles<true><<<blocks_for(8 * E, 256), 256, 0, stream()>>>(
            row.as<int>(), rowptr.as<int64_t>(), cv.as<uint64_t>(), E, nullptr, offT.as<int64_t>(),
            tg.as<uint32_t>(), tid.as<int64_t>());
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_HIP_TRY(hipEventRecord(ev.e[5], stream()));
    }
    if (Q > 0) {
        RNNG_TRY(qg.reserve((size_t)Q * sizeof(uint32_t)));
        RNNG_TRY(qid.reserve((size_t)Q * sizeof(int64_t)));
        k_tetrahedra<true><<<blocks_for(E, 256), 256, 0, stream()>>>(row.as<int>(), rowptr.as<int64_t>(),
                                                                     col.as<int>(), vrE.as<uint32_t>(), E, nullptr,
                                                                     offQ.as<int64_t>(), qg.as<uint32_t>(),
                                                                     qid.as<int64_t>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    cntT.release();
    offT.release();
    cntQ.release();
    offQ.release();
    RNNG_HIP_TRY(hipEventRecord(ev.e[2], stream()));

    // 5. (value, id) order: stable group sorts, per-group counts, placement
    if (T > 0) RNNG_TRY(sort_pairs_u32_i64_swap(tg, tid, T, gbits));
    if (Q > 0) RNNG_TRY(sort_pairs_u32_i64_swap(qg, qid, Q, gbits));
    const int64_t Ecount = maxdim >= 1 ? E : 0;
    const int64_t total = (int64_t)n + Ecount + T + Q;
    RNNG_TRY(f->value.reserve(std::max<size_t>(1, (size_t)total) * sizeof(float)));
    RNNG_TRY(f->id.reserve(std::max<size_t>(1, (size_t)total) * sizeof(int64_t)));
    float* val = f->value.as<float>();
    int64_t* id = f->id.as<int64_t>();
    if (n > 0) k_place_vertices<<<blocks_for(n, 256), 256, 0, stream()>>>(val, id, n);
    if (E > 0) {
        LocalBuf gsE, gsT, gsQ, gsize, goff;
        RNNG_TRY(gsE.reserve(((size_t)G + 1) * sizeof(int64_t)));
        RNNG_TRY(gsize.reserve((size_t)G * sizeof(int64_t)));
        RNNG_TRY(goff.reserve((size_t)G * sizeof(int64_t)));
        k_group_starts<<<blocks_for(E + 1, 256), 256, 0, stream()>>>(vrS.as<uint32_t>(), E, G, gsE.as<int64_t>());
        if (T > 0) {
            RNNG_TRY(gsT.reserve(((size_t)G + 1) * sizeof(int64_t)));
            k_group_starts<<<blocks_for(T + 1, 256), 256, 0, stream()>>>(tg.as<uint32_t>(), T, G,
                                                                       gsT.as<int64_t>());
        }
        if (Q > 0) {
            RNNG_TRY(gsQ.reserve(((size_t)G + 1) * sizeof(int64_t)));
            k_group_starts<<<blocks_for(Q + 1, 256), 256, 0, stream()>>>(qg.as<uint32_t>(), Q, G,
                                                                       gsQ.as<int64_t>());
        }
        k_group_sizes<<<blocks_for(G, 256), 256, 0, stream()>>>(gsE.as<int64_t>(), T > 0 ? gsT.as<int64_t>() : nullptr,
                                                               Q > 0 ? gsQ.as<int64_t>() : nullptr, n, G,
                                                               gsize.as<int64_t>());
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(exclusive_scan_i64(gsize.as<int64_t>(), goff.as<int64_t>(), G));
        k_place_edges<<<blocks_for(E, 256), 256, 0, stream()>>>(vrS.as<uint32_t>(), perm.as<uint32_t>(), row.as<int>(),
                                                               col.as<int>(), E, gsE.as<int64_t>(),
                                                               goff.as<int64_t>(), gval.as<float>(), n, val, id);
        if (T > 0)
            k_place_cliques<<<blocks_for(T, 256), 256, 0, stream()>>>(
                tg.as<uint32_t>(), tid.as<int64_t>(), T, gsT.as<int64_t>(), gsE.as<int64_t>(), nullptr,
                goff.as<int64_t>(), gval.as<float>(), n, val, id);
        if (Q > 0)
            k_place_cliques<<<blocks_for(Q, 256), 256, 0, stream()>>>(
                qg.as<uint32_t>(), qid.as<int64_t>(), Q, gsQ.as<int64_t>(), gsE.as<int64_t>(),
                T > 0 ? gsT.as<int64_t>() : nullptr, goff.as<int64_t>(), gval.as<float>(), n, val, id);
        RNNG_HIP_TRY(hipGetLastError());
    }
    RNNG_HIP_TRY(hipEventRecord(ev.e[3], stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    f->n = total;
    f->counts[0] = n;
    f->counts[1] = Ecount;
    f->counts[2] = T;
    f->counts[3] = Q;
    if (st) {
        st->pair_evals += gst.pair_evals;
        st->candidates += gst.candidates;
        st->edges += gst.edges;
        st->band_pairs += gst.band_pairs;
        st->band_flips += gst.band_flips;
        st->refills += gst.refills;
        st->filter_ms += ev.ms(0, 1);     // epsilon graph
        st->recheck_ms += ev.ms(1, 2);    // CSR, value groups, cliques
        st->prep_ms += ev.ms(2, 3);       // (value, id) order
        if (T > 0) {
            st->kernel_ms += ev.ms(4, 5);  // the triangle fill
            st->kernel_units += T;
        }
        st->total_ms += ev.ms(0, 3);
    }
    return RNNG_OK;
}

}  // namespace
}  // namespace rnng

using namespace rnng;

extern "C" {

int rnng_rips_build(rnng_dataset_t ds, float eps, int32_t maxdim, rnng_filtration_t* out,
                    rnng_stats* st) {
    if (!ds || !out) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    *out = nullptr;
    RNNG_TRY(ensure_init());
    rnng_filtration_s* f = new rnng_filtration_s();
    const int rc = rips_build(ds, eps, maxdim, f, st);
    if (rc != RNNG_OK) {
        rnng_filtration_destroy(f);
        return rc;
    }
    *out = f;
    return RNNG_OK;
}

int rnng_filtration_size(rnng_filtration_t f, int64_t* n, int64_t* per_dim) {
    if (!f || !n) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    *n = f->n;
    if (per_dim)
        for (int p = 0; p < 4; ++p) per_dim[p] = f->counts[p];
    return RNNG_OK;
}

int rnng_filtration_copy(rnng_filtration_t f, float* value, int64_t* id, int64_t cap) {
    if (!f || cap < 0) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    const int64_t k = std::min(cap, f->n);
    if (k <= 0) return RNNG_OK;
    if (value)
        RNNG_HIP_TRY(hipMemcpyAsync(value, f->value.ptr, (size_t)k * sizeof(float), hipMemcpyDeviceToHost, stream()));
    if (id) RNNG_HIP_TRY(hipMemcpyAsync(id, f->id.ptr, (size_t)k * sizeof(int64_t), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

int rnng_filtration_relabel(rnng_filtration_t f, const int64_t* ids, int64_t n, const uint8_t* vertex_flags,
                            uint8_t* simplex_flags) {
    if (!f || n < 0 || (n > 0 && !ids)) {
        set_error("rnng_filtration_relabel: bad arguments");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    if (f->n == 0) return RNNG_OK;
    if (f->counts[0] != n) {
        set_error("rnng_filtration_relabel: n must be the number of points the filtration was built on");
        return RNNG_ERR_ARG;
    }
    std::vector<int64_t> sorted_ids(ids, ids + n);
    std::sort(sorted_ids.begin(), sorted_ids.end());
    if (n > 0 && (sorted_ids.front() < 0 || std::adjacent_find(sorted_ids.begin(), sorted_ids.end()) != sorted_ids.end())) {
        set_error("rnng_filtration_relabel: labels must be distinct and non-negative");
        return RNNG_ERR_ARG;
    }
    // the relabelled ids must be what the reference's int64 binomials give
    const int64_t top = n > 0 ? sorted_ids.back() + 1 : 0;
    if ((f->counts[3] > 0 && top > kMaxTetraVertices) || (f->counts[2] > 0 && top > 1800000) ||
        top > 0x7fffffffll) {
        set_error("rnng_filtration_relabel: labels too large for the simplex ids of this dimension");
        return RNNG_ERR_UNSUPPORTED;
    }
    LocalBuf dmap, dvf, dsf;
    RNNG_TRY(dmap.reserve((size_t)std::max<int64_t>(n, 1) * sizeof(int64_t)));
    if (n > 0) RNNG_HIP_TRY(hipMemcpyAsync(dmap.ptr, ids, (size_t)n * sizeof(int64_t), hipMemcpyHostToDevice, stream()));
    if (vertex_flags) {
        RNNG_TRY(dvf.reserve((size_t)std::max<int64_t>(n, 1)));
        if (n > 0) RNNG_HIP_TRY(hipMemcpyAsync(dvf.ptr, vertex_flags, (size_t)n, hipMemcpyHostToDevice, stream()));
    }
    RNNG_TRY(dsf.reserve((size_t)f->n));
    k_relabel<<<blocks_for(f->n, 256), 256, 0, stream()>>>(f->id.as<int64_t>(), f->n, dmap.as<int64_t>(), n,
                                                           vertex_flags ? dvf.as<uint8_t>() : nullptr,
                                                           dsf.as<uint8_t>());
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_TRY(sort_by_value_then_id(f->value.as<float>(), f->id.as<int64_t>(), f->n, dsf.as<uint8_t>()));
    if (simplex_flags)
        RNNG_HIP_TRY(hipMemcpyAsync(simplex_flags, dsf.ptr, (size_t)f->n, hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

int rnng_filtration_destroy(rnng_filtration_t f) {
    if (!f) return RNNG_OK;
    f->value.release();
    f->id.release();
    delete f;
    return RNNG_OK;
}

}  // extern "C"
