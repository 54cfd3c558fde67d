// Vietoris-Rips filtration (vertices, edges, triangles) of the L2 epsilon
// graph, for gfx950.
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
// Pipeline: epsilon graph -> device sort by (u, v) -> CSR of upper
// neighbours -> per-edge merge intersection of the two sorted neighbour lists
// (triangles) and, for dimension 3, a three-way merge with the third vertex's
// list (tetrahedra); count, exclusive scan, fill: deterministic, no atomics ->
// filtration arrays -> device sort by (value, id).
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

// Largest vertex count whose 4-clique ids the reference computes without
// overflow: its Binom multiplies v (v-1) (v-2) (v-3) in int64
// (include/utils.hpp:126-127), which stays below 2^63 for v <= 55110.
constexpr int64_t kMaxTetraVertices = 55111;

// rowptr[u] = first edge of row u; keys sorted by (u << 32 | v).
__global__ void k_csr_rows(const uint64_t* __restrict__ keys, int64_t E, int n,
                           int64_t* __restrict__ rowptr, int* __restrict__ col) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e > E) return;
    const int u = e < E ? (int)(keys[e] >> 32) : n;
    const int pu = e > 0 ? (int)(keys[e - 1] >> 32) : -1;
    for (int r = pu + 1; r <= u; ++r) rowptr[r] = e;
    if (e < E) col[e] = (int)(keys[e] & 0xffffffffu);
}

__global__ void k_edge_rows(const uint64_t* __restrict__ keys, int64_t E, int* __restrict__ row) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) row[e] = (int)(keys[e] >> 32);
}

// Triangles (u, v, w), u < v < w, through edge e = (u, v): the common upper
// neighbours w > v of u and v.  FILL = false counts, FILL = true writes.
template <bool FILL>
__global__ __launch_bounds__(256) void k_triangles(
    const int* __restrict__ row, const int64_t* __restrict__ rowptr, const int* __restrict__ col,
    const float* __restrict__ wt, int64_t E, int64_t* __restrict__ count,
    const int64_t* __restrict__ offset, float* __restrict__ tval, int64_t* __restrict__ tid) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int u = row[e], v = col[e];
    const float wuv = wt[e];
    int64_t i1 = e + 1, e1 = rowptr[u + 1];
    int64_t i2 = rowptr[v], e2 = rowptr[v + 1];
    int64_t c = 0, out = FILL ? offset[e] : 0;
    const int64_t base = (int64_t)u + binom2(v) + ((int64_t)2 << 60);
    while (i1 < e1 && i2 < e2) {
        const int a = col[i1], b = col[i2];
        if (a == b) {
            if (FILL) {
                tval[out] = fmaxf(wuv, fmaxf(wt[i1], wt[i2]));
                tid[out] = base + binom3(a);
                ++out;
            } else {
                ++c;
            }
            ++i1;
            ++i2;
        } else if (a < b) {
            ++i1;
        } else {
            ++i2;
        }
    }
    if (!FILL) count[e] = c;
}

// Tetrahedra (u, v, w, x), u < v < w < x, through edge e = (u, v): for each
// common upper neighbour w of u and v, the common upper neighbours x > w of
// u, v and w (a three-way merge of the sorted rows).  Value = the largest of
// the six edge weights (the max over facets, include/point.hpp:1194-1208).
template <bool FILL>
__global__ __launch_bounds__(256) void k_tetrahedra(
    const int* __restrict__ row, const int64_t* __restrict__ rowptr, const int* __restrict__ col,
    const float* __restrict__ wt, int64_t E, int64_t* __restrict__ count,
    const int64_t* __restrict__ offset, float* __restrict__ tval, int64_t* __restrict__ tid) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int u = row[e], v = col[e];
    const float wuv = wt[e];
    const int64_t e1 = rowptr[u + 1], e2 = rowptr[v + 1];
    int64_t i1 = e + 1, i2 = rowptr[v];
    int64_t c = 0, out = FILL ? offset[e] : 0;
    const int64_t base = (int64_t)u + binom2(v) + ((int64_t)3 << 60);
    while (i1 < e1 && i2 < e2) {
        const int a = col[i1], b = col[i2];
        if (a < b) {
            ++i1;
            continue;
        }
        if (b < a) {
            ++i2;
            continue;
        }
        const int w = a;
        const float m3 = fmaxf(wuv, fmaxf(wt[i1], wt[i2]));
        int64_t j1 = i1 + 1, j2 = i2 + 1, j3 = rowptr[w];
        const int64_t e3 = rowptr[w + 1];
        while (j1 < e1 && j2 < e2 && j3 < e3) {
            const int x1 = col[j1], x2 = col[j2], x3 = col[j3];
            const int mn = min(x1, min(x2, x3));
            if (x1 == x2 && x2 == x3) {
                if (FILL) {
                    tval[out] = fmaxf(m3, fmaxf(wt[j1], fmaxf(wt[j2], wt[j3])));
                    tid[out] = base + binom3(w) + binom4(x1);
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

__global__ void k_fill_vertices(float* __restrict__ val, int64_t* __restrict__ id, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        val[i] = 0.0f;
        id[i] = i;
    }
}

__global__ void k_fill_edges(const int* __restrict__ row, const int* __restrict__ col,
                             const float* __restrict__ wt, int64_t E, float* __restrict__ val,
                             int64_t* __restrict__ id) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) {
        val[e] = wt[e];
        id[e] = (int64_t)row[e] + binom2(col[e]) + ((int64_t)1 << 60);
    }
}

__global__ void k_pack_ij(const int2* __restrict__ ij, uint64_t* __restrict__ keys, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) keys[e] = ((uint64_t)(uint32_t)ij[e].x << 32) | (uint32_t)ij[e].y;
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
    if (maxdim >= 1 && n >= 2) {
        const int mode = ds->d >= 16 ? RNNG_MODE_MFMA : RNNG_MODE_PLAIN;
        RNNG_TRY(l2_graph_build(ds, ds, true, eps, mode, &g, st));
        E = g.n_edges;
    }
    RNNG_HIP_TRY(hipEventRecord(ev.e[1], stream()));
    LocalBuf keys, rowptr, col, row, cnt, off;
    RNNG_TRY(keys.reserve(std::max<size_t>(1, (size_t)E) * sizeof(uint64_t)));
    RNNG_TRY(rowptr.reserve(((size_t)n + 1) * sizeof(int64_t)));
    RNNG_TRY(col.reserve(std::max<size_t>(1, (size_t)E) * sizeof(int)));
    RNNG_TRY(row.reserve(std::max<size_t>(1, (size_t)E) * sizeof(int)));
    if (E > 0) {
        k_pack_ij<<<blocks_for(E, 256), 256, 0, stream()>>>(g.ij.as<int2>(), keys.as<uint64_t>(), E);
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(sort_pairs_u64_f32(keys.as<uint64_t>(), g.w.as<float>(), E));
    }
    k_csr_rows<<<blocks_for(E + 1, 256), 256, 0, stream()>>>(keys.as<uint64_t>(), E, n,
                                                               rowptr.as<int64_t>(), col.as<int>());
    k_edge_rows<<<blocks_for(std::max<int64_t>(E, 1), 256), 256, 0, stream()>>>(keys.as<uint64_t>(), E,
                                                                                row.as<int>());
    RNNG_HIP_TRY(hipGetLastError());
    int64_t T = 0;
    if (maxdim >= 2 && E > 0) {
        RNNG_TRY(cnt.reserve((size_t)E * sizeof(int64_t)));
        RNNG_TRY(off.reserve(((size_t)E + 1) * sizeof(int64_t)));
        k_triangles<false><<<blocks_for(E, 256), 256, 0, stream()>>>(
            row.as<int>(), rowptr.as<int64_t>(), col.as<int>(), g.w.as<float>(), E, cnt.as<int64_t>(),
            nullptr, nullptr, nullptr);
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(exclusive_scan_i64(cnt.as<int64_t>(), off.as<int64_t>(), E));
        int64_t last_off = 0, last_cnt = 0;
        RNNG_HIP_TRY(hipMemcpy(&last_off, off.as<int64_t>() + (E - 1), sizeof(int64_t), hipMemcpyDeviceToHost));
        RNNG_HIP_TRY(hipMemcpy(&last_cnt, cnt.as<int64_t>() + (E - 1), sizeof(int64_t), hipMemcpyDeviceToHost));
        T = last_off + last_cnt;
    }
    // clique counts by the per-edge count / scan / fill of kernel K
    auto count_cliques = [&](auto kernel, LocalBuf& cbuf, LocalBuf& obuf, int64_t& total_out) -> int {
        RNNG_TRY(cbuf.reserve((size_t)E * sizeof(int64_t)));
        RNNG_TRY(obuf.reserve(((size_t)E + 1) * sizeof(int64_t)));
        kernel<<<blocks_for(E, 256), 256, 0, stream()>>>(row.as<int>(), rowptr.as<int64_t>(), col.as<int>(),
                                                         g.w.as<float>(), E, cbuf.as<int64_t>(), nullptr, nullptr,
                                                         nullptr);
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(exclusive_scan_i64(cbuf.as<int64_t>(), obuf.as<int64_t>(), E));
        int64_t lo = 0, lc = 0;
        RNNG_HIP_TRY(hipMemcpy(&lo, obuf.as<int64_t>() + (E - 1), sizeof(int64_t), hipMemcpyDeviceToHost));
        RNNG_HIP_TRY(hipMemcpy(&lc, cbuf.as<int64_t>() + (E - 1), sizeof(int64_t), hipMemcpyDeviceToHost));
        total_out = lo + lc;
        return RNNG_OK;
    };
    int64_t Q = 0;
    LocalBuf cnt4, off4;
    if (maxdim >= 3 && E > 0) RNNG_TRY(count_cliques(k_tetrahedra<false>, cnt4, off4, Q));
    const int64_t total = (int64_t)n + (maxdim >= 1 ? E : 0) + T + Q;
    RNNG_TRY(f->value.reserve(std::max<size_t>(1, (size_t)total) * sizeof(float)));
    RNNG_TRY(f->id.reserve(std::max<size_t>(1, (size_t)total) * sizeof(int64_t)));
    float* val = f->value.as<float>();
    int64_t* id = f->id.as<int64_t>();
    if (n > 0) k_fill_vertices<<<blocks_for(n, 256), 256, 0, stream()>>>(val, id, n);
    if (E > 0)
        k_fill_edges<<<blocks_for(E, 256), 256, 0, stream()>>>(row.as<int>(), col.as<int>(), g.w.as<float>(),
                                                                E, val + n, id + n);
    if (T > 0)
        k_triangles<true><<<blocks_for(E, 256), 256, 0, stream()>>>(
            row.as<int>(), rowptr.as<int64_t>(), col.as<int>(), g.w.as<float>(), E, nullptr,
            off.as<int64_t>(), val + n + E, id + n + E);
    if (Q > 0)
        k_tetrahedra<true><<<blocks_for(E, 256), 256, 0, stream()>>>(
            row.as<int>(), rowptr.as<int64_t>(), col.as<int>(), g.w.as<float>(), E, nullptr,
            off4.as<int64_t>(), val + n + E + T, id + n + E + T);
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_HIP_TRY(hipEventRecord(ev.e[2], stream()));
    RNNG_TRY(sort_by_value_then_id(val, id, total));
    RNNG_HIP_TRY(hipEventRecord(ev.e[3], stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    f->n = total;
    f->counts[0] = n;
    f->counts[1] = maxdim >= 1 ? E : 0;
    f->counts[2] = T;
    f->counts[3] = Q;
    if (st) {
        st->recheck_ms += ev.ms(1, 2);    // CSR + clique stage
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
    if (value) RNNG_HIP_TRY(hipMemcpy(value, f->value.ptr, (size_t)k * sizeof(float), hipMemcpyDeviceToHost));
    if (id) RNNG_HIP_TRY(hipMemcpy(id, f->id.ptr, (size_t)k * sizeof(int64_t), hipMemcpyDeviceToHost));
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
