// Edit-distance epsilon graph for gfx950.
//
// Reference semantics (gabe-raulet/RNNG): EditDistance<char>::operator()
// drivers/systolic.cpp:271-299 -- Levenshtein distance with unit costs,
// returned as a float; a pair is an edge iff that float is <= radius
// (include/search.hpp:244,399).  Distances are integers, so with
// k = floor(eps) the test is dist <= k.
//
// Three stages, all exact:
//   1. Strings are ordered by length.  |m - n| <= dist, so for row string a
//      only the contiguous range of strings at most k longer can match.
//   2. k_edit_filter compares 64-bin byte profiles of hashed 3-grams with
//      v_sad_u8.  Each edit operation changes at most 2q = 6 q-gram counts,
//      and merging q-grams into bins (or clamping counts at 255) can only
//      shrink the L1 distance, so dist >= SAD / 6: pairs with SAD > 6k are
//      not edges.
//   3. k_edit_verify runs Ukkonen's banded dynamic programme (cells with
//      |i - j| <= KB, KB >= k).  Any alignment of cost <= k stays inside that
//      band, so the banded value equals the Levenshtein distance whenever the
//      distance is <= k and exceeds k otherwise.
#include "internal.hpp"

#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

struct rnng_strings_s {
    int64_t n = 0;
    int64_t total = 0;
    rnng::DevBuf atoms;     // uint8, original order
    rnng::DevBuf offsets;   // int64 [n+1], original order
    rnng::DevBuf order;     // int32 [n]: sorted position -> original id
    rnng::DevBuf slen;      // int32 [n]: length by sorted position
    rnng::DevBuf soff;      // int64 [n]: atom offset by sorted position
    rnng::DevBuf prof;      // uint4 [n][4]: 3-gram profile by sorted position
    std::vector<int32_t> len_sorted;
};

namespace rnng {
namespace {

constexpr int ET = 64;     // filter tile edge (pairs)
constexpr int ECH = 32;    // column tiles per filter work item

// 2-bit code of a character: exact for A, C, G, T (0, 1, 3, 2); for other
// alphabets it merges characters, which keeps the profile bound valid.
__device__ __forceinline__ unsigned code2(unsigned c) { return (c >> 1) & 3u; }

__global__ __launch_bounds__(256) void k_profiles(const uint8_t* __restrict__ atoms,
                                                  const int64_t* __restrict__ soff,
                                                  const int* __restrict__ slen, int n,
                                                  uint4* __restrict__ prof) {
    __shared__ __align__(16) uint8_t cnt[256][64];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    uint8_t* c = cnt[threadIdx.x];
    for (int b = 0; b < 64; ++b) c[b] = 0;
    if (a < n) {
        const uint8_t* s = atoms + soff[a];
        const int len = slen[a];
        unsigned code = 0;
        for (int i = 0; i < len; ++i) {
            code = ((code << 2) | code2(s[i])) & 63u;
            if (i >= 2 && c[code] < 255) c[code]++;
        }
        const uint4* src = (const uint4*)c;
#pragma unroll
        for (int q = 0; q < 4; ++q) prof[(size_t)a * 4 + q] = src[q];
    }
}

__global__ __launch_bounds__(256) void k_edit_filter(
    const uint4* __restrict__ prof, const int* __restrict__ slen, const int* __restrict__ order,
    const int4* __restrict__ work, int n, int k, unsigned maxsad, int2* __restrict__ cand,
    unsigned long long cap, unsigned long long* __restrict__ counter) {
    __shared__ uint4 rp[ET][4];
    __shared__ uint4 cp[ET][4];
    __shared__ int rl[ET], cl[ET];
    const int4 wi = work[blockIdx.x];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int a0 = wi.x * ET;
    {
        const int r = tid >> 2, q = tid & 3;
        const int a = a0 + r;
        rp[r][q] = a < n ? prof[(size_t)a * 4 + q] : make_uint4(0, 0, 0, 0);
        if (q == 0) rl[r] = a < n ? slen[a] : 0;
    }
    for (int bt = wi.y; bt < wi.z; ++bt) {
        const int b0 = bt * ET;
        __syncthreads();
        {
            const int r = tid >> 2, q = tid & 3;
            const int b = b0 + r;
            cp[r][q] = b < n ? prof[(size_t)b * 4 + q] : make_uint4(0, 0, 0, 0);
            if (q == 0) cl[r] = b < n ? slen[b] : 0;
        }
        __syncthreads();
        unsigned acc[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[r][c] = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint4 x[4], y[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) x[r] = rp[ty * 4 + r][q];
#pragma unroll
            for (int c = 0; c < 4; ++c) y[c] = cp[tx * 4 + c][q];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    unsigned s = acc[r][c];
                    s = __builtin_amdgcn_sad_u8(x[r].x, y[c].x, s);
                    s = __builtin_amdgcn_sad_u8(x[r].y, y[c].y, s);
                    s = __builtin_amdgcn_sad_u8(x[r].z, y[c].z, s);
                    s = __builtin_amdgcn_sad_u8(x[r].w, y[c].w, s);
                    acc[r][c] = s;
                }
        }
        unsigned cnt = 0;
        bool ok[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int a = a0 + ty * 4 + r, b = b0 + tx * 4 + c;
                ok[r][c] = a < b && b < n && (cl[tx * 4 + c] - rl[ty * 4 + r]) <= k &&
                           acc[r][c] <= maxsad;
                cnt += ok[r][c] ? 1u : 0u;
            }
        if (__any(cnt != 0)) {
            unsigned long long pos = wave_reserve(cnt, counter);
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    if (ok[r][c]) {
                        if (pos < cap) {
                            const int i = order[a0 + ty * 4 + r], j = order[b0 + tx * 4 + c];
                            cand[pos] = make_int2(min(i, j), max(i, j));
                        }
                        ++pos;
                    }
        }
    }
}

// Banded Levenshtein, band half-width KB >= k.  D[o + KB] holds the DP cell
// (i, i + o) of the current row i (rows over s, columns over t).
template <int KB>
__device__ int banded_edit(const uint8_t* __restrict__ s, int m, const uint8_t* __restrict__ t,
                           int n, int k) {
    constexpr int W = 2 * KB + 1;
    constexpr int INF = 1 << 20;
    if (m - n > KB || n - m > KB) return INF;
    int D[W + 1];
    unsigned T[W];   // T[o + KB] = t[i - 1 + o] for the current row i (256 = outside t)
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const int o = q - KB;
        D[q] = (o >= 0 && o <= n) ? o : INF;   // row 0: D(0, j) = j
        const int tj = o;                       // row 1 needs t[o], filled by the shift below
        T[q] = (tj - 1 >= 0 && tj - 1 < n) ? t[tj - 1] : 256u;
    }
    D[W] = INF;
    for (int i = 1; i <= m; ++i) {
        // slide the t window: T[q] <- t[i - 1 + (q - KB)]
#pragma unroll
        for (int q = 0; q < W - 1; ++q) T[q] = T[q + 1];
        {
            const int tj = i - 1 + KB;
            T[W - 1] = (tj >= 0 && tj < n) ? t[tj] : 256u;
        }
        const unsigned sc = s[i - 1];
        int left = INF;   // D(i, j - 1) for the first band cell
        int mn = INF;
#pragma unroll
        for (int q = 0; q < W; ++q) {
            const int o = q - KB;
            const int j = i + o;
            int v = min(min(D[q] + (sc != T[q] ? 1 : 0), D[q + 1] + 1), left + 1);
            v = (j == 0) ? i : v;
            v = (j < 0 || j > n) ? INF : v;
            D[q] = v;
            left = v;
            mn = min(mn, v);
        }
        if (mn > k) return INF;
    }
    const int oe = n - m;
    int r = INF;
#pragma unroll
    for (int q = 0; q < W; ++q)
        if (q - KB == oe) r = D[q];
    return r;
}

template <int KB>
__global__ __launch_bounds__(256) void k_edit_verify(
    const uint8_t* __restrict__ atoms, const int64_t* __restrict__ offsets,
    const int2* __restrict__ cand, unsigned long long ccap, int k, int2* __restrict__ eij,
    float* __restrict__ ew, unsigned long long ecap, unsigned long long* __restrict__ counters) {
    const unsigned long long nc = min(counters[0], ccap);
    const int lane = lane_id();
    for (unsigned long long base = (unsigned long long)blockIdx.x * blockDim.x; base < nc;
         base += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long c = base + threadIdx.x;
        bool edge = false;
        int2 p = make_int2(0, 0);
        int dist = 0;
        if (c < nc) {
            p = cand[c];
            const int64_t oi = offsets[p.x], oj = offsets[p.y];
            const int m = (int)(offsets[p.x + 1] - oi), nn = (int)(offsets[p.y + 1] - oj);
            dist = banded_edit<KB>(atoms + oi, m, atoms + oj, nn, k);
            edge = dist <= k;
        }
        const unsigned long long msk = __ballot(edge);
        if (msk) {
            unsigned lo = 0, hi = 0;
            if (lane == 0) {
                unsigned long long b0 = atomicAdd(&counters[1], (unsigned long long)__popcll(msk));
                lo = (unsigned)b0;
                hi = (unsigned)(b0 >> 32);
            }
            lo = __shfl(lo, 0, 64);
            hi = __shfl(hi, 0, 64);
            const unsigned long long b0 = ((unsigned long long)hi << 32) | lo;
            const unsigned rank = __builtin_amdgcn_mbcnt_hi((unsigned)(msk >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((unsigned)msk, 0u));
            if (edge && b0 + rank < ecap) {
                eij[b0 + rank] = p;
                ew[b0 + rank] = (float)dist;
            }
        }
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

struct EditWs {
    DevBuf counters, cand, work;
    unsigned long long cand_cap = 0;
    std::vector<int4> work_host;
};
std::mutex g_edit_mu;
EditWs g_ews;

template <int KB>
void launch_verify(const rnng_strings_s* S, EditWs& ws, int k, rnng_graph_s* g) {
    k_edit_verify<KB><<<2048, 256, 0, stream()>>>(
        S->atoms.as<uint8_t>(), S->offsets.as<int64_t>(), ws.cand.as<int2>(), ws.cand_cap, k,
        g->ij.as<int2>(), g->w.as<float>(), (unsigned long long)g->cap,
        ws.counters.as<unsigned long long>());
}

int dispatch_verify(const rnng_strings_s* S, EditWs& ws, int k, rnng_graph_s* g) {
    if (k <= 4) launch_verify<4>(S, ws, k, g);
    else if (k <= 8) launch_verify<8>(S, ws, k, g);
    else if (k <= 12) launch_verify<12>(S, ws, k, g);
    else if (k <= 16) launch_verify<16>(S, ws, k, g);
    else if (k <= 24) launch_verify<24>(S, ws, k, g);
    else if (k <= 32) launch_verify<32>(S, ws, k, g);
    else {
        set_error("edit-distance radius above 32 is not supported");
        return RNNG_ERR_UNSUPPORTED;
    }
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

int edit_build(rnng_strings_s* S, float eps, rnng_graph_s* g, rnng_stats* st) {
    EditWs& ws = g_ews;
    const int n = (int)S->n;
    if (st) st->pair_evals += (int64_t)n * (n - 1) / 2;
    if (!(eps >= 0.0f) || n < 2) {
        RNNG_TRY(edge_buffers(g, 1));
        g->n_edges = 0;
        return RNNG_OK;
    }
    const int k = eps >= 1e9f ? 1000000000 : (int)floorf(eps);
    if (k > 32) {
        set_error("edit-distance radius above 32 is not supported");
        return RNNG_ERR_UNSUPPORTED;
    }
    // work list: row tile x column-tile chunk over the length-compatible range
    ws.work_host.clear();
    const int ntiles = (n + ET - 1) / ET;
    for (int at = 0; at < ntiles; ++at) {
        const int a_last = std::min(n - 1, at * ET + ET - 1);
        const int maxlen = S->len_sorted[a_last] + k;
        const int b_end = (int)(std::upper_bound(S->len_sorted.begin(), S->len_sorted.end(), maxlen) -
                                S->len_sorted.begin());
        const int bt_end = (b_end + ET - 1) / ET;
        for (int bt = at; bt < bt_end; bt += ECH)
            ws.work_host.push_back(make_int4(at, bt, std::min(bt_end, bt + ECH), 0));
    }
    RNNG_TRY(ws.counters.reserve(8 * sizeof(unsigned long long)));
    RNNG_TRY(ws.work.reserve(std::max<size_t>(1, ws.work_host.size()) * sizeof(int4)));
    if (!ws.work_host.empty())
        RNNG_HIP_TRY(hipMemcpyAsync(ws.work.ptr, ws.work_host.data(), ws.work_host.size() * sizeof(int4),
                                    hipMemcpyHostToDevice, stream()));
    if (ws.cand_cap == 0) {
        ws.cand_cap = std::max<unsigned long long>(1ull << 20, 128ull * (unsigned long long)n);
        RNNG_TRY(ws.cand.reserve(ws.cand_cap * sizeof(int2)));
    }
    if (g->cap == 0)
        RNNG_TRY(edge_buffers(g, g_edge_hint > 0 ? g_edge_hint
                                                 : std::max<int64_t>(1 << 20, 64 * (int64_t)n)));
    Events ev;
    RNNG_TRY(ev.create());
    for (int attempt = 0; attempt < 3; ++attempt) {
        k_zero_u64<<<1, 64, 0, stream()>>>(ws.counters.as<unsigned long long>(), 8);
        RNNG_HIP_TRY(hipEventRecord(ev.e[0], stream()));
        if (!ws.work_host.empty())
            k_edit_filter<<<(unsigned)ws.work_host.size(), 256, 0, stream()>>>(
                S->prof.as<uint4>(), S->slen.as<int>(), S->order.as<int>(), ws.work.as<int4>(), n, k,
                (unsigned)(6 * k), ws.cand.as<int2>(), ws.cand_cap, ws.counters.as<unsigned long long>());
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_HIP_TRY(hipEventRecord(ev.e[1], stream()));
        RNNG_TRY(dispatch_verify(S, ws, k, g));
        RNNG_HIP_TRY(hipEventRecord(ev.e[2], stream()));
        unsigned long long cnt[8];
        RNNG_HIP_TRY(hipMemcpyAsync(cnt, ws.counters.ptr, sizeof(cnt), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        const bool cand_over = cnt[0] > ws.cand_cap;
        const bool edge_over = (int64_t)cnt[1] > g->cap;
        if (!cand_over && !edge_over) {
            g->n_edges = (int64_t)cnt[1];
            if (st) {
                st->candidates += (int64_t)cnt[0];
                st->edges += (int64_t)cnt[1];
                st->filter_ms += ev.ms(0, 1);
                st->recheck_ms += ev.ms(1, 2);
                st->total_ms += ev.ms(0, 2);
            }
            return RNNG_OK;
        }
        if (st) st->refills++;
        if (cand_over) {
            ws.cand_cap = cnt[0] + cnt[0] / 8 + 1024;
            RNNG_TRY(ws.cand.reserve(ws.cand_cap * sizeof(int2)));
        }
        if (edge_over) RNNG_TRY(edge_buffers(g, (int64_t)cnt[1] + (int64_t)cnt[1] / 8 + 1024));
    }
    set_error("edit graph buffer refill did not converge");
    return RNNG_ERR_HIP;
}

}  // namespace
}  // namespace rnng

using namespace rnng;

extern "C" {

int rnng_strings_create(const char* atoms, const int64_t* offsets, int64_t n, rnng_strings_t* out) {
    if (!out || n < 0 || (n > 0 && (!offsets || (!atoms && offsets[n] > offsets[0])))) {
        set_error("rnng_strings_create: bad arguments");
        return RNNG_ERR_ARG;
    }
    *out = nullptr;
    if (n > 0x7fffffff - 64) {
        set_error("string sets are limited to 2^31 - 64 strings");
        return RNNG_ERR_UNSUPPORTED;
    }
    for (int64_t i = 0; i < n; ++i)
        if (offsets[i + 1] < offsets[i]) {
            set_error("rnng_strings_create: offsets must be non-decreasing");
            return RNNG_ERR_ARG;
        }
    RNNG_TRY(ensure_init());
    rnng_strings_s* S = new rnng_strings_s();
    S->n = n;
    const int64_t base = n > 0 ? offsets[0] : 0;
    S->total = n > 0 ? offsets[n] - base : 0;
    std::vector<int64_t> off(n + 1);
    for (int64_t i = 0; i <= n; ++i) off[i] = n > 0 ? offsets[i] - base : 0;
    std::vector<int32_t> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
        return off[a + 1] - off[a] < off[b + 1] - off[b];
    });
    std::vector<int32_t> len(n);
    std::vector<int64_t> soff(n);
    for (int64_t a = 0; a < n; ++a) {
        len[a] = (int32_t)(off[ord[a] + 1] - off[ord[a]]);
        soff[a] = off[ord[a]];
    }
    S->len_sorted = len;
    int rc = RNNG_OK;
    auto up = [&](DevBuf& b, const void* src, size_t bytes) {
        if (rc != RNNG_OK) return;
        rc = b.reserve(std::max<size_t>(bytes, 16));
        if (rc == RNNG_OK && bytes) {
            hipError_t e = hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, stream());
            if (e == hipSuccess) e = hipStreamSynchronize(stream());
            if (e != hipSuccess) {
                set_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
                rc = RNNG_ERR_HIP;
            }
        }
    };
    up(S->atoms, n > 0 ? atoms + base : nullptr, (size_t)S->total);
    up(S->offsets, off.data(), off.size() * sizeof(int64_t));
    up(S->order, ord.data(), ord.size() * sizeof(int32_t));
    up(S->slen, len.data(), len.size() * sizeof(int32_t));
    up(S->soff, soff.data(), soff.size() * sizeof(int64_t));
    if (rc == RNNG_OK) rc = S->prof.reserve(std::max<size_t>(16, (size_t)n * 64));
    if (rc == RNNG_OK && n > 0) {
        k_profiles<<<blocks_for(n, 256), 256, 0, stream()>>>(S->atoms.as<uint8_t>(), S->soff.as<int64_t>(),
                                                              S->slen.as<int>(), (int)n, S->prof.as<uint4>());
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(stream());
        if (e != hipSuccess) {
            set_error(std::string("k_profiles: ") + hipGetErrorString(e));
            rc = RNNG_ERR_HIP;
        }
    }
    if (rc != RNNG_OK) {
        rnng_strings_destroy(S);
        return rc;
    }
    *out = S;
    return RNNG_OK;
}

int rnng_strings_destroy(rnng_strings_t S) {
    if (!S) return RNNG_OK;
    S->atoms.release();
    S->offsets.release();
    S->order.release();
    S->slen.release();
    S->soff.release();
    S->prof.release();
    delete S;
    return RNNG_OK;
}

int rnng_edit_graph_build(rnng_strings_t S, float eps, rnng_graph_t* out, rnng_stats* st) {
    if (!S || !out) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    *out = nullptr;
    RNNG_TRY(ensure_init());
    std::lock_guard<std::mutex> lk(g_edit_mu);
    rnng_graph_s* g = new rnng_graph_s();
    const int rc = edit_build(S, eps, g, st);
    if (rc != RNNG_OK) {
        rnng_graph_destroy(g);
        return rc;
    }
    *out = g;
    return RNNG_OK;
}

int rnng_edit_eps_pairs(const char* atoms, const int64_t* offsets, int64_t n, float eps, int64_t* ij,
                        float* w, int64_t cap, int64_t* needed, rnng_stats* st) {
    if (!needed || cap < 0) {
        set_error("bad output arguments");
        return RNNG_ERR_ARG;
    }
    rnng_strings_t S = nullptr;
    RNNG_TRY(rnng_strings_create(atoms, offsets, n, &S));
    rnng_graph_t g = nullptr;
    int rc = rnng_edit_graph_build(S, eps, &g, st);
    if (rc == RNNG_OK) {
        *needed = g->n_edges;
        if (g->n_edges > cap) {
            set_error("output capacity too small");
            rc = RNNG_ERR_CAPACITY;
        } else {
            rc = rnng_graph_copy(g, ij, w, cap, 1);
        }
    }
    rnng_graph_destroy(g);
    rnng_strings_destroy(S);
    return rc;
}

}  // extern "C"
