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
//   2. k_edit_filter bounds the L1 distance of 64-bin 3-gram count profiles
//      on int8 MFMA.  Each edit operation changes at most 2q = 6 q-gram
//      counts, and merging q-grams into bins or clamping counts at 4 can only
//      shrink the L1 distance, so dist >= L1_4 / 6: pairs with L1_4 > 6k are
//      not edges.  With counts clamped at 4 and written as 4-level
//      thermometer codes t(p) (level l set iff count > l),
//      L1_4(p, q) = |p|_4 + |q|_4 - 2 <t(p), t(q)>, a 256-long int8 dot
//      product per pair.
//   3. k_edit_verify_bp decides each candidate with a bit-parallel
//      Levenshtein recurrence (Myers 1999, in Hyyro's 2001 form for the edit
//      distance) restricted to Ukkonen's diagonal band |i - j| <= k: one
//      (2k+1)-bit window of vertical deltas per column that slides down one
//      row per column once j > k + 1.  The cell just above the window takes
//      the value of its left neighbour + 1 and the cell below the window in
//      the previous column that of its upper neighbour + 1; both are upper
//      bounds of the true values, so every computed cell is >= the true
//      distance, and cells on an alignment of cost <= k (all inside the band)
//      are exact: the result equals the Levenshtein distance whenever that is
//      <= k and exceeds k otherwise.  The match vector of a column is an AND
//      over the bit-planes of the string set's dense alphabet codes (2 planes
//      for DNA).  k = 32 (a 65-bit window) uses k_edit_verify, Ukkonen's
//      banded dynamic programme with the same band; k > 32 uses
//      k_edit_verify_myers, the same recurrence unbanded over all rows of the
//      shorter string (up to 8 64-bit words, 512 atoms, one thread per
//      candidate), and for longer strings k_edit_verify_long, the same
//      recurrence over any length with one wave per candidate: lane w holds
//      word w of a band of 64 words and works on column j - w at step j (the
//      word-to-word carries pass one lane per step), bands of 4,096 rows in
//      turn with the carries between bands kept per column in memory.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "internal.hpp"
#include "mfma.hpp"

struct rnng_strings_s {
    int64_t n = 0;
    int64_t total = 0;
    rnng::DevBuf atoms;     // uint8, original order
    rnng::DevBuf offsets;   // int64 [n+1], original order
    rnng::DevBuf order;     // int32 [n]: sorted position -> original id
    rnng::DevBuf slen;      // int32 [n]: length by sorted position
    rnng::DevBuf soff;      // int64 [n]: atom offset by sorted position
    rnng::DevBuf prof;      // uint4 [n][4]: 3-gram profile by sorted position
    rnng::DevBuf mxa, myb;  // int8 thermometer fragments (MFMA layout), rows and columns
    rnng::DevBuf mnrm;      // int [n_pad]: |p|_4 by sorted position (padding: huge)
    rnng::DevBuf codes;     // uint8 dense alphabet codes, strings in sorted order, 16-byte aligned
    rnng::DevBuf coff;      // int64 [n]: offset of each sorted position's codes
    int planes = 1;         // bits of a code
    int64_t codes_len = 0;  // bytes of codes (strings, alignment padding and tail)
    int64_t n_pad = 0;
    std::vector<int32_t> len_sorted;
    bool seen[256] = {false};   // bytes that occur in the set (codes are their ranks)
};

namespace rnng {
namespace {


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

constexpr int TLEV = 4;              // thermometer levels per 3-gram bin (counts clamped at 4)
constexpr int EKS = 64 * TLEV / 32;  // int8 k-steps of 32 per pair
constexpr int ERT = 2;               // row tiles (32 strings) per wave
constexpr int ERB = 4 * ERT * 32;    // rows per work item (4 waves)
constexpr int ECT = 64;              // column tiles per work item
constexpr int EPAD = 1 << 28;        // |p|_4 of padding strings: never a candidate
constexpr int ESTAGE = 4096;         // candidates staged in LDS per work item

// Fragment layout (as the L2 filter's, with int8 operands): tile t = 32
// strings, k-step s, lane l: frag[(t*EKS + s)*64 + l] = 16 int8 of string
// 32t + (l & 31), thermometer slots 32s + 16(l >> 5) + j (bin = slot / 4,
// level = slot % 4).  Rows carry 1 per set level, columns -2, so the MFMA
// adds -2 <t(p), t(q)>.  One thread per (string, step, half).
__global__ __launch_bounds__(256) void k_edit_frags(const uint4* __restrict__ prof, int n, int64_t n_pad,
                                                    i32x4_t* __restrict__ xa, i32x4_t* __restrict__ yb,
                                                    int* __restrict__ nrm) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_pad * EKS * 2) return;
    const int64_t a = g / (EKS * 2);
    const int s = (int)(g % (EKS * 2)) >> 1, h = (int)(g & 1);
    // bins 8s + 4h .. +3 are word 2(s & 1) + h of profile chunk s / 2
    unsigned c4 = 0;
    if (a < n) {
        const uint4 ch = prof[a * 4 + (s >> 1)];
        const int w = 2 * (s & 1) + h;
        c4 = w == 0 ? ch.x : w == 1 ? ch.y : w == 2 ? ch.z : ch.w;
    }
    i32x4_t ones, twos;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const unsigned c = (c4 >> (8 * j)) & 255u;
        unsigned lv = 0;
#pragma unroll
        for (int l = 0; l < TLEV; ++l) lv |= (c > (unsigned)l ? 1u : 0u) << (8 * l);
        ones[j] = (int)lv;
        twos[j] = (int)(lv * 0xFEu);   // 0x01 -> 0xFE (-2) per byte, no carries
    }
    const int64_t at = (a >> 5) * EKS + s;
    const int lane = h * 32 + (int)(a & 31);
    xa[at * 64 + lane] = ones;
    yb[at * 64 + lane] = twos;
    if (s == 0 && h == 0) {
        int sum = EPAD;
        if (a < n) {
            sum = 0;
            for (int q = 0; q < 4; ++q) {
                const uint4 ch = prof[a * 4 + q];
                const unsigned w4[4] = {ch.x, ch.y, ch.z, ch.w};
                for (int w = 0; w < 4; ++w)
                    for (int j = 0; j < 4; ++j) sum += (int)min((w4[w] >> (8 * j)) & 255u, (unsigned)TLEV);
            }
        }
        nrm[a] = sum;
    }
}

__global__ void k_debug_mfma_i8(const i32x4_t* __restrict__ a, const i32x4_t* __restrict__ b,
                                int* __restrict__ d) {
    const int lane = threadIdx.x;
    i32x16_t acc;
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0;
    acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[lane], b[lane], acc, 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 16; ++q) d[lane * 16 + q] = acc[q];
}

// One work item = rows [ERB wi.x, +ERB) (two row tiles per wave, fragments in
// registers) against column tiles [wi.y, wi.z), which the workgroup stages in
// LDS (double buffered, one barrier per tile).  The accumulator of pair
// (a, b) is |p_a|_4 + |p_b|_4 - 6k - 1 - 2 <t(p_a), t(p_b)>, negative iff
// L1_4 <= 6k.  The row part |p_a|_4 is the first MFMA's C operand (no
// per-value start adds), so the accumulator ends at |p_a|_4 - 2 <t, t> and
// the pair survives iff that is below -(|p_b|_4 - 6k - 1); a lane's 16 values
// share one column, so the hot loop tests the minimum of the lane's
// accumulators.  Tiles with a survivor take the candidate path, which
// keeps a < b < n (sorted positions) inside the length window and stages
// (a, b) in LDS.  It reads no global memory: on this filter most tiles hold a
// candidate, and a load per set bit would stall the whole workgroup at the
// next barrier, so row lengths sit in LDS and column lengths are fetched a
// tile ahead.
// Rows and columns may come from two string sets (a bipartite join: rows nr
// strings of one, columns nc of the other, each in its own length order); a
// self join passes the same set twice and keeps a < b only.
__global__ __launch_bounds__(256) void k_edit_filter(
    const i32x4_t* __restrict__ xa, const i32x4_t* __restrict__ yb, const int* __restrict__ rnrm,
    const int* __restrict__ cnrm, const int* __restrict__ rslen, const int* __restrict__ cslen,
    const int4* __restrict__ work, int nr, int nc, int self_join, int k, int bias,
    int2* __restrict__ cand, unsigned long long cap, unsigned long long* __restrict__ counter) {
    constexpr int NCH = EKS * 64;   // 16-byte fragments per column tile
    __shared__ i32x4_t bs[2][NCH];
    __shared__ int2 stage[ESTAGE];
    __shared__ int srl[4][64];
    int4 wi;
    if (!load_work_item(work, wi)) return;   // padding item: no memory is touched
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    WaveSink sk{stage + wave * (ESTAGE / 4), 0u, (unsigned)(ESTAGE / 4), cand, cap, counter};
    const int rt0 = wi.x * (4 * ERT) + wave * ERT;

    i32x4_t a[ERT][EKS];
#pragma unroll
    for (int r = 0; r < ERT; ++r)
#pragma unroll
        for (int s = 0; s < EKS; ++s) a[r][s] = xa[((size_t)(rt0 + r) * EKS + s) * 64 + lane];
    i32x16_t xr[ERT];
#pragma unroll
    for (int r = 0; r < ERT; ++r)
#pragma unroll
        for (int q = 0; q < 16; ++q) xr[r][q] = rnrm[(rt0 + r) * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3)];
    int yc = cnrm[wi.y * 32 + (lane & 31)] + bias;
    // the lengths of the wave's 64 rows sit in LDS (read by the candidate
    // path), each lane keeps the length of its column, fetched a tile ahead
    srl[wave][lane] = rt0 * 32 + lane < nr ? rslen[rt0 * 32 + lane] : 0;
    int clen = wi.y * 32 + (lane & 31) < nc ? cslen[wi.y * 32 + (lane & 31)] : 0;
    bs[0][tid] = yb[(size_t)wi.y * NCH + tid];
    bs[0][tid + 256] = yb[(size_t)wi.y * NCH + tid + 256];
    __syncthreads();

    for (int ct = wi.y; ct < wi.z; ++ct) {
        const int buf = (ct - wi.y) & 1;
        const bool has_next = ct + 1 < wi.z;
        i32x4_t p0 = {0, 0, 0, 0}, p1 = {0, 0, 0, 0};
        int ycn = 0, clen_next = 0;
        if (has_next) {
            p0 = yb[(size_t)(ct + 1) * NCH + tid];
            p1 = yb[(size_t)(ct + 1) * NCH + tid + 256];
            const int cn = (ct + 1) * 32 + (lane & 31);
            ycn = cnrm[cn] + bias;
            clen_next = cn < nc ? cslen[cn] : 0;
        }
        i32x16_t acc[ERT];
#pragma unroll
        for (int s = 0; s < EKS; ++s) {
            const i32x4_t bb = bs[buf][s * 64 + lane];
#pragma unroll
            for (int r = 0; r < ERT; ++r)
                acc[r] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[r][s], bb, s == 0 ? xr[r] : acc[r], 0, 0, 0);
        }
        const int nyc = -yc;
        int mn = acc[0][0];
#pragma unroll
        for (int r = 0; r < ERT; ++r)
#pragma unroll
            for (int q = (r == 0 ? 1 : 0); q < 16; ++q) mn = min(mn, acc[r][q]);

        if (__any(mn < nyc)) {
            // Candidate path: the lanes walk their own set sign bits together,
            // one per lane and round, each checking its row's length (from
            // LDS) against the column's; a round's survivors go to the wave's
            // slice of the stage under one ballot.  The mask's shifts depend
            // on an opaque zero, so the compiler cannot hoist them into every
            // tile of the hot loop.
            int opaque;
            asm volatile("s_mov_b32 %0, 0" : "=s"(opaque));
            unsigned mask = 0;
#pragma unroll
            for (int r = 0; r < ERT; ++r)
#pragma unroll
                for (int q = 0; q < 16; ++q) mask |= (acc[r][q] < nyc ? 1u : 0u) << (r * 16 + q + opaque);
            const int col = ct * 32 + (lane & 31);
            const int rowb = 4 * (lane >> 5);
            if (col >= nc) mask = 0;
#pragma unroll 1
            for (;;) {
                const bool has = mask != 0;
                if (!__any(has)) break;
                bool pass = false;
                int row = 0;
                if (has) {
                    const int rq = __builtin_ctz(mask);
                    mask &= mask - 1;
                    const int q = rq & 15;
                    const int rl_i = rowb + (rq >> 4) * 32 + 8 * (q >> 2) + (q & 3);   // row within the wave
                    row = rt0 * 32 + rl_i;
                    // a self join's lengths rise with the sorted position, so
                    // clen >= the row's length there; padding rows never pass
                    const int dl = clen - srl[wave][rl_i];
                    pass = self_join ? (row < col && dl <= k) : (row < nr && dl <= k && -dl <= k);
                }
                sk.put(pass, __ballot(pass), row, col, lane);
            }
        }
        if (has_next) {
            bs[buf ^ 1][tid] = p0;
            bs[buf ^ 1][tid + 256] = p1;
        }
        __syncthreads();
        yc = ycn;
        clen = clen_next;
    }
    sk.flush(lane);   // the wave's staged candidates
}

// One side (rows or columns) of an edit join, strings in length order:
// atoms/soff address the raw bytes, codes/coff the dense alphabet codes
// (16-byte aligned per string), order maps a sorted position to its id.
struct EditSide {
    const uint8_t* atoms;
    const int64_t* soff;
    const uint8_t* codes;
    const int64_t* coff;
    const int* slen;
    const int* order;
};

// The output edge of candidate (a, b): a self join's ids in ascending order,
// a bipartite join's as (row id, column id).
__device__ __forceinline__ int2 edge_ids(const EditSide& R, const EditSide& C, int2 q, int self_join) {
    const int i = R.order[q.x], j = C.order[q.y];
    return self_join ? make_int2(min(i, j), max(i, j)) : make_int2(i, j);
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

// Candidates are sorted positions (a, b) of the row and column sides.
template <int KB>
__global__ __launch_bounds__(256) void k_edit_verify(
    const EditSide R, const EditSide C, int self_join, const int2* __restrict__ cand, unsigned long long ccap,
    int k, int2* __restrict__ eij, float* __restrict__ ew, unsigned long long ecap,
    unsigned long long* __restrict__ counters) {
    const unsigned long long nc = min(counters[0], ccap);
    const int lane = lane_id();
    for (unsigned long long base = (unsigned long long)blockIdx.x * blockDim.x; base < nc;
         base += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long c = base + threadIdx.x;
        bool edge = false;
        int2 p = make_int2(0, 0);
        int dist = 0;
        if (c < nc) {
            const int2 q = cand[c];
            const int m = R.slen[q.x], nn = C.slen[q.y];
            if (m - nn <= k && nn - m <= k) {   // |m - n| <= dist
                const uint8_t* sr = R.atoms + R.soff[q.x];
                const uint8_t* sc = C.atoms + C.soff[q.y];
                // the shorter string is the pattern; one call, so lanes of either order run
                // the same code
                const bool sw = m > nn;
                dist = banded_edit<KB>(sw ? sc : sr, sw ? nn : m, sw ? sr : sc, sw ? m : nn, k);
                edge = dist <= k;
            }
            if (edge) p = edge_ids(R, C, q, self_join);
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

// Bit-planes of 32 consecutive codes (two aligned 16-byte loads): bit u of
// w[q] = bit q of code u.  Per dword, (x >> q) & 0x01010101 leaves one bit per
// byte and the multiply by 0x01020408 gathers the four into bits 24..27
// (distinct partial products, no carries).
template <int NP>
__device__ __forceinline__ void plane_words_of(const uint4 a, const uint4 b, unsigned (&w)[NP]) {
    const unsigned d[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        unsigned v = 0;
#pragma unroll
        for (int r = 0; r < 8; ++r) v |= (((d[r] >> q) & 0x01010101u) * 0x01020408u) >> 24 << (4 * r);
        w[q] = v;
    }
}

template <int NP>
__device__ __forceinline__ void plane_words(const uint8_t* __restrict__ c32, unsigned (&w)[NP]) {
    plane_words_of<NP>(*reinterpret_cast<const uint4*>(c32), *reinterpret_cast<const uint4*>(c32 + 16), w);
}

// Bit-parallel banded Levenshtein of pattern p (rows, length m) against
// text t (columns, length n): D[m][n] if <= k, else a value > k.  VT holds
// the 2k+1-bit window (uint32_t for k <= 15, uint64_t for k <= 31); NP is the
// number of code bit-planes.  Both strings start 16-byte aligned and may be
// read up to 128 bytes past their end (codes there only reach rows below m
// or columns after n, which no kept cell depends on).  Columns run in groups
// of 32: the text's plane words give each column's code bits, and the
// pattern rows entering the window (row j + k for column j) come from one
// funnel shift by k of two consecutive pattern plane words.
template <class VT, int NP>
__device__ int banded_edit_bp(const uint8_t* __restrict__ p, int m, const uint8_t* __restrict__ t, int n,
                              int k) {
    if (m - n > k || n - m > k) return k + 1;
    const int w = 2 * k + 1;
    const VT W = w >= (int)(8 * sizeof(VT)) ? ~(VT)0 : (((VT)1 << w) - 1);
    unsigned pc[NP], pn[NP];
    plane_words<NP>(p, pc);        // pattern codes 0..31
    plane_words<NP>(p + 32, pn);   // 32..63
    VT pl[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) pl[q] = (sizeof(VT) == 8 ? ((VT)pc[q] | ((VT)pn[q] << 32)) : (VT)pc[q]) & W;
    VT Pv = W, Mv = 0;
    int T = 0;       // value of the cell just above the window
    int R = 1;       // first row of the window
    // the text's next 32 codes are read a group ahead (the loads land under
    // the current group's columns)
    uint4 ta = *reinterpret_cast<const uint4*>(t), tb = *reinterpret_cast<const uint4*>(t + 16);
    for (int g = 0; 32 * g < n; ++g) {
        unsigned tw[NP], ew[NP];
        plane_words_of<NP>(ta, tb, tw);
        if (32 * (g + 1) < n) {
            ta = *reinterpret_cast<const uint4*>(t + 32 * (g + 1));
            tb = *reinterpret_cast<const uint4*>(t + 32 * (g + 1) + 16);
        }
#pragma unroll
        for (int q = 0; q < NP; ++q) ew[q] = __builtin_amdgcn_alignbit(pn[q], pc[q], (unsigned)k);
        const int cols = min(32, n - 32 * g);
        for (int u = 0; u < cols; ++u) {
            const int j = 32 * g + u + 1;
            if (j > k + 1) {
                T += 1 + (int)(Pv & 1) - (int)(Mv & 1);
                Pv = (Pv >> 1) | ((VT)1 << (w - 1));
                Mv >>= 1;
                ++R;
#pragma unroll
                for (int q = 0; q < NP; ++q) pl[q] = (pl[q] >> 1) | ((VT)((ew[q] >> u) & 1u) << (w - 1));
            } else {
                T += 1;
            }
            VT Eq = W;
#pragma unroll
            for (int q = 0; q < NP; ++q) Eq &= pl[q] ^ ((VT)((tw[q] >> u) & 1u) - (VT)1);
            const VT Xv = Eq | Mv;
            const VT Xh = ((((Eq & Pv) + Pv) & W) ^ Pv) | Eq;
            VT Ph = (Mv | ~(Xh | Pv)) & W;
            VT Mh = Pv & Xh;
            Ph = ((Ph << 1) | 1) & W;
            Mh = (Mh << 1) & W;
            Pv = (Mh | ~(Xv | Ph)) & W;
            Mv = Ph & Xv;
        }
#pragma unroll
        for (int q = 0; q < NP; ++q) pc[q] = pn[q];
        plane_words<NP>(p + 32 * (g + 2), pn);
    }
    const int nb = m - R + 1;   // rows R..m of the final window
    const VT mask = nb <= 0 ? (VT)0 : nb >= (int)(8 * sizeof(VT)) ? ~(VT)0 : (((VT)1 << nb) - 1);
    int pcnt, mcnt;
    if (sizeof(VT) == 8) {
        pcnt = __popcll((unsigned long long)(Pv & mask));
        mcnt = __popcll((unsigned long long)(Mv & mask));
    } else {
        pcnt = __popc((unsigned)(Pv & mask));
        mcnt = __popc((unsigned)(Mv & mask));
    }
    return T + pcnt - mcnt;
}

template <class VT, int NP>
__global__ __launch_bounds__(256) void k_edit_verify_bp(
    const EditSide R, const EditSide C, int self_join, const int2* __restrict__ cand, unsigned long long ccap,
    int k, int2* __restrict__ eij, float* __restrict__ ew, unsigned long long ecap,
    unsigned long long* __restrict__ counters) {
    // each wave's edges gather in its own LDS buffer and take output slots
    // with one atomic per EBUF edges (the counters share one cache line)
    constexpr int EBUF = 256;
    __shared__ int2 bij_all[4][EBUF];
    __shared__ float bw_all[4][EBUF];
    const unsigned long long nc = min(counters[0], ccap);
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    int2* bij = bij_all[wave];
    float* bw = bw_all[wave];
    int en = 0;                        // edges in the buffer (the same in every lane)
    unsigned long long cols_run = 0;   // text columns this lane's candidates ran
    auto flush = [&]() {
        unsigned lo = 0, hi = 0;
        if (lane == 0 && en) {
            const unsigned long long b0 = atomicAdd(&counters[1], (unsigned long long)en);
            lo = (unsigned)b0;
            hi = (unsigned)(b0 >> 32);
        }
        lo = __shfl(lo, 0, 64);
        hi = __shfl(hi, 0, 64);
        const unsigned long long b0 = ((unsigned long long)hi << 32) | lo;
        for (int i = lane; i < en; i += 64)
            if (b0 + i < ecap) {
                eij[b0 + i] = bij[i];
                ew[b0 + i] = bw[i];
            }
        en = 0;
        __builtin_amdgcn_wave_barrier();
    };
    for (unsigned long long base = (unsigned long long)blockIdx.x * blockDim.x; base < nc;
         base += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long c = base + threadIdx.x;
        bool edge = false;
        int2 pr = make_int2(0, 0);
        int dist = 0;
        if (c < nc) {
            const int2 q = cand[c];
            const int m = R.slen[q.x], nn = C.slen[q.y];
            // the shorter string is the pattern; one call, so lanes of either order run the
            // same code
            const uint8_t* pr_ = R.codes + R.coff[q.x];
            const uint8_t* pc_ = C.codes + C.coff[q.y];
            const bool sw = m > nn;
            dist = banded_edit_bp<VT, NP>(sw ? pc_ : pr_, sw ? nn : m, sw ? pr_ : pc_, sw ? m : nn, k);
            edge = dist <= k;
            if (edge) pr = edge_ids(R, C, q, self_join);
            if (m - nn <= k && nn - m <= k) cols_run += (unsigned long long)max(m, nn);
        }
        const unsigned long long msk = __ballot(edge);
        if (msk) {
            const int ne = __popcll(msk);
            if (en + ne > EBUF) flush();
            const int o = en + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(msk >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((unsigned)msk, 0u));
            if (edge) {
                bij[o] = pr;
                bw[o] = (float)dist;
            }
            en += ne;
            __builtin_amdgcn_wave_barrier();
        }
    }
    flush();
    // text columns the band ran (counters[2]): one atomic per wave
    for (int o = 32; o > 0; o >>= 1) cols_run += __shfl_xor(cols_run, o, 64);
    if (lane == 0 && cols_run) atomicAdd(&counters[2], cols_run);
}

// Exact Levenshtein distance of pattern p (rows, length m <= 64 W) and text t
// (columns, length n), for radii beyond the banded kernels: Myers' bit-parallel
// column step over all m rows, W 64-bit words of vertical deltas (Pv = +1,
// Mv = -1) with the carries of the addition and of the shifts passed word to
// word; D[0][j] = j enters as the +1 shifted into row 1's horizontal delta, and
// the score follows row m (D[m][0] = m).  Rows past m (pattern codes read past
// its end) only feed rows below them, so they never reach row m.  Returns
// D[m][n], or a value > k as soon as D[m][j] - (n - j) > k after a group of
// 32 columns (D[m][n] >= D[m][j] - (n - j)).  Same read-past-end allowance as
// banded_edit_bp (at most 31 bytes past either string).
template <int W, int NP>
__device__ int myers_edit(const uint8_t* __restrict__ p, int m, const uint8_t* __restrict__ t, int n, int k) {
    if (m == 0) return n;
    uint64_t pl[NP][W];
#pragma unroll
    for (int w = 0; w < W; ++w)
#pragma unroll
        for (int q = 0; q < NP; ++q) pl[q][w] = 0;
#pragma unroll
    for (int u = 0; u < 2 * W; ++u) {
        if (32 * u < m) {
            unsigned pw[NP];
            plane_words<NP>(p + 32 * u, pw);
#pragma unroll
            for (int q = 0; q < NP; ++q) pl[q][u >> 1] |= (uint64_t)pw[q] << (32 * (u & 1));
        }
    }
    uint64_t Pv[W], Mv[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        Pv[w] = ~0ull;
        Mv[w] = 0ull;
    }
    const int wl = (m - 1) >> 6;   // word and bit of row m
    const int bl = (m - 1) & 63;
    int score = m;
    for (int g = 0; 32 * g < n; ++g) {
        unsigned tw[NP];
        plane_words<NP>(t + 32 * g, tw);
        const int cols = min(32, n - 32 * g);
        for (int u = 0; u < cols; ++u) {
            uint64_t cy = 0, ph_in = 1, mh_in = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                uint64_t Eq = ~0ull;
#pragma unroll
                for (int q = 0; q < NP; ++q) Eq &= pl[q][w] ^ ((uint64_t)((tw[q] >> u) & 1u) - 1ull);
                const uint64_t Xv = Eq | Mv[w];
                const uint64_t a = Eq & Pv[w];
                const uint64_t s1 = a + Pv[w];
                const uint64_t s2 = s1 + cy;
                cy = (uint64_t)((s1 < a) | (s2 < s1));
                const uint64_t Xh = (s2 ^ Pv[w]) | Eq;
                const uint64_t Ph = Mv[w] | ~(Xh | Pv[w]);
                const uint64_t Mh = Pv[w] & Xh;
                if (w == wl) score += (int)((Ph >> bl) & 1ull) - (int)((Mh >> bl) & 1ull);
                const uint64_t Phs = (Ph << 1) | ph_in;
                const uint64_t Mhs = (Mh << 1) | mh_in;
                ph_in = Ph >> 63;
                mh_in = Mh >> 63;
                Pv[w] = Mhs | ~(Xv | Phs);
                Mv[w] = Phs & Xv;
            }
        }
        if (score - (n - 32 * g - cols) > k) return k + 1;
    }
    return score;
}

template <int W, int NP>
__global__ __launch_bounds__(256) void k_edit_verify_myers(
    const EditSide R, const EditSide C, int self_join, const int2* __restrict__ cand, unsigned long long ccap,
    int k, int2* __restrict__ eij, float* __restrict__ ew, unsigned long long ecap,
    unsigned long long* __restrict__ counters) {
    const unsigned long long nc = min(counters[0], ccap);
    const int lane = lane_id();
    for (unsigned long long base = (unsigned long long)blockIdx.x * blockDim.x; base < nc;
         base += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long c = base + threadIdx.x;
        bool edge = false;
        int2 pr = make_int2(0, 0);
        int dist = 0;
        if (c < nc) {
            const int2 q = cand[c];
            const int m = R.slen[q.x], nn = C.slen[q.y];
            // (a shorter string of more than 64 W atoms: k_edit_verify_long)
            if (m - nn <= k && nn - m <= k && min(m, nn) <= 64 * W) {
                // the shorter string is the pattern (the distance is symmetric); one call,
                // so lanes of either order run the same code
                const uint8_t* pr_ = R.codes + R.coff[q.x];
                const uint8_t* pc_ = C.codes + C.coff[q.y];
                const bool sw = m > nn;
                dist = myers_edit<W, NP>(sw ? pc_ : pr_, sw ? nn : m, sw ? pr_ : pc_, sw ? m : nn, k);
                edge = dist <= k;
            }
            if (edge) pr = edge_ids(R, C, q, self_join);
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
                eij[b0 + rank] = pr;
                ew[b0 + rank] = (float)dist;
            }
        }
    }
}

// The multiword recurrence of myers_edit for a pattern of any length, as a
// wavefront over one wave: in band b (pattern words 64 b .. 64 b + 63, rows
// 4096 b + 1 ..), lane w holds word 64 b + w and at step s works on text
// column s - w, taking the addition carry and the two horizontal deltas of
// the word below from lane w - 1's previous step.  Lane 0 takes them from the
// DP's top row (band 0: D[0][j] = j, a +1 horizontal delta) or from the band
// below, whose last word left them per column in scr (three bit-planes of
// (n + 31) / 32 words, written and read in column order with device-coherent
// accesses and a fence between bands).  The score follows row m (on the lane
// holding word (m - 1) / 64 of the last band): D[m][n], or a value > k as soon
// as D[m][j] - (n - j) > k.  Every argument is wave-uniform; pattern bytes are
// read at most 31 past the pattern's end, as in myers_edit.
template <int NP>
__device__ int myers_wave(const uint8_t* __restrict__ p, int m, const uint8_t* __restrict__ t, int n, int k,
                          int lane, unsigned* __restrict__ scr) {
    if (m == 0) return n;
    const int W = (m + 63) >> 6;           // pattern words
    const int nb = (W + 63) >> 6;          // bands of 64 words
    const int wl = (m - 1) >> 6, bl = (m - 1) & 63;
    const int nw32 = (n + 31) >> 5;
    int score = m;                         // D[m][0]; kept by row m's lane
    for (int b = 0; b < nb; ++b) {
        const int w = 64 * b + lane;
        const bool active = w < W;
        const int last = min(63, W - 1 - 64 * b);   // last active lane of the band
        const bool top = b == nb - 1;
        const int wlane = wl - 64 * b;     // row m's lane (top band)
        uint64_t pl[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) pl[q] = 0;
        if (active) {
            unsigned lo[NP], hi[NP];
            plane_words<NP>(p + 64 * w, lo);
#pragma unroll
            for (int q = 0; q < NP; ++q) hi[q] = 0u;
            if (64 * w + 32 < m) plane_words<NP>(p + 64 * w + 32, hi);
#pragma unroll
            for (int q = 0; q < NP; ++q) pl[q] = (uint64_t)lo[q] | ((uint64_t)hi[q] << 32);
        }
        uint64_t Pv = ~0ull, Mv = 0ull;
        unsigned cin = 0u;                 // carries from lane - 1 (bit 0 addition, 1 +1 delta, 2 -1 delta)
        unsigned acc0 = 0u, acc1 = 0u, acc2 = 0u;   // last lane: carries for the band above
        unsigned bin0 = 0u, bin1 = 0u, bin2 = 0u;   // lane 0: carries from the band below
        const int steps = n + last;
        for (int s = 0; s < steps; ++s) {
            const int col = s - lane;
            const bool live = active && col >= 0 && col < n;
            unsigned in = cin;
            if (lane == 0) {
                if (b == 0) {
                    in = 2u;
                } else if (col < n) {
                    const int sh = col & 31;
                    if (sh == 0) {
                        const unsigned* src = scr + (col >> 5);
                        bin0 = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        bin1 = __hip_atomic_load(src + nw32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        bin2 = __hip_atomic_load(src + 2 * nw32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    in = ((bin0 >> sh) & 1u) | (((bin1 >> sh) & 1u) << 1) | (((bin2 >> sh) & 1u) << 2);
                }
            }
            unsigned out = 0u;
            if (live) {
                const unsigned tc = t[col];
                uint64_t Eq = ~0ull;
#pragma unroll
                for (int q = 0; q < NP; ++q) Eq &= pl[q] ^ ((uint64_t)((tc >> q) & 1u) - 1ull);
                const uint64_t Xv = Eq | Mv;
                const uint64_t a = Eq & Pv;
                const uint64_t s1 = a + Pv;
                const uint64_t s2 = s1 + (uint64_t)(in & 1u);
                const unsigned cyo = (unsigned)((s1 < a) | (s2 < s1));
                const uint64_t Xh = (s2 ^ Pv) | Eq;
                const uint64_t Ph = Mv | ~(Xh | Pv);
                const uint64_t Mh = Pv & Xh;
                if (top && lane == wlane) score += (int)((Ph >> bl) & 1ull) - (int)((Mh >> bl) & 1ull);
                const uint64_t Phs = (Ph << 1) | (uint64_t)((in >> 1) & 1u);
                const uint64_t Mhs = (Mh << 1) | (uint64_t)((in >> 2) & 1u);
                Pv = Mhs | ~(Xv | Phs);
                Mv = Phs & Xv;
                out = cyo | ((unsigned)(Ph >> 63) << 1) | ((unsigned)(Mh >> 63) << 2);
                if (!top && lane == 63) {
                    const int sh = col & 31;
                    acc0 |= (out & 1u) << sh;
                    acc1 |= ((out >> 1) & 1u) << sh;
                    acc2 |= ((out >> 2) & 1u) << sh;
                    if (sh == 31 || col == n - 1) {
                        unsigned* dst = scr + (col >> 5);
                        __hip_atomic_store(dst, acc0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(dst + nw32, acc1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(dst + 2 * nw32, acc2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        acc0 = acc1 = acc2 = 0u;
                    }
                }
            }
            cin = __shfl_up(out, 1, 64);
            if (top && (s & 31) == 31) {
                // row m's lane has finished column s - wlane
                const int sc = __shfl(score, wlane, 64);
                const int cw = s - wlane;
                if (cw >= 0 && cw < n && sc - (n - 1 - cw) > k) return k + 1;
            }
        }
        if (!top) __threadfence();
    }
    return __shfl(score, wl - 64 * (nb - 1), 64);
}

// Candidates whose shorter string has more than short_max atoms (the others
// are k_edit_verify_myers's), one wave each.  scr: 3 * scr_words words per
// wave (only patterns of more than 4,096 atoms use it).
template <int NP>
__global__ __launch_bounds__(256) void k_edit_verify_long(
    const EditSide R, const EditSide C, int self_join, const int2* __restrict__ cand, unsigned long long ccap,
    int k, int short_max, int2* __restrict__ eij, float* __restrict__ ew, unsigned long long ecap,
    unsigned long long* __restrict__ counters, unsigned* __restrict__ scratch, int64_t scr_words) {
    const unsigned long long nc = min(counters[0], ccap);
    const int lane = lane_id();
    const unsigned wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const unsigned nwaves = gridDim.x * (blockDim.x >> 6);
    unsigned* scr = scratch ? scratch + (size_t)wave * 3 * (size_t)scr_words : nullptr;
    for (unsigned long long c = wave; c < nc; c += nwaves) {
        const int2 q = cand[c];
        const int m0 = R.slen[q.x], n0 = C.slen[q.y];
        if (m0 - n0 > k || n0 - m0 > k || min(m0, n0) <= short_max) continue;
        // the shorter string is the pattern
        const bool sw = m0 > n0;
        const uint8_t* pr_ = R.codes + R.coff[q.x];
        const uint8_t* pc_ = C.codes + C.coff[q.y];
        const int dist = myers_wave<NP>(sw ? pc_ : pr_, sw ? n0 : m0, sw ? pr_ : pc_, sw ? m0 : n0, k, lane, scr);
        if (lane == 0 && dist <= k) {
            const unsigned long long e = atomicAdd(&counters[1], 1ull);
            if (e < ecap) {
                eij[e] = edge_ids(R, C, q, self_join);
                ew[e] = (float)dist;
            }
        }
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

struct EditWs {
    DevBuf counters, cand, work, scratch;
    unsigned long long cand_cap = 0;
    std::vector<int4> work_host;
};
std::mutex g_edit_mu;
EditWs g_ews;

// The two sides of a join and what the verify kernels are chosen by.
struct EditJoin {
    EditSide R, C;
    int self_join = 1;
    int planes = 1;
    int maxlen = 0;   // longest string of either side
};

template <int KB>
void launch_verify(const EditJoin& J, EditWs& ws, int k, rnng_graph_s* g) {
    k_edit_verify<KB><<<2048, 256, 0, stream()>>>(J.R, J.C, J.self_join, ws.cand.as<int2>(), ws.cand_cap, k,
                                                  g->ij.as<int2>(), g->w.as<float>(), (unsigned long long)g->cap,
                                                  ws.counters.as<unsigned long long>());
}

template <class VT, int NP>
void launch_verify_bp(const EditJoin& J, EditWs& ws, int k, rnng_graph_s* g) {
    k_edit_verify_bp<VT, NP><<<2048, 256, 0, stream()>>>(J.R, J.C, J.self_join, ws.cand.as<int2>(), ws.cand_cap, k,
                                                         g->ij.as<int2>(), g->w.as<float>(),
                                                         (unsigned long long)g->cap,
                                                         ws.counters.as<unsigned long long>());
}

template <class VT>
void launch_verify_planes(const EditJoin& J, EditWs& ws, int k, rnng_graph_s* g) {
    if (J.planes <= 2) launch_verify_bp<VT, 2>(J, ws, k, g);
    else if (J.planes <= 3) launch_verify_bp<VT, 3>(J, ws, k, g);
    else if (J.planes <= 4) launch_verify_bp<VT, 4>(J, ws, k, g);
    else launch_verify_bp<VT, 8>(J, ws, k, g);
}

template <int W, int NP>
void launch_verify_myers(const EditJoin& J, EditWs& ws, int k, rnng_graph_s* g) {
    k_edit_verify_myers<W, NP><<<2048, 256, 0, stream()>>>(J.R, J.C, J.self_join, ws.cand.as<int2>(), ws.cand_cap,
                                                           k, g->ij.as<int2>(), g->w.as<float>(),
                                                           (unsigned long long)g->cap,
                                                           ws.counters.as<unsigned long long>());
}

template <int W>
void launch_myers_planes(const EditJoin& J, EditWs& ws, int k, rnng_graph_s* g) {
    if (J.planes <= 2) launch_verify_myers<W, 2>(J, ws, k, g);
    else if (J.planes <= 3) launch_verify_myers<W, 3>(J, ws, k, g);
    else if (J.planes <= 4) launch_verify_myers<W, 4>(J, ws, k, g);
    else launch_verify_myers<W, 8>(J, ws, k, g);
}

constexpr int kMyersMaxLen = 512;   // longest pattern of the thread-per-candidate kernel (8 words)
constexpr int kLongWaves = 1024;     // waves of k_edit_verify_long

template <int NP>
int launch_verify_long(const EditJoin& J, EditWs& ws, int k, rnng_graph_s* g) {
    // carries between bands of 4,096 rows: three bit-planes of a text's
    // columns per wave, for patterns longer than one band
    int64_t words = 0;
    unsigned* scr = nullptr;
    if (J.maxlen > 4096) {
        words = ((int64_t)J.maxlen + 31) / 32;
        RNNG_TRY(ws.scratch.reserve((size_t)kLongWaves * 3 * (size_t)words * sizeof(unsigned)));
        scr = ws.scratch.as<unsigned>();
    }
    k_edit_verify_long<NP><<<kLongWaves / 4, 256, 0, stream()>>>(
        J.R, J.C, J.self_join, ws.cand.as<int2>(), ws.cand_cap, k, kMyersMaxLen, g->ij.as<int2>(), g->w.as<float>(),
        (unsigned long long)g->cap, ws.counters.as<unsigned long long>(), scr, words);
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

int dispatch_verify(const EditJoin& J, EditWs& ws, int k, rnng_graph_s* g) {
    if (k > 32) {
        if (J.maxlen <= 64) launch_myers_planes<1>(J, ws, k, g);
        else if (J.maxlen <= 128) launch_myers_planes<2>(J, ws, k, g);
        else if (J.maxlen <= 256) launch_myers_planes<4>(J, ws, k, g);
        else launch_myers_planes<8>(J, ws, k, g);
        RNNG_HIP_TRY(hipGetLastError());
        if (J.maxlen > kMyersMaxLen) {   // pairs whose shorter string is longer than 512 atoms
            if (J.planes <= 2) return launch_verify_long<2>(J, ws, k, g);
            if (J.planes <= 3) return launch_verify_long<3>(J, ws, k, g);
            if (J.planes <= 4) return launch_verify_long<4>(J, ws, k, g);
            return launch_verify_long<8>(J, ws, k, g);
        }
        return RNNG_OK;
    }
    if (k <= 15) {
        launch_verify_planes<uint32_t>(J, ws, k, g);
        RNNG_HIP_TRY(hipGetLastError());
        return RNNG_OK;
    }
    if (k <= 31) {
        launch_verify_planes<uint64_t>(J, ws, k, g);
        RNNG_HIP_TRY(hipGetLastError());
        return RNNG_OK;
    }
    launch_verify<32>(J, ws, k, g);
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

__global__ void k_recode(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n,
                         const uint8_t* __restrict__ lut) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = lut[src[i]];
}

EditSide side_of(const rnng_strings_s* S, const uint8_t* codes) {
    EditSide e;
    e.atoms = S->atoms.as<uint8_t>();
    e.soff = S->soff.as<int64_t>();
    e.codes = codes;
    e.coff = S->coff.as<int64_t>();
    e.slen = S->slen.as<int>();
    e.order = S->order.as<int>();
    return e;
}

// Pairs (a in A, b in B) with Levenshtein distance <= eps; a self join (A is
// B) keeps a < b.  Strings of both sets are in length order, so each row
// block's partners form one contiguous column range.
int edit_join(const rnng_strings_s* A, const rnng_strings_s* B, bool self, float eps, rnng_graph_s* g,
              rnng_stats* st) {
    EditWs& ws = g_ews;
    const int nr = (int)A->n, nc = (int)B->n;
    if (st) st->pair_evals += self ? (int64_t)nr * (nr - 1) / 2 : (int64_t)nr * nc;
    if (!(eps >= 0.0f) || (self ? nr < 2 : (nr == 0 || nc == 0))) {
        RNNG_TRY(edge_buffers(g, 1));
        g->n_edges = 0;
        return RNNG_OK;
    }
    const int k = eps >= 1e9f ? 1000000000 : (int)floorf(eps);
    EditJoin J;
    J.self_join = self ? 1 : 0;
    J.maxlen = std::max(A->len_sorted.empty() ? 0 : A->len_sorted.back(),
                        B->len_sorted.empty() ? 0 : B->len_sorted.back());
    // codes of both sides in one alphabet: each set numbers the bytes it
    // holds, so two sets with different byte sets are renumbered by the union
    LocalBuf rc_a, rc_b, lut;
    const uint8_t* ca = A->codes.as<uint8_t>();
    const uint8_t* cb = B->codes.as<uint8_t>();
    J.planes = A->planes;
    if (!self && std::memcmp(A->seen, B->seen, sizeof(A->seen)) != 0) {
        uint8_t uni[256], la[256] = {0}, lb[256] = {0};
        int sigma = 0;
        for (int c = 0; c < 256; ++c) uni[c] = (A->seen[c] || B->seen[c]) ? (uint8_t)sigma++ : 0;
        int qa = 0, qb = 0;
        for (int c = 0; c < 256; ++c) {
            if (A->seen[c]) la[qa++] = uni[c];
            if (B->seen[c]) lb[qb++] = uni[c];
        }
        J.planes = 1;
        while ((1 << J.planes) < sigma) ++J.planes;
        RNNG_TRY(lut.reserve(512));
        RNNG_HIP_TRY(hipMemcpyAsync(lut.ptr, la, 256, hipMemcpyHostToDevice, stream()));
        RNNG_HIP_TRY(hipMemcpyAsync(lut.as<uint8_t>() + 256, lb, 256, hipMemcpyHostToDevice, stream()));
        RNNG_TRY(rc_a.reserve((size_t)A->codes_len));
        RNNG_TRY(rc_b.reserve((size_t)B->codes_len));
        k_recode<<<blocks_for(A->codes_len, 256), 256, 0, stream()>>>(ca, rc_a.as<uint8_t>(), A->codes_len,
                                                                      lut.as<uint8_t>());
        k_recode<<<blocks_for(B->codes_len, 256), 256, 0, stream()>>>(cb, rc_b.as<uint8_t>(), B->codes_len,
                                                                      lut.as<uint8_t>() + 256);
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));   // the host tables above are on the stack
        ca = rc_a.as<uint8_t>();
        cb = rc_b.as<uint8_t>();
    }
    J.R = side_of(A, ca);
    J.C = side_of(B, cb);
    // the clamped 3-gram L1 of two strings is at most 512, so any radius above
    // 100 passes every length-compatible pair (and -(6k + 1) stays in range)
    const int kf = std::min(k, 100);
    const int64_t kw = std::min(k, 1 << 30);
    // work list: row block x column-tile chunk over the length-compatible range
    ws.work_host.clear();
    const int nrb = (int)(A->n_pad / ERB);
    const std::vector<int32_t>& lb_ = B->len_sorted;
    for (int rb = 0; rb < nrb; ++rb) {
        if ((int64_t)rb * ERB >= nr) break;
        const int a_first = rb * ERB, a_last = std::min(nr - 1, rb * ERB + ERB - 1);
        const int64_t maxlen = (int64_t)A->len_sorted[a_last] + kw;
        const int b_end = (int)(std::upper_bound(lb_.begin(), lb_.end(), (int32_t)std::min<int64_t>(maxlen, INT32_MAX)) -
                                lb_.begin());
        int ct0 = rb * (ERB / 32);
        if (!self) {
            const int64_t minlen = (int64_t)A->len_sorted[a_first] - kw;
            const int b_begin = (int)(std::lower_bound(lb_.begin(), lb_.end(),
                                                       (int32_t)std::max<int64_t>(minlen, INT32_MIN)) -
                                      lb_.begin());
            ct0 = b_begin / 32;
        }
        const int ct_end = (b_end + 31) / 32;
        for (int ct = ct0; ct < ct_end; ct += ECT)
            ws.work_host.push_back(make_int4(rb, ct, std::min(ct_end, ct + ECT), 0));
    }
    xcd_permute(ws.work_host);
    RNNG_TRY(ws.counters.reserve(8 * sizeof(unsigned long long)));
    RNNG_TRY(ws.work.reserve(std::max<size_t>(1, ws.work_host.size()) * sizeof(int4)));
    if (!ws.work_host.empty())
        RNNG_HIP_TRY(hipMemcpyAsync(ws.work.ptr, ws.work_host.data(), ws.work_host.size() * sizeof(int4),
                                    hipMemcpyHostToDevice, stream()));
    const int nmax = std::max(nr, nc);
    if (ws.cand_cap == 0) {
        ws.cand_cap = std::max<unsigned long long>(1ull << 20, 128ull * (unsigned long long)nmax);
        RNNG_TRY(ws.cand.reserve(ws.cand_cap * sizeof(int2)));
    }
    if (g->cap == 0)
        RNNG_TRY(edge_buffers(g, g_edge_hint > 0 ? g_edge_hint
                                                 : std::max<int64_t>(1 << 20, 64 * (int64_t)nmax)));
    Events ev;
    RNNG_TRY(ev.create());
    for (int attempt = 0; attempt < 3; ++attempt) {
        k_zero_u64<<<1, 64, 0, stream()>>>(ws.counters.as<unsigned long long>(), 8);
        RNNG_HIP_TRY(hipEventRecord(ev.e[0], stream()));
        if (!ws.work_host.empty())
            k_edit_filter<<<(unsigned)ws.work_host.size(), 256, 0, stream()>>>(
                A->mxa.as<i32x4_t>(), B->myb.as<i32x4_t>(), A->mnrm.as<int>(), B->mnrm.as<int>(),
                A->slen.as<int>(), B->slen.as<int>(), ws.work.as<int4>(), nr, nc, J.self_join, k, -(6 * kf + 1),
                ws.cand.as<int2>(), ws.cand_cap, ws.counters.as<unsigned long long>());
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_HIP_TRY(hipEventRecord(ev.e[1], stream()));
        RNNG_TRY(dispatch_verify(J, ws, k, g));
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
                st->kernel_ms += ev.ms(1, 2);
                st->kernel_units += (int64_t)cnt[2];
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

// ---------------------------------------------------------------------------
// String sets from device-resident atoms (original order) and offsets
// ---------------------------------------------------------------------------
__global__ void k_str_lengths(const int64_t* __restrict__ off, int64_t n, unsigned* __restrict__ len,
                              int* __restrict__ id) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        len[i] = (unsigned)(off[i + 1] - off[i]);
        id[i] = (int)i;
    }
}

__global__ void k_str_sorted(const int* __restrict__ order, const int64_t* __restrict__ off,
                             const unsigned* __restrict__ slen, int64_t n, int64_t* __restrict__ soff,
                             int64_t* __restrict__ alen) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) {
        soff[a] = off[order[a]];
        alen[a] = ((int64_t)slen[a] + 15) / 16 * 16;
    }
}

// bytes that occur: one LDS table per block, one global OR per byte value
__global__ __launch_bounds__(256) void k_byte_seen(const uint8_t* __restrict__ atoms, int64_t total,
                                                   unsigned* __restrict__ seen) {
    __shared__ unsigned s[256];
    s[threadIdx.x] = 0u;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (int64_t)gridDim.x * 256)
        s[atoms[i]] = 1u;
    __syncthreads();
    if (s[threadIdx.x]) atomicOr(&seen[threadIdx.x], 1u);
}

// dense codes of each string in length order, one wave per string
__global__ __launch_bounds__(256) void k_write_codes(const uint8_t* __restrict__ atoms,
                                                     const int64_t* __restrict__ soff,
                                                     const unsigned* __restrict__ slen,
                                                     const int64_t* __restrict__ coff, int64_t n,
                                                     const uint8_t* __restrict__ lut, uint8_t* __restrict__ codes) {
    const int lane = lane_id();
    for (int64_t a = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; a < n;
         a += ((int64_t)gridDim.x * blockDim.x) >> 6) {
        const uint8_t* src = atoms + soff[a];
        uint8_t* dst = codes + coff[a];
        const unsigned len = slen[a];
        for (unsigned i = lane; i < len; i += 64) dst[i] = lut[src[i]];
    }
}

__global__ void k_fill_bytes(uint32_t* __restrict__ p, int64_t words) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) p[i] = 0u;
}

// S->atoms (total bytes) and S->offsets (n + 1, from 0) hold the strings in
// their original order; builds the length order (a stable sort, so equal
// lengths keep their ids' order), the dense alphabet codes, the 3-gram
// profiles and the filter fragments on the device.
int strings_from_device(rnng_strings_s* S, int64_t n, int64_t total) {
    S->n = n;
    S->total = total;
    RNNG_TRY(S->order.reserve((size_t)std::max<int64_t>(n, 1) * sizeof(int32_t)));
    RNNG_TRY(S->slen.reserve((size_t)std::max<int64_t>(n, 1) * sizeof(int32_t)));
    RNNG_TRY(S->soff.reserve((size_t)std::max<int64_t>(n, 1) * sizeof(int64_t)));
    LocalBuf alen, seen, lut;
    RNNG_TRY(alen.reserve((size_t)std::max<int64_t>(n, 1) * sizeof(int64_t)));
    RNNG_TRY(S->coff.reserve((size_t)std::max<int64_t>(n, 1) * sizeof(int64_t)));
    if (n > 0) {
        k_str_lengths<<<blocks_for(n, 256), 256, 0, stream()>>>(S->offsets.as<int64_t>(), n, S->slen.as<unsigned>(),
                                                                 S->order.as<int>());
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(sort_pairs_u32_i32(S->slen.as<uint32_t>(), S->order.as<int32_t>(), n, 32));
        k_str_sorted<<<blocks_for(n, 256), 256, 0, stream()>>>(S->order.as<int>(), S->offsets.as<int64_t>(),
                                                                S->slen.as<unsigned>(), n, S->soff.as<int64_t>(),
                                                                alen.as<int64_t>());
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(exclusive_scan_i64(alen.as<int64_t>(), S->coff.as<int64_t>(), n));
        S->len_sorted.resize((size_t)n);
        RNNG_HIP_TRY(hipMemcpyAsync(S->len_sorted.data(), S->slen.ptr, (size_t)n * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, stream()));
    }
    // alphabet: the bytes that occur, numbered in byte order
    RNNG_TRY(seen.reserve(256 * sizeof(unsigned)));
    k_fill_bytes<<<1, 256, 0, stream()>>>(seen.as<uint32_t>(), 256);
    if (total > 0)
        k_byte_seen<<<(unsigned)std::min<int64_t>(4096, (total + 255) / 256), 256, 0, stream()>>>(
            S->atoms.as<uint8_t>(), total, seen.as<unsigned>());
    RNNG_HIP_TRY(hipGetLastError());
    unsigned hs[256];
    int64_t last[2] = {0, 0};
    RNNG_HIP_TRY(hipMemcpyAsync(hs, seen.ptr, sizeof(hs), hipMemcpyDeviceToHost, stream()));
    if (n > 0) {
        RNNG_HIP_TRY(hipMemcpyAsync(&last[0], S->coff.as<int64_t>() + (n - 1), sizeof(int64_t), hipMemcpyDeviceToHost,
                                    stream()));
        RNNG_HIP_TRY(hipMemcpyAsync(&last[1], alen.as<int64_t>() + (n - 1), sizeof(int64_t), hipMemcpyDeviceToHost,
                                    stream()));
    }
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    uint8_t code[256];
    int sigma = 0;
    for (int c = 0; c < 256; ++c) {
        S->seen[c] = hs[c] != 0;
        code[c] = S->seen[c] ? (uint8_t)sigma++ : 0;
    }
    S->planes = 1;
    while ((1 << S->planes) < sigma) ++S->planes;
    // codes: each string 16-byte aligned in length order; the verify kernels
    // read up to 128 bytes past a string's end
    S->codes_len = last[0] + last[1] + 256;
    RNNG_TRY(S->codes.reserve((size_t)(S->codes_len + 3) / 4 * 4));
    RNNG_TRY(lut.reserve(256));
    RNNG_HIP_TRY(hipMemcpyAsync(lut.ptr, code, 256, hipMemcpyHostToDevice, stream()));
    k_fill_bytes<<<blocks_for((S->codes_len + 3) / 4, 256), 256, 0, stream()>>>(S->codes.as<uint32_t>(),
                                                                               (S->codes_len + 3) / 4);
    if (n > 0)
        k_write_codes<<<(unsigned)std::min<int64_t>(8192, (n + 3) / 4), 256, 0, stream()>>>(
            S->atoms.as<uint8_t>(), S->soff.as<int64_t>(), S->slen.as<unsigned>(), S->coff.as<int64_t>(), n,
            lut.as<uint8_t>(), S->codes.as<uint8_t>());
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_TRY(S->prof.reserve(std::max<size_t>(16, (size_t)n * 64)));
    if (n > 0) {
        k_profiles<<<blocks_for(n, 256), 256, 0, stream()>>>(S->atoms.as<uint8_t>(), S->soff.as<int64_t>(),
                                                              S->slen.as<int>(), (int)n, S->prof.as<uint4>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    // MFMA filter fragments, padded to whole row blocks
    S->n_pad = (n + ERB - 1) / ERB * ERB;
    if (n > 0) {
        // 64 bins x TLEV levels = one int8 per slot: n_pad x 256 bytes per form
        const size_t fb = (size_t)S->n_pad * 64 * TLEV;
        RNNG_TRY(S->mxa.reserve(fb));
        RNNG_TRY(S->myb.reserve(fb));
        RNNG_TRY(S->mnrm.reserve((size_t)S->n_pad * sizeof(int)));
        k_edit_frags<<<blocks_for(S->n_pad * EKS * 2, 256), 256, 0, stream()>>>(
            S->prof.as<uint4>(), (int)n, S->n_pad, S->mxa.as<i32x4_t>(), S->myb.as<i32x4_t>(), S->mnrm.as<int>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

// ---------------------------------------------------------------------------
// .seqs files (include/point.hpp:170-205): one string per line, a trailing
// '\r' dropped, empty lines skipped, ids in file order
// ---------------------------------------------------------------------------
__global__ void k_newline_flags(const uint8_t* __restrict__ b, int64_t size, uint8_t* __restrict__ f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < size) f[i] = b[i] == (uint8_t)'\n' ? 1 : 0;
}

// line k spans [nl[k-1] + 1, nl[k]) (the first from 0, the last to the end)
__global__ void k_line_spans(const uint8_t* __restrict__ b, int64_t size, const int64_t* __restrict__ nl,
                             int64_t nnl, longlong2* __restrict__ span, uint8_t* __restrict__ keep) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nnl) return;
    const int64_t s = k == 0 ? 0 : nl[k - 1] + 1;
    int64_t e = k < nnl ? nl[k] : size;
    if (e > s && b[e - 1] == (uint8_t)'\r') --e;
    span[k] = make_longlong2(s, e);
    keep[k] = e > s ? 1 : 0;
}

__global__ void k_span_lengths(const longlong2* __restrict__ span, int64_t start, int64_t n,
                               int64_t* __restrict__ len) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) len[i] = i < n ? span[start + i].y - span[start + i].x : 0;
}

// one wave per line: its bytes to the string's place in the atom buffer
__global__ __launch_bounds__(256) void k_gather_lines(const uint8_t* __restrict__ b,
                                                      const longlong2* __restrict__ span, int64_t start, int64_t n,
                                                      const int64_t* __restrict__ off, uint8_t* __restrict__ atoms) {
    const int lane = lane_id();
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n;
         i += ((int64_t)gridDim.x * blockDim.x) >> 6) {
        const longlong2 sp = span[start + i];
        uint8_t* dst = atoms + off[i];
        for (int64_t j = sp.x + lane; j < sp.y; j += 64) dst[j - sp.x] = b[j];
    }
}

int read_seqs_device(const char* path, int64_t start, int64_t count, rnng_strings_s* S) {
    int64_t size = 0;
    RNNG_TRY(upload_file_bytes(path, nullptr, &size, 0));
    LocalBuf raw, flags, nl, cnt, tmp, span, keep, spank;
    RNNG_TRY(raw.reserve((size_t)std::max<int64_t>(size, 16)));
    RNNG_TRY(upload_file_bytes(path, raw.as<uint8_t>(), &size, (int64_t)raw.bytes));
    // newline positions
    RNNG_TRY(flags.reserve((size_t)std::max<int64_t>(size, 1)));
    RNNG_TRY(nl.reserve((size_t)std::max<int64_t>(size, 1) * sizeof(int64_t)));
    RNNG_TRY(cnt.reserve(2 * sizeof(unsigned long long)));
    if (size > 0) {
        k_newline_flags<<<blocks_for(size, 256), 256, 0, stream()>>>(raw.as<uint8_t>(), size, flags.as<uint8_t>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    unsigned long long hc[2] = {0, 0};
    if (size > 0) {
        rocprim::counting_iterator<int64_t> pos(0);
        size_t tb = 0;
        RNNG_HIP_TRY(rocprim::select(nullptr, tb, pos, flags.as<uint8_t>(), nl.as<int64_t>(),
                                     cnt.as<unsigned long long>(), (size_t)size, stream()));
        RNNG_TRY(tmp.reserve(tb));
        RNNG_HIP_TRY(rocprim::select(tmp.ptr, tb, pos, flags.as<uint8_t>(), nl.as<int64_t>(),
                                     cnt.as<unsigned long long>(), (size_t)size, stream()));
        RNNG_HIP_TRY(hipMemcpyAsync(hc, cnt.ptr, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    }
    const int64_t nnl = (int64_t)hc[0];
    // non-empty lines
    const int64_t nseg = nnl + 1;
    RNNG_TRY(span.reserve((size_t)nseg * sizeof(longlong2)));
    RNNG_TRY(spank.reserve((size_t)nseg * sizeof(longlong2)));
    RNNG_TRY(keep.reserve((size_t)nseg));
    k_line_spans<<<blocks_for(nseg, 256), 256, 0, stream()>>>(raw.as<uint8_t>(), size, nl.as<int64_t>(), nnl,
                                                              span.as<longlong2>(), keep.as<uint8_t>());
    RNNG_HIP_TRY(hipGetLastError());
    {
        size_t tb = 0;
        RNNG_HIP_TRY(rocprim::select(nullptr, tb, span.as<longlong2>(), keep.as<uint8_t>(), spank.as<longlong2>(),
                                     cnt.as<unsigned long long>() + 1, (size_t)nseg, stream()));
        LocalBuf t2;
        RNNG_TRY(t2.reserve(tb));
        RNNG_HIP_TRY(rocprim::select(t2.ptr, tb, span.as<longlong2>(), keep.as<uint8_t>(), spank.as<longlong2>(),
                                     cnt.as<unsigned long long>() + 1, (size_t)nseg, stream()));
        RNNG_HIP_TRY(hipMemcpyAsync(&hc[1], cnt.as<unsigned long long>() + 1, sizeof(unsigned long long),
                                    hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    }
    const int64_t lines = (int64_t)hc[1];
    if (start < 0 || start > lines) {
        set_error("start beyond the last string of the file");
        return RNNG_ERR_ARG;
    }
    if (count < 0) count = lines - start;
    if (start + count > lines) {
        set_error("start + count beyond the last string of the file");
        return RNNG_ERR_ARG;
    }
    if (count > 0x7fffffff - 64) {
        set_error("string sets are limited to 2^31 - 64 strings");
        return RNNG_ERR_UNSUPPORTED;
    }
    const int64_t n = count;
    LocalBuf len;
    RNNG_TRY(len.reserve((size_t)(n + 1) * sizeof(int64_t)));
    RNNG_TRY(S->offsets.reserve((size_t)(n + 1) * sizeof(int64_t)));
    k_span_lengths<<<blocks_for(n + 1, 256), 256, 0, stream()>>>(spank.as<longlong2>(), start, n, len.as<int64_t>());
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_TRY(exclusive_scan_i64(len.as<int64_t>(), S->offsets.as<int64_t>(), n + 1));
    int64_t total = 0;
    RNNG_HIP_TRY(hipMemcpyAsync(&total, S->offsets.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost,
                                stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    RNNG_TRY(S->atoms.reserve((size_t)std::max<int64_t>(total, 16)));
    if (n > 0) {
        k_gather_lines<<<(unsigned)std::min<int64_t>(8192, (n + 3) / 4), 256, 0, stream()>>>(
            raw.as<uint8_t>(), spank.as<longlong2>(), start, n, S->offsets.as<int64_t>(), S->atoms.as<uint8_t>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    return strings_from_device(S, n, total);
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
        if (offsets[i + 1] < offsets[i] || offsets[i + 1] - offsets[i] > 0x7fffffff) {
            set_error("rnng_strings_create: offsets must be non-decreasing, strings below 2^31 atoms");
            return RNNG_ERR_ARG;
        }
    RNNG_TRY(ensure_init());
    rnng_strings_s* S = new rnng_strings_s();
    const int64_t base = n > 0 ? offsets[0] : 0;
    const int64_t total = n > 0 ? offsets[n] - base : 0;
    std::vector<int64_t> off((size_t)n + 1);
    for (int64_t i = 0; i <= n; ++i) off[(size_t)i] = n > 0 ? offsets[i] - base : 0;
    int rc = S->atoms.reserve((size_t)std::max<int64_t>(total, 16));
    if (rc == RNNG_OK) rc = S->offsets.reserve(off.size() * sizeof(int64_t));
    if (rc == RNNG_OK && total > 0 &&
        hipMemcpyAsync(S->atoms.ptr, atoms + base, (size_t)total, hipMemcpyHostToDevice, stream()) != hipSuccess) {
        set_error("rnng_strings_create: upload failed");
        rc = RNNG_ERR_HIP;
    }
    if (rc == RNNG_OK &&
        hipMemcpyAsync(S->offsets.ptr, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream()) !=
            hipSuccess) {
        set_error("rnng_strings_create: upload failed");
        rc = RNNG_ERR_HIP;
    }
    if (rc == RNNG_OK) rc = strings_from_device(S, n, total);
    if (rc != RNNG_OK) {
        rnng_strings_destroy(S);
        return rc;
    }
    *out = S;
    return RNNG_OK;
}

int rnng_strings_read_file(const char* path, int64_t start, int64_t count, rnng_strings_t* out) {
    if (!path || !out) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    *out = nullptr;
    RNNG_TRY(ensure_init());
    rnng_strings_s* S = new rnng_strings_s();
    const int rc = read_seqs_device(path, start, count, S);
    if (rc != RNNG_OK) {
        rnng_strings_destroy(S);
        return rc;
    }
    *out = S;
    return RNNG_OK;
}

int rnng_strings_size(rnng_strings_t S, int64_t* n, int64_t* atoms) {
    if (!S || !n) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    *n = S->n;
    if (atoms) *atoms = S->total;
    return RNNG_OK;
}

int rnng_strings_copy(rnng_strings_t S, char* atoms, int64_t* offsets) {
    if (!S) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    if (atoms && S->total > 0)
        RNNG_HIP_TRY(hipMemcpyAsync(atoms, S->atoms.ptr, (size_t)S->total, hipMemcpyDeviceToHost, stream()));
    if (offsets)
        RNNG_HIP_TRY(hipMemcpyAsync(offsets, S->offsets.ptr, (size_t)(S->n + 1) * sizeof(int64_t),
                                    hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

int rnng_debug_mfma_i8(const int8_t* a, const int8_t* b, int32_t* d) {
    if (!a || !b || !d) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    LocalBuf da, db, dd;
    RNNG_TRY(da.reserve(1024));
    RNNG_TRY(db.reserve(1024));
    RNNG_TRY(dd.reserve(64 * 16 * sizeof(int)));
    RNNG_HIP_TRY(hipMemcpyAsync(da.ptr, a, 1024, hipMemcpyHostToDevice, stream()));
    RNNG_HIP_TRY(hipMemcpyAsync(db.ptr, b, 1024, hipMemcpyHostToDevice, stream()));
    k_debug_mfma_i8<<<1, 64, 0, stream()>>>(da.as<i32x4_t>(), db.as<i32x4_t>(), dd.as<int>());
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_HIP_TRY(hipMemcpyAsync(d, dd.ptr, 64 * 16 * sizeof(int), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
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
    S->mxa.release();
    S->myb.release();
    S->mnrm.release();
    S->codes.release();
    S->coff.release();
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
    const int rc = edit_join(S, S, true, eps, g, st);
    if (rc != RNNG_OK) {
        rnng_graph_destroy(g);
        return rc;
    }
    *out = g;
    return RNNG_OK;
}

int rnng_edit_graph_build_bipartite(rnng_strings_t a, rnng_strings_t b, float eps, rnng_graph_t* out,
                                    rnng_stats* st) {
    if (!a || !b || !out) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    *out = nullptr;
    RNNG_TRY(ensure_init());
    std::lock_guard<std::mutex> lk(g_edit_mu);
    rnng_graph_s* g = new rnng_graph_s();
    const int rc = edit_join(a, b, false, eps, g, st);
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
