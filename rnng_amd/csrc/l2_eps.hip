// L2 epsilon-range kernels for gfx950 and their host orchestration.
//
// Reference semantics (gabe-raulet/RNNG):
//   L2Distance<float>::operator()  drivers/systolic.cpp:253-269  -- sequential
//       FP32 sum over dims 0..d-1 of fl(fl(p_k - q_k)^2), not fused, then sqrt
//   edge test dist <= radius       include/search.hpp:244,399 (radius is float)
//   all-pairs query over a block   drivers/systolic.cpp:123-147
//
// Two distance-stage algorithms:
//   PLAIN  k_l2_plain: LDS-tiled FP32 kernel that evaluates every pair in the
//          reference order (sub, mul, add, contraction off) and compares the
//          squared sum with T(eps), the largest float whose sqrtf is <= eps.
//   MFMA   k_l2_mfma_filter: BF16 32x32x16 MFMA over augmented vectors that
//          produces  acc = |x~|^2 + |y~|^2 - 2 x~.y~ - C0  for every pair, where
//          C0 bounds (from above) what acc can be for any reference edge.  A
//          pair is a candidate iff acc < 0, tested as the sign bit of the OR of
//          the accumulators, so the epilogue costs half a VALU op per pair.
//          k_l2_recheck then recomputes every candidate in the reference's FP32
//          order (that decides the edge) and sends pairs within 1e-5*eps of eps
//          to an FP64 check that is counted and reported.
#include "common.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "mfma.hpp"

namespace rnng {

int64_t g_edge_hint = 0;   // initial edge-buffer size of new graphs (0: default)

int edge_buffers(rnng_graph_s* g, int64_t cap) {
    if (cap < 1) cap = 1;
    RNNG_TRY(g->ij.reserve((size_t)cap * sizeof(int2)));
    RNNG_TRY(g->w.reserve((size_t)cap * sizeof(float)));
    g->cap = cap;
    return RNNG_OK;
}

float sq_threshold_host(float eps) {
    if (!(eps >= 0.0f)) return -1.0f;
    if (std::isinf(eps)) return INFINITY;
    float v = (float)((double)eps * (double)eps);
    if (std::isinf(v)) v = FLT_MAX;
    while (v > 0.0f && std::sqrt(v) > eps) v = std::nextafter(v, -INFINITY);
    for (;;) {
        float up = std::nextafter(v, INFINITY);
        if (std::isinf(up) || std::sqrt(up) > eps) break;
        v = up;
    }
    return v;
}

namespace {

// ------------------------------------------------------------------------
// Small utility kernels
// ------------------------------------------------------------------------
__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u32(unsigned* p, int n, unsigned v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Order-preserving float <-> uint mapping for atomic min/max.
__device__ __forceinline__ unsigned f2ord(float f) {
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(unsigned u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// ------------------------------------------------------------------------
// PLAIN: reference-order FP32 tiled kernel
// ------------------------------------------------------------------------
constexpr int PB = 128;   // pairs tile edge per workgroup
constexpr int PK = 16;    // dims staged per LDS round

// Tile t of the upper triangle (row-major, diagonal included) of an nb x nb grid.
__device__ __forceinline__ void decode_upper(long long t, int nb, int& bi, int& bj) {
    const double a = 2.0 * nb + 1.0;
    long long b = (long long)((a - sqrt(a * a - 8.0 * (double)t)) * 0.5);
    auto start = [nb](long long r) { return r * nb - r * (r - 1) / 2; };
    if (b < 0) b = 0;
    while (b > 0 && start(b) > t) --b;
    while (b + 1 < nb && start(b + 1) <= t) ++b;
    bi = (int)b;
    bj = (int)(b + (t - start(b)));
}

// block_base: a dispatch's grid is limited to 2^32 work-items, so large
// launches are split into chunks of blocks that each start at block_base.
__global__ __launch_bounds__(256) void k_l2_plain(
    const float* __restrict__ X, int nx, const float* __restrict__ Y, int ny, int d, float T,
    int self_join, int nbx, int nby, long long block_base, int2* __restrict__ out_ij,
    float* __restrict__ out_w, unsigned long long cap, unsigned long long* __restrict__ counter) {
#pragma clang fp contract(off)
    __shared__ float As[PK][PB + 4];
    __shared__ float Bs[PK][PB + 4];

    int bi, bj;
    const long long blk = block_base + (long long)blockIdx.x;
    if (self_join) {
        decode_upper(blk, nbx, bi, bj);
    } else {
        bi = (int)(blk / nby);
        bj = (int)(blk % nby);
    }
    const int row0 = bi * PB, col0 = bj * PB;
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int lr = tid >> 1, lk = (tid & 1) * 8;   // loader: row and k offset

    float acc[8][8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[r][c] = 0.0f;

    for (int k0 = 0; k0 < d; k0 += PK) {
        {
            const int gr = row0 + lr, gc = col0 + lr;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = k0 + lk + e;
                As[lk + e][lr] = (gr < nx && k < d) ? X[(size_t)gr * d + k] : 0.0f;
                Bs[lk + e][lr] = (gc < ny && k < d) ? Y[(size_t)gc * d + k] : 0.0f;
            }
        }
        __syncthreads();
        const int kend = min(PK, d - k0);   // staged dims beyond d are zero: skip them
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

    // Padding dims contribute exact +0.0, so acc equals the reference sum.
    unsigned cnt = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int i = row0 + (r < 4 ? ty * 4 + r : 64 + ty * 4 + r - 4);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int j = col0 + (c < 4 ? tx * 4 + c : 64 + tx * 4 + c - 4);
            const bool ok = i < nx && j < ny && (!self_join || i < j) && acc[r][c] <= T;
            cnt += ok ? 1u : 0u;
        }
    }
    if (!__any(cnt != 0)) return;
    unsigned long long pos = wave_reserve(cnt, counter);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int i = row0 + (r < 4 ? ty * 4 + r : 64 + ty * 4 + r - 4);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int j = col0 + (c < 4 ? tx * 4 + c : 64 + tx * 4 + c - 4);
            const bool ok = i < nx && j < ny && (!self_join || i < j) && acc[r][c] <= T;
            if (ok) {
                if (pos < cap) {
                    out_ij[pos] = make_int2(i, j);
                    out_w[pos] = sqrtf(acc[r][c]);
                }
                ++pos;
            }
        }
    }
}

// ------------------------------------------------------------------------
// MFMA path: data preparation
// ------------------------------------------------------------------------
// Fragment layout (shared by the x-form and the y-form):
//   tile t = 32 points, k-step s = 16 dims, lane l in 0..63:
//   frag[(t*KS + s)*64 + l] = 8 bf16 of point 32t + (l & 31),
//                             dims 16s + 8(l >> 5) .. +7
// which is exactly the per-lane operand of v_mfma_f32_32x32x16_bf16 (A: row
// = lane&31; B: column = lane&31; the 16 dims of a step split by lane>>5), so
// a wave loads one 1 KB fragment with one fully coalesced 16-byte load.
//
// x~ = bf16(x - mu) (rows), y~ = -2 bf16(y - mu) (columns), dims >= d zero,
// KS = ceil(d / 16).  The squared norms |x~|^2 and |y~/2|^2 are FP32 arrays
// xn (rows) and ycn (columns, with -C0 added): the accumulator of pair (i, j)
// starts at xn[i] + ycn[j] and the MFMAs add x~.y~, so it ends at
// |x' - y'|^2 - C0 and its sign bit marks a candidate.  Padding points carry
// 2^100 as their norm, so any pair with one is positive.

constexpr int NSTAT = 8;
// stat slots (unsigned, ordered-float or flags) in ds_stat[]
enum { ST_INT = 0, ST_MAXE = 1, ST_MAXN = 2, ST_MAXRES = 3, ST_MAXDEV = 4, ST_NONFIN = 5 };


// Per-dimension sums, minima and maxima, and the all-integer flag.  A block
// covers COLSTAT_ROWS rows; for d <= 256 its threads split into P = 256 / d
// row phases (thread t: column t % d, rows r0 + t / d, + P, ...), so every
// thread streams coalesced row segments, and the phases are combined in LDS
// before one set of atomics per column and block.  Wider rows loop over
// columns with a stride of 256.
constexpr int COLSTAT_ROWS = 1024;

__global__ __launch_bounds__(256) void k_colstats(const float* __restrict__ X, int n, int d,
                                                  double* __restrict__ colsum,
                                                  unsigned* __restrict__ colmin,
                                                  unsigned* __restrict__ colmax,
                                                  unsigned* __restrict__ stat) {
    __shared__ double ssum[256];
    __shared__ float smin[256], smax[256];
    const int t = threadIdx.x;
    const int r0 = blockIdx.x * COLSTAT_ROWS;
    const int r1 = min(n, r0 + COLSTAT_ROWS);
    bool isint = true;
    if (d <= 256) {
        const int P = 256 / d;
        const int k = t % d, ph = t / d;
        double s = 0.0;
        float mn = INFINITY, mx = -INFINITY;
        if (ph < P) {
#pragma unroll 4
            for (int r = r0 + ph; r < r1; r += P) {
                const float v = X[(size_t)r * d + k];
                s += (double)v;
                mn = fminf(mn, v);
                mx = fmaxf(mx, v);
                isint = isint && (v == rintf(v));
            }
        }
        ssum[t] = s;
        smin[t] = mn;
        smax[t] = mx;
        __syncthreads();
        if (t < d) {
            for (int q = 1; q < P; ++q) {
                s += ssum[t + q * d];
                mn = fminf(mn, smin[t + q * d]);
                mx = fmaxf(mx, smax[t + q * d]);
            }
            atomicAdd(&colsum[t], s);
            atomicMin(&colmin[t], f2ord(mn));
            atomicMax(&colmax[t], f2ord(mx));
        }
    } else {
        for (int k = t; k < d; k += 256) {
            double s = 0.0;
            float mn = INFINITY, mx = -INFINITY;
            for (int r = r0; r < r1; ++r) {
                const float v = X[(size_t)r * d + k];
                s += (double)v;
                mn = fminf(mn, v);
                mx = fmaxf(mx, v);
                isint = isint && (v == rintf(v));
            }
            atomicAdd(&colsum[k], s);
            atomicMin(&colmin[k], f2ord(mn));
            atomicMax(&colmax[k], f2ord(mx));
        }
    }
    if (!__all(isint) && lane_id() == 0) atomicAnd(&stat[ST_INT], 0u);
}

// mu_k = per-dimension mean; rounded to an integer when every value is one so
// that x - mu stays exact.  Also records the largest |x - mu| and the widest
// per-dimension range (both as ordered floats), and flags a column whose sum is
// not finite (it holds a NaN or an infinity: the double sum of floats cannot
// overflow).
__global__ void k_center(const double* __restrict__ colsum, const unsigned* __restrict__ colmin,
                         const unsigned* __restrict__ colmax, int n, int d,
                         float* __restrict__ mu, unsigned* __restrict__ stat,
                         unsigned* __restrict__ range_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d) return;
    const bool isint = stat[ST_INT] != 0;
    double m = n > 0 ? colsum[k] / (double)n : 0.0;
    if (!isfinite(m)) atomicOr(&stat[ST_NONFIN], 1u);
    float muk = isint ? (float)rint(m) : (float)m;
    mu[k] = muk;
    const float mn = ord2f(colmin[k]), mx = ord2f(colmax[k]);
    const float dev = fmaxf(fabsf(mx - muk), fabsf(mn - muk));
    atomicMax(&stat[ST_MAXDEV], f2ord(dev));
    atomicMax(range_out, f2ord(mx - mn));
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// One workgroup per 64 points: the points' rows are staged in LDS with
// contiguous loads, each point's norm and conversion error reduced over the
// wave on the way, then the x~ and/or y~ fragments (either output may be null) are
// written 16 bytes per lane, contiguous within each 1 KB fragment, and the
// norms (FP32, 2^100 for padding) to xn and/or yn.
// idmap (may be null): padded position -> point index (-1 = padding); without
// it position p holds point p for p < n.
constexpr int PP = 64;   // points per prep workgroup
constexpr float kPadNorm = 1.2676506e30f;   // 2^100

__global__ __launch_bounds__(256) void k_prep_frags(
    const float* __restrict__ X, int n, int n_pad, int d, int KS, const float* __restrict__ mu,
    const int* __restrict__ idmap, uint16_t* __restrict__ xa, uint16_t* __restrict__ yb,
    float* __restrict__ xn, float* __restrict__ yn, unsigned* __restrict__ stat, int l16) {
    extern __shared__ float xs[];          // [PP][d + 1]: x~ = bf16(x - mu) as float
    __shared__ int preal[PP];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int p0 = blockIdx.x * PP;
    const int ld = d + 1;
    if (tid < PP) {
        const int p = p0 + tid;
        int src = -1;
        if (p < n_pad) src = idmap ? idmap[p] : (p < n ? p : -1);
        preal[tid] = src;
    }
    __syncthreads();
    // rows to LDS as x~, with each row's norm and conversion error summed over
    // the wave (in double) on the way: X is read once
    float e_max = 0.0f, n_max = 0.0f;
    for (int r = wave; r < PP; r += 4) {
        const int src = preal[r];
        double nrm = 0.0, err = 0.0;
        for (int k = lane; k < d; k += 64) {
            float v = 0.0f;
            if (src >= 0) {
                const double exact = (double)X[(size_t)src * d + k] - (double)mu[k];
                v = bf16_to_f32(f32_to_bf16_rn((float)exact));
                nrm += (double)v * (double)v;
                const double de = (double)v - exact;
                err += de * de;
            }
            xs[r * ld + k] = v;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            nrm += __shfl_xor(nrm, o, 64);
            err += __shfl_xor(err, o, 64);
        }
        if (lane == 0 && p0 + r < n_pad) {
            const float nv = src >= 0 ? (float)nrm : kPadNorm;
            if (xn) xn[p0 + r] = nv;
            if (yn) yn[p0 + r] = nv;
            if (src >= 0) {
                e_max = fmaxf(e_max, (float)sqrt(err) * (1.0f + 1e-6f));
                n_max = fmaxf(n_max, (float)nrm * (1.0f + 1e-6f));
            }
        }
    }
    __syncthreads();
    const int nchunk = (PP / 32) * KS * 64;
    for (int c = tid; c < nchunk; c += 256) {
        const int tt = c / (KS * 64), rem = c - tt * (KS * 64);
        const int st = rem >> 6, l = rem & 63;
        const int r = tt * 32 + (l & 31);
        if (p0 + tt * 32 >= n_pad) continue;
        const int kb = 16 * st + 8 * (l >> 5);
        uint16_t xv[8], yv[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = kb + e;
            const float xf = k < d ? xs[r * ld + k] : 0.0f;
            xv[e] = f32_to_bf16_rn(xf);
            yv[e] = f32_to_bf16_rn(-2.0f * xf);
        }
        // chunk ch = 2 st + (l >> 5) (dims 8 ch .. 8 ch + 7) of tile point l & 31
        const int ch = 2 * st + (l >> 5), pt = l & 31;
        const int word = l16 ? (((pt >> 4) * (KS / 2) + (ch >> 2)) * 64 + 16 * (ch & 3) + (pt & 15)) : rem;
        const size_t slot = ((size_t)((p0 >> 5) + tt) * KS * 64 + word) * 8;
        if (xa) *(uint4*)(xa + slot) = *(const uint4*)xv;
        if (yb) *(uint4*)(yb + slot) = *(const uint4*)yv;
    }
    // the workgroup's maxima (non-negative floats: their bits order as
    // unsigned), then one atomic each, and only where the running maximum (a
    // relaxed read) is lower: every workgroup's atomics on the same two words
    // otherwise serialise at one L2 channel
    e_max = wave_max(e_max);
    n_max = wave_max(n_max);
    __shared__ float wmax[2][4];
    if (lane == 0) {
        wmax[0][wave] = e_max;
        wmax[1][wave] = n_max;
    }
    __syncthreads();
    if (tid == 0) {
        const unsigned eb = __float_as_uint(fmaxf(fmaxf(wmax[0][0], wmax[0][1]), fmaxf(wmax[0][2], wmax[0][3])));
        const unsigned nb = __float_as_uint(fmaxf(fmaxf(wmax[1][0], wmax[1][1]), fmaxf(wmax[1][2], wmax[1][3])));
        if (eb > __hip_atomic_load(&stat[ST_MAXE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&stat[ST_MAXE], eb);
        if (nb > __hip_atomic_load(&stat[ST_MAXN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&stat[ST_MAXN], nb);
    }
}

int launch_prep(const float* X, int n, int n_pad, int d, int KS, const float* mu, const int* idmap,
                uint16_t* xa, uint16_t* yb, float* xn, float* yn, unsigned* stat, bool l16) {
    if (n_pad <= 0) return RNNG_OK;
    const size_t lds = (size_t)PP * (d + 1) * sizeof(float);
    k_prep_frags<<<blocks_for(n_pad, PP), 256, lds, stream()>>>(X, n, n_pad, d, KS, mu, idmap, xa, yb, xn, yn,
                                                                stat, l16 ? 1 : 0);
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

// C0: an upper bound of what the MFMA can compute for |x' - y'|^2 for any pair
// whose reference FP32 distance is <= the band radius.
__global__ void k_finalize_c0(const unsigned* __restrict__ stat, const unsigned* __restrict__ range,
                              int d, int KS, float Tb, MfmaParams* __restrict__ out) {
    const bool isint = stat[ST_INT] != 0;
    const double maxdev = (double)ord2f(stat[ST_MAXDEV]);
    const double R = (double)ord2f(*range);
    const double maxE = (double)__uint_as_float(stat[ST_MAXE]);
    const double maxN = (double)__uint_as_float(stat[ST_MAXN]);
    const double two24 = 16777216.0;
    double c0;
    int int_mode = 0;
    if (isint && maxdev <= 256.0 && (double)d * R * R < two24 &&
        (double)Tb + 4.0 * maxN + 2.0 < two24) {
        // Every norm, the start value xn + (yn - C0) and every MFMA partial sum
        // is an integer below 2^24: all exact, so acc = |x - y|^2 - C0 and with
        // the integer C0 = floor(Tb) + 1, acc < 0 <=> |x - y|^2 <= Tb.
        c0 = floor((double)Tb) + 1.0;
        int_mode = 1;
    } else {
        const double u = 1.0 / two24;
        const double g_ref = (d + 3) * u * 1.01;
        const double smax = (double)Tb / (1.0 - g_ref);
        const double r = sqrt(smax) + 2.0 * maxE;
        const double base = r * r;
        // FP32 accumulation over KS*16 products and the start value, plus the
        // roundings of xn, yn, yn - C0 and their sum (each <= u times maxN or C0)
        const double nterms = KS * 16.0 + 2.0;
        c0 = base + 2.0 * nterms * u * (base + 4.0 * maxN) + 8.0 * u * (maxN + base);
        c0 = c0 * (1.0 + 1.0 / 1048576.0) + 1e-30;
    }
    float f = (float)c0;
    if ((double)f < c0) f = nextafterf(f, INFINITY);
    out->negc0 = -f;
    out->int_mode = int_mode;
    out->c0 = c0;
}

// Column start values: ycn[j] = yn[j] - C0 for real points (padding keeps
// its 2^100).
__global__ void k_column_bias(const float* __restrict__ yn, float* __restrict__ ycn, int n_pad,
                              const MfmaParams* __restrict__ prm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    const float v = yn[i];
    ycn[i] = v >= kPadNorm ? v : v + prm->negc0;
}

// ------------------------------------------------------------------------
// MFMA filter kernel
// ------------------------------------------------------------------------
// Workgroup = 4 waves = one row block of 4*RT row tiles (32 rows each); wave w
// owns row tiles RT*w .. RT*w+RT-1 and keeps their fragments in registers for
// the whole work item.  The work item's column tiles stream through.  Per
// column tile and wave: KS coalesced 1 KB fragment loads, RT*KS MFMAs, an OR
// over the RT*16 accumulators and one wave-uniform branch.
//   USE_LDS: the workgroup stages each column tile once in LDS (double
//   buffered, one barrier per tile) instead of every wave loading it from
//   L1, cutting vector-cache traffic by 4x; RT = 4 halves it again by using
//   every B fragment for 4 MFMAs.
// Output row/column of accumulator register r in lane l (32x32 MFMA):
//   row = 8*(r/4) + 4*(l/32) + r%4, column = l%32.


template <int KS, int RT, bool USE_LDS, int CT = 1>
__global__ __launch_bounds__(256) void k_l2_mfma_filter(
    const bf16x8_t* __restrict__ xa, const bf16x8_t* __restrict__ yb, const float* __restrict__ xn,
    const float* __restrict__ ycn, const int4* __restrict__ work, int n_rows, int n_cols, int self_join,
    const MfmaParams* __restrict__ prm, int2* __restrict__ cand, unsigned long long cap,
    unsigned long long* __restrict__ counter) {
    // CT column tiles per step (USE_LDS only): RT*CT independent accumulator
    // chains and one barrier / epilogue per RT*CT*KS MFMAs.
    constexpr int NCH = KS * 64;                // 16-byte fragments per column tile
    constexpr int PER = (CT * NCH + 255) / 256; // fragments staged per thread per step
    __shared__ bf16x8_t bs[USE_LDS ? 2 : 1][USE_LDS ? CT * NCH : 1];
    // Candidates are staged in LDS (one LDS atomic per ballot) and written out
    // with a single global reservation per work item; a single global counter
    // hit once per ballot serialises at one L2 address.
    __shared__ int2 stage[LSTAGE];
    __shared__ unsigned scount;
    __shared__ unsigned long long sbase;
    int4 wi;
    if (!load_work_item(work, wi)) return;   // padding item: no memory is touched
    const int tid = threadIdx.x;
    if (tid == 0) scount = 0u;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const int rt0 = wi.x * (4 * RT) + wave * RT;

    bf16x8_t a[RT][KS];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int s = 0; s < KS; ++s) a[r][s] = xa[((size_t)(rt0 + r) * KS + s) * 64 + lane];
    // row part of the accumulator start: the norm of each register's row
    float xr[RT][16];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int q = 0; q < 16; ++q) xr[r][q] = xn[(rt0 + r) * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3)];
    // column tiles beyond wi.z read a valid tile (wi.z - 1) and are masked out
    const int last = wi.z - 1;
    // column part (norm - C0) of this step's column tiles, loaded a step ahead
    float yc[CT];
#pragma unroll
    for (int cc = 0; cc < CT; ++cc) yc[cc] = ycn[min(wi.y + cc, last) * 32 + (lane & 31)];

    bf16x8_t b[USE_LDS ? 1 : KS];
    if (USE_LDS) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int c = tid + 256 * j;
            if (c < CT * NCH) {
                const int ctile = min(wi.y + c / NCH, last);
                bs[0][c] = yb[(size_t)ctile * NCH + (c % NCH)];
            }
        }
    } else {
#pragma unroll
        for (int s = 0; s < KS; ++s) b[s] = yb[((size_t)wi.y * KS + s) * 64 + lane];
    }
    __syncthreads();

    for (int ct = wi.y; ct < wi.z; ct += CT) {
        const int buf = ((ct - wi.y) / CT) & 1;
        const bool has_next = ct + CT < wi.z;
        float ycn_next[CT];
#pragma unroll
        for (int cc = 0; cc < CT; ++cc) ycn_next[cc] = ycn[min(ct + CT + cc, last) * 32 + (lane & 31)];
        bf16x8_t pre[USE_LDS ? PER : KS];
        if (USE_LDS) {
            if (has_next) {
#pragma unroll
                for (int j = 0; j < PER; ++j) {
                    const int c = tid + 256 * j;
                    if (c < CT * NCH) {
                        const int ctile = min(ct + CT + c / NCH, last);
                        pre[j] = yb[(size_t)ctile * NCH + (c % NCH)];
                    }
                }
            }
        } else {
            const int ctn = has_next ? ct + 1 : ct;
#pragma unroll
            for (int s = 0; s < KS; ++s) pre[s] = yb[((size_t)ctn * KS + s) * 64 + lane];
        }

        f32x16_t acc[CT][RT];
#pragma unroll
        for (int cc = 0; cc < CT; ++cc)
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int q = 0; q < 16; ++q) acc[cc][r][q] = xr[r][q] + yc[cc];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
#pragma unroll
            for (int cc = 0; cc < CT; ++cc) {
                const bf16x8_t bb = USE_LDS ? bs[buf][cc * NCH + s * 64 + lane] : b[USE_LDS ? 0 : s];
#pragma unroll
                for (int r = 0; r < RT; ++r)
                    acc[cc][r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[r][s], bb, acc[cc][r], 0, 0, 0);
            }
        }
        unsigned o = 0;
#pragma unroll
        for (int cc = 0; cc < CT; ++cc)
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int q = 0; q < 16; ++q) o |= __float_as_uint(acc[cc][r][q]);

        if (__any((int)o < 0)) {
            // Rare path: sign bits -> one mask per lane and column tile, then a
            // rolled loop with one ballot per accumulator register, staged
            // through LDS (one LDS atomic per ballot).  The mask's shifts
            // depend on an opaque zero, so the compiler cannot speculate them
            // into every tile of the hot loop.
            const int rowb = rt0 * 32 + 4 * (lane >> 5);
            int opaque;
            asm volatile("s_mov_b32 %0, 0" : "=s"(opaque));
#pragma unroll
            for (int cc = 0; cc < CT; ++cc) {
                unsigned long long mask = 0;
#pragma unroll
                for (int r = 0; r < RT; ++r)
#pragma unroll
                    for (int q = 0; q < 16; ++q)
                        mask |= (unsigned long long)(__float_as_uint(acc[cc][r][q]) >> 31) << (r * 16 + q + opaque);
                const int col = (ct + cc) * 32 + (lane & 31);
                const int lim = (ct + cc < wi.z && col < n_cols)
                                    ? (self_join ? min(n_rows, col) : n_rows) : 0;
                // only the registers with a sign bit set in some lane are visited
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
                    const int row = rowb + (rq >> 4) * 32 + 8 * (q >> 2) + (q & 3);
                    const bool p = ((mask >> rq) & 1ull) && row < lim;
                    const unsigned long long m = __ballot(p);
                    if (m) {
                        unsigned slot = 0;
                        if (lane == 0) slot = atomicAdd(&scount, (unsigned)__popcll(m));
                        slot = __shfl(slot, 0, 64) +
                               __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                        if (p) {
                            if (slot < LSTAGE) {
                                stage[slot] = make_int2(row, col);
                            } else {
                                const unsigned long long g = atomicAdd(counter, 1ull);
                                if (g < cap) cand[g] = make_int2(row, col);
                            }
                        }
                    }
                }
            }
        }
        if (USE_LDS) {
            if (has_next) {
#pragma unroll
                for (int j = 0; j < PER; ++j) {
                    const int c = tid + 256 * j;
                    if (c < CT * NCH) bs[buf ^ 1][c] = pre[j];
                }
            }
            __syncthreads();
        } else {
#pragma unroll
            for (int s = 0; s < KS; ++s) b[USE_LDS ? 0 : s] = pre[s];
        }
#pragma unroll
        for (int cc = 0; cc < CT; ++cc) yc[cc] = ycn_next[cc];
    }
    __syncthreads();
    const unsigned nst = min(scount, (unsigned)LSTAGE);
    if (tid == 0) sbase = nst ? atomicAdd(counter, (unsigned long long)nst) : 0ull;
    __syncthreads();
    for (unsigned i = tid; i < nst; i += 256)
        if (sbase + i < cap) cand[sbase + i] = stage[i];
}

// Group variant: the workgroup loads G column tiles per step straight from
// global memory into LDS (global_load_lds_dwordx4: no staging registers),
// double buffered, one barrier per G tiles.  Each wave holds its two row
// tiles' fragments in registers and runs the G tiles from LDS.  The two
// buffers are distinct LDS objects and the loop is unrolled by two, so the
// compiler can see that reads of one buffer never alias the asynchronous
// loads into the other and does not wait for those loads before reading.
template <int KS, int G>
__device__ __forceinline__ void filter_g_load(const bf16x8_t* __restrict__ yb, bf16x8_t* dst, int ct0,
                                              int last, int wave, int lane) {
    constexpr int NJ = G * KS;
#pragma unroll
    for (int jj = 0; jj < (NJ + 3) / 4; ++jj) {
        const int j = jj * 4 + wave;
        if (j >= NJ) break;
        const int tile = min(ct0 + j / KS, last);
        const bf16x8_t* src = yb + ((size_t)tile * KS + (j % KS)) * 64 + lane;
        __builtin_amdgcn_global_load_lds((const void*)src,
                                         (__attribute__((address_space(3))) void*)(dst + j * 64), 16, 0, 0);
    }
}

template <int KS, int G, int RT>
__device__ __forceinline__ void filter_g_compute(
    const bf16x8_t* bsrc, const bf16x8_t (&a)[RT][KS], const float (&xr)[RT][16], const float* __restrict__ ycn,
    int ct, int zend, int lane, int rowb, int n_rows, int n_cols, int self_join, int2* stage, unsigned* scount,
    int2* __restrict__ cand, unsigned long long cap, unsigned long long* __restrict__ counter) {
    float yc[G];
#pragma unroll
    for (int cc = 0; cc < G; ++cc) yc[cc] = ycn[min(ct + cc, zend - 1) * 32 + (lane & 31)];
#pragma unroll
    for (int cc = 0; cc < G; ++cc) {
        f32x16_t acc[RT];
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[r][q] = xr[r][q] + yc[cc];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const bf16x8_t bb = bsrc[(cc * KS + s) * 64 + lane];
#pragma unroll
            for (int r = 0; r < RT; ++r)
                acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[r][s], bb, acc[r], 0, 0, 0);
        }
        unsigned o = 0;
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int q = 0; q < 16; ++q) o |= __float_as_uint(acc[r][q]);
        if (__any((int)o < 0)) {
            // Rare path: the sign bits become one 32-bit mask per lane, then a
            // rolled loop emits one ballot per accumulator register, staged
            // through LDS.  The shifts depend on an opaque zero so that they
            // stay out of the hot loop.
            int opaque;
            asm volatile("s_mov_b32 %0, 0" : "=s"(opaque));
            unsigned mask = 0;
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int q = 0; q < 16; ++q) mask |= (__float_as_uint(acc[r][q]) >> 31) << (r * 16 + q + opaque);
            const int col = (ct + cc) * 32 + (lane & 31);
            const int lim = (ct + cc < zend && col < n_cols) ? (self_join ? min(n_rows, col) : n_rows) : 0;
            // only the registers with a sign bit set in some lane are visited
            unsigned any = mask;
#pragma unroll
            for (int o2 = 1; o2 < 64; o2 <<= 1) any |= __shfl_xor(any, o2, 64);
            any = __builtin_amdgcn_readfirstlane(any);
#pragma unroll 1
            while (any) {
                const int rq = __builtin_ctz(any);
                any &= any - 1;
                const int q = rq & 15;
                const int row = rowb + (rq >> 4) * 32 + 8 * (q >> 2) + (q & 3);
                const bool p = ((mask >> rq) & 1u) && row < lim;
                const unsigned long long m = __ballot(p);
                if (m) {
                    unsigned slot = 0;
                    if (lane == 0) slot = atomicAdd(scount, (unsigned)__popcll(m));
                    slot = __shfl(slot, 0, 64) +
                           __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                    if (p) {
                        if (slot < LSTAGE) {
                            stage[slot] = make_int2(row, col);
                        } else {
                            const unsigned long long g = atomicAdd(counter, 1ull);
                            if (g < cap) cand[g] = make_int2(row, col);
                        }
                    }
                }
            }
        }
    }
}

template <int KS, int G>
__global__ __launch_bounds__(256) void k_l2_mfma_filter_g(
    const bf16x8_t* __restrict__ xa, const bf16x8_t* __restrict__ yb, const float* __restrict__ xn,
    const float* __restrict__ ycn, const int4* __restrict__ work, int n_rows, int n_cols, int self_join,
    const MfmaParams* __restrict__ prm, int2* __restrict__ cand, unsigned long long cap,
    unsigned long long* __restrict__ counter) {
    constexpr int RT = 2;
    constexpr int NCH = KS * 64;       // fragments per column tile
    __shared__ bf16x8_t bsA[G * NCH];
    __shared__ bf16x8_t bsB[G * NCH];
    __shared__ int2 stage[LSTAGE];
    __shared__ unsigned scount;
    __shared__ unsigned long long sbase;
    int4 wi;
    if (!load_work_item(work, wi)) return;   // padding item: no memory is touched
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    if (tid == 0) scount = 0u;
    const int rt0 = wi.x * (4 * RT) + wave * RT;
    const int last = wi.z - 1;

    filter_g_load<KS, G>(yb, bsA, wi.y, last, wave, lane);
    bf16x8_t a[RT][KS];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int s = 0; s < KS; ++s) a[r][s] = xa[((size_t)(rt0 + r) * KS + s) * 64 + lane];
    float xr[RT][16];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int q = 0; q < 16; ++q) xr[r][q] = xn[(rt0 + r) * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3)];
    const int rowb = rt0 * 32 + 4 * (lane >> 5);

    int ct = wi.y;
    for (;;) {
        __builtin_amdgcn_s_waitcnt(0x3f70);   // vmcnt(0): this wave's share of the group landed
        __syncthreads();                        // every share landed; bsB is no longer read
        if (ct + G < wi.z) filter_g_load<KS, G>(yb, bsB, ct + G, last, wave, lane);
        filter_g_compute<KS, G, RT>(bsA, a, xr, ycn, ct, wi.z, lane, rowb, n_rows, n_cols, self_join, stage,
                                    &scount, cand, cap, counter);
        ct += G;
        if (ct >= wi.z) break;
        __builtin_amdgcn_s_waitcnt(0x3f70);
        __syncthreads();                        // bsA is no longer read
        if (ct + G < wi.z) filter_g_load<KS, G>(yb, bsA, ct + G, last, wave, lane);
        filter_g_compute<KS, G, RT>(bsB, a, xr, ycn, ct, wi.z, lane, rowb, n_rows, n_cols, self_join, stage,
                                    &scount, cand, cap, counter);
        ct += G;
        if (ct >= wi.z) break;
    }
    __syncthreads();
    const unsigned nst = min(scount, (unsigned)LSTAGE);
    if (tid == 0) sbase = nst ? atomicAdd(counter, (unsigned long long)nst) : 0ull;
    __syncthreads();
    for (unsigned i = tid; i < nst; i += 256)
        if (sbase + i < cap) cand[sbase + i] = stage[i];
}

// ------------------------------------------------------------------------
// FP32 recheck in the reference's order + FP64 band check
// ------------------------------------------------------------------------
// counters: [0] candidates, [1] edges, [2] band pairs, [3] band flips
// Each wave takes RC = 16 candidates per round.  The 32 rows they need are
// first all loaded into registers (two rows per 64-lane instruction when
// d % 4 == 0), then written to LDS rows padded to a multiple of 4 floats whose
// stride keeps the 16 computing lanes on distinct banks; then 16 lanes each
// run one candidate's sequential sum from LDS.  Rounds are uniform over the
// block (barriers), candidates past the end are masked.
constexpr int RC = 16;   // candidates per wave per round

__host__ __device__ constexpr int recheck_ld(int d) { return ((d + 3) & ~3) + 4; }

// ------------------------------------------------------------------------
// Integer data: exact int8 MFMA filter
// ------------------------------------------------------------------------
// When every coordinate is an integer and every dimension spans at most 255
// (o_k = the dimension's minimum), x' = x - o - 128 lies in [-128, 127] and
// the squared distance |x - y|^2 = h_x + h_y - 2 x'.y' (h = |x'|^2) is an
// exact integer below 2^24, which is also what the reference's FP32 sum
// gives; an edge is d^2 <= T' = floor(T(eps)).  Rows carry x' and columns
// -1 - y' = 127 - (y - o), both int8, so one v_mfma_i32_32x32x32_i8 step
// adds A.B = -x'.y' - sum(x') over 32 dimensions (half the steps of BF16).
// With u = h_x and v = h_y - T', d^2 <= T' <=> x'.y' >= ceil((u + v) / 2),
// and ceil((u + v) / 2) >= floor(u/2) + floor(v/2): the filter keeps every
// pair with x'.y' >= floor(u/2) + floor(v/2) (all edges, and pairs at most
// two units past the threshold): with r_x = sum(x') + floor(h_x / 2) - 1 and
// c_y = floor((h_y - T') / 2) the accumulator starts at r_x, and the pair is
// a candidate iff it ends below -c_y.  Every candidate is rechecked in the
// reference's FP32 order.

// Start terms of padding positions.  Real rows end at |acc| <= |r_x| + |A.B|
// <= 1.2e6 + 2.4e6 (d <= 144) and -c_y <= Tq / 2 + 1 <= 2^29 + 1, so a padding
// row (acc >= 2^30 - 2.4e6) or a padding column (-c_y = -2^28) is never a
// candidate, whatever eps: padding positions inside the cell path's padded
// order are not masked by the row/column limits, and 2^30 + A.B stays below
// 2^31.
constexpr int I8PAD_ROW = 1 << 30;
constexpr int I8PAD_COL = 1 << 28;

__device__ __forceinline__ int floor_half(int v) { return v >= 0 ? v / 2 : -((1 - v) / 2); }

// One workgroup per 32-point tile (padded position a holds point idmap[a],
// or a itself without a map): wave w writes k-steps s = w, w + 4, ...
// (32 dims each) of the tile's int8 fragments; lane l = 32 h + (a & 31) holds
// dims 32 s + 16 h .. +15 of point a, read as 64 contiguous bytes (float4
// loads when rows are 16-byte aligned) and stored as one 16-byte word at
// fragment index l (coalesced).  The row and column start terms come from the
// per-lane sums of x'^2 and x', reduced over waves and halves in LDS.
// Optional: hsq receives the exact |x'|^2 of every point and xrow its x' as
// point-major int8 rows of KS8 * 32 bytes (for the exact recheck).
__global__ __launch_bounds__(256) void k_prep_i8(const float* __restrict__ X, int n, int npad, int d, int KS8,
                                                 const int* __restrict__ idmap, int vec4,
                                                 const unsigned* __restrict__ colmin, int Tq,
                                                 i32x4_t* __restrict__ xa, i32x4_t* __restrict__ yb,
                                                 int* __restrict__ rterm, int* __restrict__ cterm,
                                                 int* __restrict__ hsq, i32x4_t* __restrict__ xrow, int l16) {
    __shared__ int off[160];      // o_k + 128 per dimension (d <= 160)
    __shared__ int sh[4][64], ssx[4][64];
    const int tile = blockIdx.x;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int a = tile * 32 + (lane & 31);
    const int half = lane >> 5;
    // padded position a holds point src (idmap: -1 = padding)
    const int src = idmap ? idmap[a] : (a < n ? a : -1);
    const bool real = src >= 0;
    for (int k = tid; k < d; k += 256) off[k] = (int)ord2f(colmin[k]) + 128;
    __syncthreads();
    int h = 0, sx = 0;
    for (int s = wave; s < KS8; s += 4) {
        const int k0 = 32 * s + 16 * half;
        float v[16];
        if (real && vec4 && k0 + 16 <= d) {
            const float4* row = reinterpret_cast<const float4*>(X + (size_t)src * d + k0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 f = row[q];
                v[4 * q] = f.x;
                v[4 * q + 1] = f.y;
                v[4 * q + 2] = f.z;
                v[4 * q + 3] = f.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = (real && k0 + j < d) ? X[(size_t)src * d + k0 + j] : 0.0f;
        }
        unsigned wa[4] = {0u, 0u, 0u, 0u}, wb[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int k = k0 + j;
            int xv = 0, yv = 0;
            if (real && k < d) {
                const int xp = (int)v[j] - off[k];
                xv = xp;
                yv = -1 - xp;
                h += xp * xp;
                sx += xp;
            }
            wa[j >> 2] |= (unsigned)(xv & 0xff) << (8 * (j & 3));
            wb[j >> 2] |= (unsigned)(yv & 0xff) << (8 * (j & 3));
        }
        // chunk c = 2 s + half (dims 16 c .. 16 c + 15) of tile point p = lane & 31
        const int c = 2 * s + half, p = lane & 31;
        const int word = l16 ? (((p >> 4) * (KS8 / 2) + (c >> 2)) * 64 + 16 * (c & 3) + (p & 15)) : s * 64 + lane;
        const size_t idx = (size_t)tile * KS8 * 64 + word;
        if (xa) xa[idx] = i32x4_t{(int)wa[0], (int)wa[1], (int)wa[2], (int)wa[3]};
        if (yb) yb[idx] = i32x4_t{(int)wb[0], (int)wb[1], (int)wb[2], (int)wb[3]};
        if (xrow) xrow[(size_t)a * (2 * KS8) + 2 * s + half] = i32x4_t{(int)wa[0], (int)wa[1], (int)wa[2], (int)wa[3]};
    }
    sh[wave][lane] = h;
    ssx[wave][lane] = sx;
    __syncthreads();
    if (tid < 32 && a < npad) {   // here a = tile * 32 + tid
        int ht = 0, st = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            ht += sh[w][tid] + sh[w][tid + 32];
            st += ssx[w][tid] + ssx[w][tid + 32];
        }
        if (rterm) rterm[a] = real ? st + (ht >> 1) - 1 : I8PAD_ROW;
        if (cterm) cterm[a] = real ? floor_half(ht - Tq) : I8PAD_COL;
        if (hsq) hsq[a] = real ? ht : 0;
    }
}

// Exact recheck of int8-filter candidates from point-major int8 rows of x'
// (KS8 * 32 bytes per point, 128 B at d = 128 instead of 512 B of FP32):
//   d^2 = h_r + h_c - 2 x'_r.x'_c,
// the exact integer the reference's FP32 sum yields on such data (every
// partial sum is an integer below 2^24), so the edge test d^2 <= T, the
// weight sqrtf(d^2) and the FP64 band check (sqrt of the same integer) match
// the FP32 recheck bit for bit.  Eight lanes per candidate: lane k of a group
// reads 16-byte words k, k + 8, ... of both rows (one coalesced 128-byte
// segment per point), and the partial dot products are summed over the group.
__global__ __launch_bounds__(256) void k_recheck_i8(const i32x4_t* __restrict__ xr8, const i32x4_t* __restrict__ yr8,
                                                    const int* __restrict__ hr, const int* __restrict__ hc, int KS8,
                                                    const int2* __restrict__ cand, unsigned long long ccap, int Tq,
                                                    float eps, uint8_t* __restrict__ flag, float* __restrict__ cw,
                                                    unsigned long long* __restrict__ counters) {
    const unsigned long long nc = min(counters[0], ccap);
    const double band = 1e-5 * (double)eps;
    const int sub = threadIdx.x & 7;
    const int nw = 2 * KS8;   // 16-byte words per row
    for (unsigned long long base = ((unsigned long long)blockIdx.x * blockDim.x + threadIdx.x) >> 3; ;
         base += ((unsigned long long)gridDim.x * blockDim.x) >> 3) {
        // uniform over the wave: its 8 groups take candidates base .. base + 7
        if ((base & ~7ull) >= nc) break;
        const bool live = base < nc;
        int dot = 0;
        int2 rc = make_int2(0, 0);
        if (live) {
            rc = cand[base];
            const i32x4_t* pr = xr8 + (size_t)rc.x * nw;
            const i32x4_t* pc = yr8 + (size_t)rc.y * nw;
            for (int j = sub; j < nw; j += 8) {
                const i32x4_t u = pr[j];
                const i32x4_t v = pc[j];
                dot = __builtin_amdgcn_sdot4(u.x, v.x, dot, false);
                dot = __builtin_amdgcn_sdot4(u.y, v.y, dot, false);
                dot = __builtin_amdgcn_sdot4(u.z, v.z, dot, false);
                dot = __builtin_amdgcn_sdot4(u.w, v.w, dot, false);
            }
        }
        dot += __shfl_xor(dot, 1, 8);
        dot += __shfl_xor(dot, 2, 8);
        dot += __shfl_xor(dot, 4, 8);
        if (live && sub == 0) {
            const int d2 = hr[rc.x] + hc[rc.y] - 2 * dot;
            const bool edge = d2 <= Tq;
            const float w = sqrtf((float)d2);
            if (fabs((double)w - (double)eps) <= band) {
                atomicAdd(&counters[2], 1ull);
                const bool in64 = sqrt((double)d2) <= (double)eps;
                if (in64 != edge) atomicAdd(&counters[3], 1ull);
            }
            flag[base] = edge ? 1 : 0;
            cw[base] = w;
        }
    }
}

// Shared pieces of the filter kernels with row terms as the first MFMA's C
// operand.  A wave owns RT = 2 row tiles: their fragments (a) and start terms
// in the accumulator layout (xr) stay in registers for the whole work item;
// the accumulator ends at r_x + (products) and the pair is a candidate iff it
// is below -c_y (c_y: the column's start term).  A lane's 16 values of a tile
// share one column, so the test is on their minimum.
//   MfI8:   int8 fragments, exact int32 accumulation (integer data).
//   MfBF16: BF16 fragments, FP32 accumulation; the C0 bound of the start terms
//           covers partial sums up to |r_x| + sum |products| <= 3 maxN, and
//           the final comparison is exact.
struct MfI8 {
    using frag = i32x4_t;
    using accv = i32x16_t;
    using accv16 = i32x4_t;
    using term = int;
    static constexpr term kHigh = 0x7fffffff, kLow = -0x7fffffff - 1;
    __device__ static accv mfma(frag a, frag b, accv c) { return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0); }
    __device__ static accv16 mfma16(frag a, frag b, accv16 c) {
        return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
    }
    __device__ static term vmin(term a, term b) { return min(a, b); }
    // v < t; v - t stays inside int32 for these start terms
    __device__ static unsigned below(term v, term t) { return (unsigned)(v - t) >> 31; }
};
struct MfBF16 {
    using frag = bf16x8_t;
    using accv = f32x16_t;
    using accv16 = f32x4_t;
    using term = float;
    static constexpr term kHigh = 3.402823466e38f, kLow = -3.402823466e38f;
    __device__ static accv mfma(frag a, frag b, accv c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static accv16 mfma16(frag a, frag b, accv16 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static term vmin(term a, term b) { return fminf(a, b); }
    __device__ static unsigned below(term v, term t) { return v < t ? 1u : 0u; }
};

constexpr int MRT = 2;   // row tiles per wave (the ring kernel also runs 4)

template <class T, int KS, int M = MRT>
__device__ __forceinline__ void mf_load_rows(const typename T::frag* __restrict__ xa,
                                             const typename T::term* __restrict__ rterm, int rt0, int lane,
                                             typename T::frag (&a)[M][KS], typename T::accv (&xr)[M]) {
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
        for (int s = 0; s < KS; ++s) a[r][s] = xa[((size_t)(rt0 + r) * KS + s) * 64 + lane];
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
        for (int q = 0; q < 16; ++q) xr[r][q] = rterm[(rt0 + r) * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3)];
}

// one column tile held in LDS (b)
template <class T, int KS, int M = MRT>
__device__ __forceinline__ void mf_tile(const typename T::frag* b, const typename T::frag (&a)[M][KS],
                                        const typename T::accv (&xr)[M], int lane, typename T::accv (&acc)[M]) {
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const typename T::frag bb = b[s * 64 + lane];
#pragma unroll
        for (int r = 0; r < M; ++r) acc[r] = T::mfma(a[r][s], bb, s == 0 ? xr[r] : acc[r]);
    }
}

// Minimum of a lane's 16 M values: four independent chains of 4 M (the
// compiler fuses pairs into 3-input minimums), then combined, so the
// dependent depth after the last MFMA is a quarter of the values plus two.
template <class T, int M = MRT>
__device__ __forceinline__ typename T::term mf_min(const typename T::accv (&acc)[M]) {
    static_assert(M == 2 || M == 4, "four chains");
    constexpr int L = 4 * M;   // values per chain
    typename T::term c[4];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        c[h] = acc[(h * L) >> 4][(h * L) & 15];
#pragma unroll
        for (int q = 1; q < L; ++q) c[h] = T::vmin(c[h], acc[(h * L + q) >> 4][(h * L + q) & 15]);
    }
    return T::vmin(T::vmin(c[0], c[1]), T::vmin(c[2], c[3]));
}

// Where the candidate walk puts a wave's candidates of one round (m: the
// wave's ballot, p: this lane has one).
// A workgroup's LDS stage shared by its waves: one LDS reservation per round,
// past the stage one global reservation per round for the wave's overflow.
struct SharedSink {
    int2* stage;
    unsigned stage_cap;
    unsigned* scount;
    int2* __restrict__ cand;
    unsigned long long cap;
    unsigned long long* __restrict__ counter;
    __device__ __forceinline__ void put(bool p, unsigned long long m, int row, int col, int lane) {
        unsigned slot = 0;
        if (lane == 0) slot = atomicAdd(scount, (unsigned)__popcll(m));
        slot = __shfl(slot, 0, 64) +
               __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (p && slot < stage_cap) stage[slot] = make_int2(row, col);
        // past the stage (dense tiles, e.g. clusters in cell order): one global
        // reservation per round for the wave's overflowing candidates
        const bool ovf = p && slot >= stage_cap;
        const unsigned long long mo = __ballot(ovf);
        if (mo) {
            unsigned long long gb = 0;
            if (lane == 0) gb = atomicAdd(counter, (unsigned long long)__popcll(mo));
            gb = __shfl(gb, 0, 64);
            if (ovf) {
                const unsigned long long g =
                    gb + __builtin_amdgcn_mbcnt_hi((unsigned)(mo >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mo, 0u));
                if (g < cap) cand[g] = make_int2(row, col);
            }
        }
    }
};

// The lanes walk their own set bits of mask together (value i of a lane is
// row rowb + RS * (i >> 2) + (i & 3)): one ballot per round, each round's
// candidates to the sink.
template <class mask_t, int RS, class Sink>
__device__ __forceinline__ void mf_walk(mask_t mask, int rowb, int col, int lane, Sink& sk) {
#pragma unroll 1
    for (;;) {
        const bool p = mask != 0;
        const unsigned long long m = __ballot(p);
        if (m == 0) break;
        int row = 0;
        if (p) {
            const int rq = sizeof(mask_t) == 8 ? __builtin_ctzll((unsigned long long)mask) : __builtin_ctz((unsigned)mask);
            mask &= mask - 1;
            row = rowb + RS * (rq >> 2) + (rq & 3);
        }
        sk.put(p, m, row, col, lane);
    }
}

// Rare path: candidates of one tile pair, staged in LDS.  Each lane builds
// the mask of its 16 M values below nyc (value i = r*16+q of a lane is row
// rowb + 8*(i >> 2) + (i & 3), rising with i), drops the rows at or past its
// limit (the row/column limit of a self join, padding) as a bit count, then
// the lanes walk their own set bits together: one ballot and one LDS
// reservation per round.  The mask's shifts depend on an opaque zero so they
// stay out of the hot loop.
template <class T, int M = MRT, class Sink>
__device__ __forceinline__ void mf_emit(const typename T::accv (&acc)[M], typename T::term nyc, int ct, int rt0,
                                        int lane, int n_rows, int n_cols, int self_join, Sink& sk) {
    static_assert(M == 2 || M == 4, "32 or 64 values per lane");
    using mask_t = typename std::conditional<M == 4, unsigned long long, unsigned>::type;
    constexpr int NV = 16 * M;   // values per lane
    const int rowb = rt0 * 32 + 4 * (lane >> 5);
    int opaque;
    asm volatile("s_mov_b32 %0, 0" : "=s"(opaque));
    mask_t mask = 0;
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
        for (int q = 0; q < 16; ++q) mask |= (mask_t)T::below(acc[r][q], nyc) << (r * 16 + q + opaque);
    const int col = ct * 32 + (lane & 31);
    const int lim = col < n_cols ? (self_join ? min(n_rows, col) : n_rows) : 0;
    // values with row offset 8*(i >> 2) + (i & 3) < L are the first k
    const int L = lim - rowb;
    const int k = L <= 0 ? 0 : (L >= 2 * NV ? NV : 4 * (L >> 3) + min(4, L & 7));
    mask &= k >= NV ? ~(mask_t)0 : (((mask_t)1 << k) - (mask_t)1);
    mf_walk<mask_t, 8>(mask, rowb, col, lane, sk);
}

// 16x16 MFMA forms of the filter (v_mfma_i32_16x16x64_i8 and
// v_mfma_f32_16x16x32_bf16), on the 16-row fragment layout: a 32-point tile
// is two 16-point halves of KH = KS / 2 k-steps each; word (half hb, k-step
// s, lane l) holds point 16 hb + (l & 15) and its k-chunk l >> 4 of k-step s
// (16 int8 or 8 bf16 values; the same 16-byte words as the 32x32 layout,
// permuted).  A wave keeps its 64 rows as four 16-row blocks: a[r][s] and the
// row start terms in the 16x16 accumulator layout (lane l: rows 4 (l >> 4) +
// q of the block, column l & 15).  A 32-column tile runs as two 16-column
// halves with their own column terms; each half's test is made while the
// next half's MFMAs run.
template <class T, int KS, int M>
__device__ __forceinline__ void mf16_load_rows(const typename T::frag* __restrict__ xa,
                                               const typename T::term* __restrict__ rterm, int rt0, int lane,
                                               typename T::frag (&a)[2 * M][KS / 2],
                                               typename T::accv16 (&xr)[2 * M]) {
    constexpr int KH = KS / 2;
#pragma unroll
    for (int r = 0; r < 2 * M; ++r)
#pragma unroll
        for (int s = 0; s < KH; ++s) a[r][s] = xa[((size_t)(rt0 + (r >> 1)) * KS + (r & 1) * KH + s) * 64 + lane];
#pragma unroll
    for (int r = 0; r < 2 * M; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) xr[r][q] = rterm[rt0 * 32 + 16 * r + 4 * (lane >> 4) + q];
}

// one 16-column half whose column fragments bb (one per k-step) were read
// from LDS
template <class T, int KS, int M>
__device__ __forceinline__ void mf16_half(const typename T::frag (&bb)[KS / 2],
                                          const typename T::frag (&a)[2 * M][KS / 2],
                                          const typename T::accv16 (&xr)[2 * M], typename T::accv16 (&acc)[2 * M]) {
#pragma unroll
    for (int s = 0; s < KS / 2; ++s)
#pragma unroll
        for (int r = 0; r < 2 * M; ++r) acc[r] = T::mfma16(a[r][s], bb[s], s == 0 ? xr[r] : acc[r]);
}

// one 16-column half held in LDS (b: its first k-step), each k-step's
// fragment read just before its MFMAs
template <class T, int KS, int M>
__device__ __forceinline__ void mf16_half_lds(const typename T::frag* b, const typename T::frag (&a)[2 * M][KS / 2],
                                              const typename T::accv16 (&xr)[2 * M], int lane,
                                              typename T::accv16 (&acc)[2 * M]) {
#pragma unroll
    for (int s = 0; s < KS / 2; ++s) {
        const typename T::frag bb = b[s * 64 + lane];
#pragma unroll
        for (int r = 0; r < 2 * M; ++r) acc[r] = T::mfma16(a[r][s], bb, s == 0 ? xr[r] : acc[r]);
    }
}

// minimum of a lane's 8 M values (four chains)
template <class T, int M>
__device__ __forceinline__ typename T::term mf16_min(const typename T::accv16 (&acc)[2 * M]) {
    constexpr int NV = 8 * M, L = NV / 4;   // values, per chain
    typename T::term c[4];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        c[h] = acc[(h * L) >> 2][(h * L) & 3];
#pragma unroll
        for (int q = 1; q < L; ++q) c[h] = T::vmin(c[h], acc[(h * L + q) >> 2][(h * L + q) & 3]);
    }
    return T::vmin(T::vmin(c[0], c[1]), T::vmin(c[2], c[3]));
}

// Rare path of a half: value i = 4 r + q of a lane is row rowb + 16 (i >> 2) +
// (i & 3) (rising with i) of column col; the rows at or past the lane's limit
// are dropped as a bit count, as in mf_emit.
template <class T, int M, class Sink>
__device__ __forceinline__ void mf16_emit(const typename T::accv16 (&acc)[2 * M], typename T::term nyc, int col,
                                          int rt0, int lane, int n_rows, int n_cols, int self_join, Sink& sk) {
    static_assert(M >= 2 && M <= 4, "16 to 32 values per lane");
    constexpr int NV = 8 * M;
    const int rowb = rt0 * 32 + 4 * (lane >> 4);
    int opaque;
    asm volatile("s_mov_b32 %0, 0" : "=s"(opaque));
    unsigned mask = 0;
#pragma unroll
    for (int r = 0; r < 2 * M; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) mask |= T::below(acc[r][q], nyc) << (r * 4 + q + opaque);
    const int lim = col < n_cols ? (self_join ? min(n_rows, col) : n_rows) : 0;
    // values with row offset 16*(i >> 2) + (i & 3) < L are the first k
    const int L = lim - rowb;
    const int k = L <= 0 ? 0 : (L >= 4 * NV ? NV : 4 * (L >> 4) + min(4, L & 15));
    mask &= k >= 32 ? ~0u : (1u << k) - 1u;
    mf_walk<unsigned, 16>(mask, rowb, col, lane, sk);
}

// The int8 counterpart of k_l2_mfma_filter<KS, 2, true, 1>: same work list,
// tiling (4 waves x 2 row tiles, column tiles double-buffered in LDS through
// registers, one barrier per tile), candidate path and output.
template <int KS8>
__global__ __launch_bounds__(256, KS8 <= 4 ? 4 : 1) void k_l2_mfma_filter_i8(
    const i32x4_t* __restrict__ xa, const i32x4_t* __restrict__ yb, const int* __restrict__ rterm,
    const int* __restrict__ cterm, const int4* __restrict__ work, int n_rows, int n_cols, int self_join,
    int2* __restrict__ cand, unsigned long long cap, unsigned long long* __restrict__ counter) {
    constexpr int NCH = KS8 * 64;
    constexpr int PER = (NCH + 255) / 256;
    __shared__ i32x4_t bs[2][NCH];
    __shared__ int2 stage[LSTAGE];
    __shared__ unsigned scount;
    __shared__ unsigned long long sbase;
    int4 wi;
    if (!load_work_item(work, wi)) return;   // padding item: no memory is touched
    const int tid = threadIdx.x;
    if (tid == 0) scount = 0u;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const int rt0 = wi.x * (4 * MRT) + wave * MRT;
    i32x4_t a[MRT][KS8];
    i32x16_t xr[MRT];
    mf_load_rows<MfI8, KS8>(xa, rterm, rt0, lane, a, xr);
    const int last = wi.z - 1;
    int yc = cterm[wi.y * 32 + (lane & 31)];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int c = tid + 256 * j;
        if (c < NCH) bs[0][c] = yb[(size_t)wi.y * NCH + c];
    }
    // every prologue load done here, so the loop's own waits count only its
    // prefetches (vmcnt 0; expcnt and lgkmcnt left at their maxima)
    __builtin_amdgcn_s_waitcnt(0xF70);
    __syncthreads();

    for (int ct = wi.y; ct < wi.z; ++ct) {
        const int buf = (ct - wi.y) & 1;
        const bool has_next = ct + 1 < wi.z;
        const int ycn = cterm[min(ct + 1, last) * 32 + (lane & 31)];
        i32x4_t pre[PER];
        if (has_next) {
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int c = tid + 256 * j;
                if (c < NCH) pre[j] = yb[(size_t)(ct + 1) * NCH + c];
            }
        }
        i32x16_t acc[MRT];
        mf_tile<MfI8, KS8>(bs[buf], a, xr, lane, acc);
        if (__any(mf_min<MfI8>(acc) < -yc)) {
            SharedSink sk{stage, (unsigned)LSTAGE, &scount, cand, cap, counter};
            mf_emit<MfI8>(acc, -yc, ct, rt0, lane, n_rows, n_cols, self_join, sk);
        }
        if (has_next) {
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int c = tid + 256 * j;
                if (c < NCH) bs[buf ^ 1][c] = pre[j];
            }
        }
        __syncthreads();
        yc = ycn;
    }
    const unsigned nst = min(scount, (unsigned)LSTAGE);
    if (tid == 0) sbase = nst ? atomicAdd(counter, (unsigned long long)nst) : 0ull;
    __syncthreads();
    for (unsigned i = tid; i < nst; i += 256)
        if (sbase + i < cap) cand[sbase + i] = stage[i];
}

// Barrier of the direct-to-LDS ring: wait until at most n of this wave's
// vector-memory operations are outstanding (n <= 15, wave-uniform) and its LDS
// operations are done, then s_barrier.  A __syncthreads() would wait for every
// direct-to-LDS load in flight (its fence covers them), including the ones
// issued for the group after next; the LDS data the loop reads after this
// barrier are the stages, whose loads the vmcnt wait covers.  The empty asm
// statements keep the compiler from moving memory operations across it.
__device__ __forceinline__ void ring_barrier(int n) {
    switch (n) {
        case 0: __builtin_amdgcn_s_waitcnt(0x070); break;
        case 1: __builtin_amdgcn_s_waitcnt(0x071); break;
        case 2: __builtin_amdgcn_s_waitcnt(0x072); break;
        case 3: __builtin_amdgcn_s_waitcnt(0x073); break;
        case 4: __builtin_amdgcn_s_waitcnt(0x074); break;
        case 5: __builtin_amdgcn_s_waitcnt(0x075); break;
        case 6: __builtin_amdgcn_s_waitcnt(0x076); break;
        case 7: __builtin_amdgcn_s_waitcnt(0x077); break;
        case 8: __builtin_amdgcn_s_waitcnt(0x078); break;
        case 9: __builtin_amdgcn_s_waitcnt(0x079); break;
        case 10: __builtin_amdgcn_s_waitcnt(0x07a); break;
        case 11: __builtin_amdgcn_s_waitcnt(0x07b); break;
        case 12: __builtin_amdgcn_s_waitcnt(0x07c); break;
        case 13: __builtin_amdgcn_s_waitcnt(0x07d); break;
        case 14: __builtin_amdgcn_s_waitcnt(0x07e); break;
        default: __builtin_amdgcn_s_waitcnt(0x07f); break;
    }
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// The column tiles of a tile-list work item, held in registers for the
// whole item: lane l of r[k] holds ctl[first + 64 k + l] (items have at most
// 256 tiles).  tile_at(pos) reads the tile at list position pos (uniform
// across the wave) with one lane read, no memory access.
struct TileIdx {
    int r[4];
    int first;
};

__device__ __forceinline__ TileIdx load_tile_idx(const int* __restrict__ ctl, int first, int end, int lane) {
    TileIdx ti;
    ti.first = first;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int p = first + 64 * k + lane;
        ti.r[k] = p < end ? ctl[p] : 0;
    }
    return ti;
}

__device__ __forceinline__ int tile_at(const TileIdx& ti, int pos) {
    const int o = __builtin_amdgcn_readfirstlane(pos - ti.first);
    const int k = o >> 6;
    const int v = k == 0 ? ti.r[0] : k == 1 ? ti.r[1] : k == 2 ? ti.r[2] : ti.r[3];
    return __builtin_amdgcn_readlane(v, o & 63);
}

// This wave's share of the loads of one group (G column tiles from t0,
// clamped to the last tile): fragment chunks j = wave, wave + 4, ... straight
// into LDS (global_load_lds_dwordx4), and, from wave 3, the group's column
// start terms (global_load_lds_dword, two tiles per instruction, (G + 1) / 2
// instructions).
// With a tile list (TL), t0 and last are positions in the work item's list
// of column tiles, whose indices ti holds.
template <class T, int KS, int G, int NW = 4, bool TL = false>
__device__ __forceinline__ void mf_group_load(const typename T::frag* __restrict__ yb,
                                              const typename T::term* __restrict__ cterm, typename T::frag* dst,
                                              typename T::term* cdst, int t0, int last, int wave, int lane,
                                              const TileIdx& ti) {
    constexpr int NJ = G * KS;
#pragma unroll
    for (int jj = 0; jj < (NJ + NW - 1) / NW; ++jj) {
        const int j = jj * NW + wave;
        if (j >= NJ) break;
        const int pos = min(t0 + j / KS, last);
        const int tile = TL ? tile_at(ti, pos) : pos;
        const typename T::frag* src = yb + ((size_t)tile * KS + (j % KS)) * 64 + lane;
        __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)(dst + j * 64),
                                         16, 0, 0);
    }
    if (wave == NW - 1) {
#pragma unroll
        for (int c2 = 0; c2 < (G + 1) / 2; ++c2) {
            int tile;
            if constexpr (TL) {
                const int t_lo = tile_at(ti, min(t0 + 2 * c2, last));
                const int t_hi = tile_at(ti, min(t0 + 2 * c2 + 1, last));
                tile = (lane >> 5) ? t_hi : t_lo;
            } else {
                tile = min(t0 + 2 * c2 + (lane >> 5), last);
            }
            const typename T::term* src = cterm + tile * 32 + (lane & 31);
            __builtin_amdgcn_global_load_lds((const void*)src,
                                             (__attribute__((address_space(3))) void*)(cdst + 64 * c2), 4, 0, 0);
        }
    }
}

// Direct-to-LDS ring: column tiles go from global memory straight into a ring
// of S LDS stages of G tiles each (no staging registers), issued S - 1 groups
// ahead of the group being computed, so a group's loads have S - 1 groups'
// time to land; one barrier per group.  The stages are distinct LDS objects
// and the loop is unrolled by S, so the compiler sees that reads of one stage
// never alias the asynchronous loads into the others.
//
// Work items (row block, first, end, 0) cover column tiles first .. end - 1,
// or with a tile list (TL) the tiles ctl[first .. end - 1], at most 256 of
// them (held in registers for the item's whole run).  A persistent
// grid (PQ) takes items queue[0]++ (in list order) in each workgroup until
// the list is exhausted; otherwise workgroup b runs item b.
//
// L16: the 16x16 MFMA forms on the 16-row fragment layout (M = 2 row tiles
// per wave as four 16-row blocks; KS even).
template <class T, int KS, int G, int MINW, int S = 3, int NW = 4, bool TL = false, bool PQ = false, int M = MRT,
          bool L16 = false>
__global__ __launch_bounds__(64 * NW, MINW) void k_l2_mfma_filter_ring(
    const typename T::frag* __restrict__ xa, const typename T::frag* __restrict__ yb,
    const typename T::term* __restrict__ rterm, const typename T::term* __restrict__ cterm,
    const int4* __restrict__ work, int n_rows, int n_cols, int self_join, int2* __restrict__ cand,
    unsigned long long cap, unsigned long long* __restrict__ counter, const int* __restrict__ ctl,
    unsigned n_items, unsigned* __restrict__ queue) {
    using frag = typename T::frag;
    using term = typename T::term;
    using accv = typename T::accv;
    using accv16 = typename T::accv16;
    constexpr int NCH = KS * 64;
    constexpr int NJ = G * KS;
    constexpr int NC = 64 * ((G + 1) / 2);   // start terms per stage (two tiles per load)
    static_assert(G >= 1 && G <= 16 && (S == 2 || S == 3), "ring shape");
    static_assert(!L16 || (M >= 2 && M <= 4 && KS % 2 == 0), "16x16 forms: 2-4 row tiles per wave, even k-steps");
    // 16x16 forms: read a tile's fragments for both halves at once where the
    // registers allow it (int8 with two row tiles per wave, all-pairs lists)
    constexpr bool PRE = std::is_same<T, MfI8>::value && !TL && M == 2;
    __shared__ frag bs0[G * NCH], bs1[G * NCH], bs2[S == 3 ? G * NCH : 1];
    __shared__ term cs0[NC], cs1[NC], cs2[S == 3 ? NC : 1];
    // candidates staged in LDS, WS per wave (at least one round of 64); the
    // four-wave int8 shapes with four tiles per stage keep 512 in all so that
    // four workgroups fit a CU
    constexpr unsigned STG = (std::is_same<T, MfI8>::value && G >= 4 && NW <= 4) ? 512u : (unsigned)LSTAGE;
    constexpr unsigned WS = (STG / NW) / 64 * 64 < 64 ? 64u : (STG / NW) / 64 * 64;
    __shared__ int2 stage[NW * WS];
    __shared__ unsigned s_item;
    const int tid = threadIdx.x;
    // wave-uniform (scalar) in the tile-list forms, whose list positions
    // derive from it, and in the 16x16 forms (its load addresses then stay in
    // scalar registers)
    const int wave = TL || L16 ? __builtin_amdgcn_readfirstlane(tid >> 6) : tid >> 6;
    const int lane = tid & 63;
    unsigned item = blockIdx.x;
    for (;; ) {
    if constexpr (PQ) {
        if (tid == 0) s_item = atomicAdd(queue, 1u);
        __syncthreads();
        item = s_item;
        __syncthreads();   // every wave has read s_item before the next round writes it
        if (item >= n_items) break;
    }
    int4 wi;
    if (load_work_item(work, item, wi)) {   // padding items touch no memory
    WaveSink sk{stage + wave * WS, 0u, WS, cand, cap, counter};
    const int rt0 = wi.x * (NW * M) + wave * M;
    const int last = wi.z - 1;
    const int ng = (wi.z - wi.y + G - 1) / G;
    // loads this wave issues per group (the same in every lane)
    const int nld = __builtin_amdgcn_readfirstlane((wave < NJ ? (NJ - 1 - wave) / NW + 1 : 0) +
                                                   (wave == NW - 1 ? (G + 1) / 2 : 0));

    // issue order: rows, group 0, group 1 (the barrier's count relies on it)
    // the item's tile indices first (the group loads wait for them), then
    // the rows
    TileIdx ti{};
    if constexpr (TL) ti = load_tile_idx(ctl, wi.y, wi.z, lane);
    frag a[L16 ? 1 : M][KS];
    accv xr[L16 ? 1 : M];
    frag a16[L16 ? 2 * M : 1][L16 ? KS / 2 : 1];
    accv16 xr16[L16 ? 2 * M : 1];
    // L16: the half whose test is pending (initially none: no value is below kLow)
    accv16 accp[L16 ? 2 * M : 1];
    term nycp = T::kLow;
    int colp = 0;
    if constexpr (L16) {
        mf16_load_rows<T, KS, M>(xa, rterm, rt0, lane, a16, xr16);
#pragma unroll
        for (int r = 0; r < 2 * M; ++r) accp[r] = accv16{T::kHigh, T::kHigh, T::kHigh, T::kHigh};
    } else {
        mf_load_rows<T, KS, M>(xa, rterm, rt0, lane, a, xr);
    }
    asm volatile("" ::: "memory");
    mf_group_load<T, KS, G, NW, TL>(yb, cterm, bs0, cs0, wi.y, last, wave, lane, ti);
    asm volatile("" ::: "memory");
    if (S == 3 && ng > 1) mf_group_load<T, KS, G, NW, TL>(yb, cterm, bs1, cs1, wi.y + G, last, wave, lane, ti);
    // rows and group 0 landed (with three stages group 1 may still be in
    // flight)
    ring_barrier(S == 3 && ng > 1 ? nld : 0);

    // group g from stage cur; the loads of group g + S - 1 go to stage nxt
    // (group g - 1's, free since the last barrier); then group g + 1 must
    // have landed (with two stages: the loads just issued)
    auto step = [&](const frag* cur, const term* ccur, frag* nxt, term* cnxt, int g) {
        const int t0 = wi.y + g * G;
        const bool issue = g + S - 1 < ng;
        if (issue) mf_group_load<T, KS, G, NW, TL>(yb, cterm, nxt, cnxt, t0 + (S - 1) * G, last, wave, lane, ti);
#pragma unroll
        for (int cc = 0; cc < G; ++cc) {
            if (cc > 0 && t0 + cc > last) break;
            if constexpr (L16) {
                const int ct = TL ? tile_at(ti, t0 + cc) : t0 + cc;
                // with PRE both halves' column fragments are read before the
                // first half's MFMAs (the second half's reads land under
                // them); else each k-step's just before its MFMAs
                frag bb[PRE ? 2 : 1][KS / 2];
                if constexpr (PRE) {
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int s = 0; s < KS / 2; ++s) bb[h][s] = cur[cc * NCH + (h * (KS / 2) + s) * 64 + lane];
                }
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const term nyc = -ccur[cc * 32 + 16 * h + (lane & 15)];
                    accv16 acc[2 * M];
                    if constexpr (PRE) mf16_half<T, KS, M>(bb[h], a16, xr16, acc);
                    else mf16_half_lds<T, KS, M>(cur + cc * NCH + h * (KS / 2) * 64, a16, xr16, lane, acc);
                    // the previous half's test, under this half's MFMAs
                    if (__any(mf16_min<T, M>(accp) < nycp))
                        mf16_emit<T, M>(accp, nycp, colp, rt0, lane, n_rows, n_cols, self_join, sk);
#pragma unroll
                    for (int r = 0; r < 2 * M; ++r) accp[r] = acc[r];
                    nycp = nyc;
                    colp = ct * 32 + 16 * h + (lane & 15);
                }
            } else {
                const term nyc = -ccur[cc * 32 + (lane & 31)];
                accv acc[M];
                mf_tile<T, KS, M>(cur + cc * NCH, a, xr, lane, acc);
                if (__any(mf_min<T, M>(acc) < nyc)) {
                    const int ct = TL ? tile_at(ti, t0 + cc) : t0 + cc;
                    mf_emit<T, M>(acc, nyc, ct, rt0, lane, n_rows, n_cols, self_join, sk);
                }
            }
        }
        ring_barrier(S == 3 && issue ? nld : 0);
    };
    int g = 0;
    if constexpr (S == 3) {
        for (;;) {
            step(bs0, cs0, bs2, cs2, g);
            if (++g >= ng) break;
            step(bs1, cs1, bs0, cs0, g);
            if (++g >= ng) break;
            step(bs2, cs2, bs1, cs1, g);
            if (++g >= ng) break;
        }
    } else {
        for (;;) {
            step(bs0, cs0, bs1, cs1, g);
            if (++g >= ng) break;
            step(bs1, cs1, bs0, cs0, g);
            if (++g >= ng) break;
        }
    }
    if constexpr (L16) {   // the last half's test
        if (__any(mf16_min<T, M>(accp) < nycp))
            mf16_emit<T, M>(accp, nycp, colp, rt0, lane, n_rows, n_cols, self_join, sk);
    }
    sk.flush(lane);   // the wave's staged candidates
    }   // item
    if constexpr (!PQ) break;
    }   // items of a persistent workgroup
}

// Output per candidate c: flag[c] = 1 if it is an edge, w[c] = its float
// distance; a stable device select then compacts the edges (no atomics).
template <bool VEC4>
__global__ __launch_bounds__(256) void k_l2_recheck(
    const float* __restrict__ X, const float* __restrict__ Y, int d,
    const int2* __restrict__ cand, unsigned long long ccap, float T, float eps,
    uint8_t* __restrict__ flag, float* __restrict__ cw,
    unsigned long long* __restrict__ counters) {
#pragma clang fp contract(off)
    extern __shared__ float rows_all[];   // [4 waves][2*RC][ld]
    const unsigned long long nc = min(counters[0], ccap);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ld = recheck_ld(d);
    float* rows = rows_all + (size_t)wave * (2 * RC) * ld;
    const double band = 1e-5 * (double)eps;
    for (unsigned long long base = (unsigned long long)blockIdx.x * (4 * RC); base < nc;
         base += (unsigned long long)gridDim.x * (4 * RC)) {
        const unsigned long long c0 = base + (unsigned long long)wave * RC;
        int2 p = make_int2(-1, -1);
        if (lane < RC && c0 + lane < nc) p = cand[c0 + lane];
        if (VEC4 && d <= 128) {
            // row r = 2*i + (lane >> 5), float4 column lane & 31
            const int half = lane >> 5, k4 = lane & 31;
            const int nk4 = d >> 2;
            float4 v[RC];
#pragma unroll
            for (int i = 0; i < RC; ++i) {
                const int r = 2 * i + half;
                const int srcl = r < RC ? r : r - RC;
                const int id = __shfl(r < RC ? p.x : p.y, srcl, 64);
                v[i] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (id >= 0 && k4 < nk4) v[i] = *(const float4*)((r < RC ? X : Y) + (size_t)id * d + 4 * k4);
            }
#pragma unroll
            for (int i = 0; i < RC; ++i) {
                const int r = 2 * i + half;
                if (k4 < nk4) *(float4*)(rows + r * ld + 4 * k4) = v[i];
            }
        } else {
            for (int r = 0; r < 2 * RC; ++r) {
                const int srcl = r < RC ? r : r - RC;
                const int id = __shfl(r < RC ? p.x : p.y, srcl, 64);
                if (id < 0) continue;
                const float* src = (r < RC ? X : Y) + (size_t)id * d;
                for (int k = lane; k < d; k += 64) rows[r * ld + k] = src[k];
            }
        }
        __syncthreads();
        bool edge = false;
        float w = 0.0f;
        if (lane < RC && p.x >= 0) {
            const float* x = rows + lane * ld;
            const float* y = rows + (RC + lane) * ld;
            float acc = 0.0f;
            int k = 0;
            for (; k + 4 <= d; k += 4) {
                const float4 xv = *(const float4*)(x + k);
                const float4 yv = *(const float4*)(y + k);
                float t;
                t = xv.x - yv.x; acc = acc + t * t;
                t = xv.y - yv.y; acc = acc + t * t;
                t = xv.z - yv.z; acc = acc + t * t;
                t = xv.w - yv.w; acc = acc + t * t;
            }
            for (; k < d; ++k) {
                const float t = x[k] - y[k];
                acc = acc + t * t;
            }
            edge = acc <= T;
            w = sqrtf(acc);
            if (fabs((double)w - (double)eps) <= band) {
                atomicAdd(&counters[2], 1ull);
                double s2 = 0.0;
                for (int kk = 0; kk < d; ++kk) {
                    const double t = (double)x[kk] - (double)y[kk];
                    s2 += t * t;
                }
                const bool in64 = sqrt(s2) <= (double)eps;
                if (in64 != edge) atomicAdd(&counters[3], 1ull);
            }
        }
        if (lane < RC && c0 + lane < nc) {
            flag[c0 + lane] = edge ? 1 : 0;
            cw[c0 + lane] = w;
        }
        __syncthreads();
    }
}

__global__ void k_pack_keys(const int2* __restrict__ ij, uint64_t* __restrict__ keys, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) keys[e] = ((uint64_t)(uint32_t)ij[e].x << 32) | (uint32_t)ij[e].y;
}

// Directed form of an undirected edge list: (i, j) and (j, i) per edge and a
// zero-weight self loop per vertex, as sort keys (u << 32 | v).
__global__ void k_directed_keys(const int2* __restrict__ ij, const float* __restrict__ w, int64_t E, int64_t n,
                                uint64_t* __restrict__ keys, float* __restrict__ ww) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 2 * E) {
        const int2 e = ij[t >> 1];
        const uint32_t a = (uint32_t)((t & 1) ? e.y : e.x), b = (uint32_t)((t & 1) ? e.x : e.y);
        keys[t] = ((uint64_t)a << 32) | b;
        ww[t] = w[t >> 1];
    } else if (t < 2 * E + n) {
        const uint32_t v = (uint32_t)(t - 2 * E);
        keys[t] = ((uint64_t)v << 32) | v;
        ww[t] = 0.0f;
    }
}

// rowptr[r] = first position of row r in the sorted keys (r = 0..n), col = low words
__global__ void k_csr_from_keys(const uint64_t* __restrict__ keys, int64_t total, int64_t n,
                                int64_t* __restrict__ rowptr, int* __restrict__ col) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e > total) return;
    const int64_t r = e < total ? (int64_t)(keys[e] >> 32) : n;
    const int64_t pr = e > 0 ? (int64_t)(keys[e - 1] >> 32) : -1;
    for (int64_t x = pr + 1; x <= r; ++x) rowptr[x] = e;
    if (e < total) col[e] = (int)(keys[e] & 0xffffffffu);
}

__global__ void k_unpack_keys(const uint64_t* __restrict__ keys, int2* __restrict__ ij, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) ij[e] = make_int2((int)(keys[e] >> 32), (int)(keys[e] & 0xffffffffu));
}

// sqrtf here is __builtin_sqrtf, which gfx950 code generation expands into
// v_sqrt_f32 plus an FMA residual correction: the correctly rounded square
// root the reference gets from std::sqrt(float).  (__fsqrt_rn lowers to a bare
// scaled v_sqrt_f32, which can be off by one ulp.)
__global__ void k_debug_sqrt(const float* __restrict__ in, float* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = sqrtf(in[i]);
}

// Plain device copy of known size, used to calibrate the HBM byte counters
// (FETCH_SIZE / WRITE_SIZE) against a kernel whose traffic is exactly known.
__global__ __launch_bounds__(256) void k_debug_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                    int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------------------------
// Debug: MFMA output layout
// ------------------------------------------------------------------------
__global__ void k_mfma_layout(int* rows, int* cols) {
    const int lane = threadIdx.x;
    bf16x8_t a, b;
    // A[m][k]: k==0 -> m+1, k==1 -> 1 ; B[k][n]: k==0 -> 1, k==1 -> 64*(n+1)
    // D[m][n] = (m+1) + 64*(n+1), decodable with exact small integers.
    const int m = lane & 31, kb = (lane >> 5) * 8;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = kb + e;
        a[e] = (__bf16)(k == 0 ? (float)(m + 1) : (k == 1 ? 1.0f : 0.0f));
        b[e] = (__bf16)(k == 0 ? 1.0f : (k == 1 ? 64.0f * (float)(m + 1) : 0.0f));
    }
    f32x16_t c;
#pragma unroll
    for (int r = 0; r < 16; ++r) c[r] = 0.0f;
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int v = (int)c[r];
        rows[lane * 16 + r] = (v % 64) - 1;
        cols[lane * 16 + r] = (v / 64) - 1;
    }
}

// ------------------------------------------------------------------------
// Host orchestration
// ------------------------------------------------------------------------
struct Workspace {
    DevBuf counters;   // u64[8]
    DevBuf cand;       // int2
    DevBuf cflag;      // uint8 per candidate: recheck says edge
    DevBuf cw;         // float per candidate: its distance
    unsigned long long cand_cap = 0;
    DevBuf xa, yb;     // fragments
    DevBuf xn, yn, ycn;   // row norms, column norms, column norms - C0
    DevBuf hx, hy;        // exact |x'|^2 per row / column point (int8 path)
    DevBuf xr8, yr8;      // x' as point-major int8 rows (int8 path recheck)
    DevBuf colsum, colmin, colmax, stat, range, mu, params;
    DevBuf work;
    std::vector<int4> work_host;
    int work_key[5] = {-1, -1, -1, -1, -1};
};

// Filter variants: {row tiles per wave, stage column tiles in LDS, column
// tiles per step, waves per workgroup}.
struct FilterVariant {
    int rt;
    bool lds;
    int ct;
    int nw;
};
// Variants 7, 8, 9 and 10 are the group kernel (direct-to-LDS loads) with G = 1..4;
// 11 and 12 the direct-to-LDS ring (three stages of 1 or 2 tiles), 13 and 14
// the ring with two stages of 2 or 3 tiles; 15 two stages of 4 tiles shared by
// 8 waves, 16 and 17 two stages of 4 or 8 tiles shared by 12 waves; 18 and 19
// the 16x16 MFMA forms of 17 and 13 (16-row fragment layout).
constexpr FilterVariant kVariants[] = {{2, false, 1, 4}, {2, true, 1, 4}, {4, true, 1, 4}, {3, true, 1, 4},
                                       {2, true, 2, 4},  {1, true, 4, 4}, {1, true, 2, 4},
                                       {2, true, 1, 4},  {2, true, 2, 4}, {2, true, 3, 4}, {2, true, 4, 4},
                                       {2, true, 1, 4},  {2, true, 2, 4}, {2, true, 2, 4}, {2, true, 3, 4},
                                       {2, true, 4, 8},  {2, true, 4, 12}, {2, true, 8, 12},
                                       {2, true, 8, 12}, {2, true, 2, 4}};
constexpr int kNumVariants = 20;
int g_variant = -1;   // -1: from RNNG_L2_VARIANT or the default

int filter_variant() {
    if (g_variant < 0) {
        g_variant = 19;
        if (const char* e = getenv("RNNG_L2_VARIANT")) {
            const int v = atoi(e);
            if (v >= 0 && v < kNumVariants) g_variant = v;
        }
    }
    return g_variant;
}

std::mutex g_ws_mu;
Workspace g_ws;
int64_t g_cand_hint = 0;   // initial candidate-buffer size (0: default)


// Row blocks of the kernel's row tiles x column chunks of chunk_tiles() tiles,
// chunk-major so that the workgroups resident at one time share column
// fragments in L2 (RNNG_L2_CHUNK overrides the default of 192).
int g_chunk = 0;

int chunk_tiles() {
    if (g_chunk == 0) {
        g_chunk = 192;
        if (const char* e = getenv("RNNG_L2_CHUNK")) {
            const int v = atoi(e);
            if (v >= 32 && v <= 4096) g_chunk = v;
        }
    }
    return g_chunk;
}

// Workgroups are dispatched to the 8 XCDs round-robin by block index, and each
// XCD has its own L2.  With xcd_remap the list is permuted so that XCD x runs
// the contiguous x-th eighth of the chunk-major list: the column chunk that
// its resident workgroups share is then fetched into one L2 instead of eight.
int g_xcd_remap = -1;

int xcd_remap() {
    if (g_xcd_remap < 0) {
        g_xcd_remap = 1;
        if (const char* e = getenv("RNNG_L2_XCD")) g_xcd_remap = atoi(e) ? 1 : 0;
    }
    return g_xcd_remap;
}

int build_work(Workspace& ws, int n_rowblocks, int rb_tiles, int n_coltiles, int self_join) {
    const int remap = xcd_remap();
    const int CH = chunk_tiles();
    if (ws.work_key[0] == n_rowblocks && ws.work_key[1] == n_coltiles &&
        ws.work_key[2] == self_join && ws.work_key[3] == CH * 2 + remap && ws.work_key[4] == rb_tiles)
        return RNNG_OK;
    std::vector<int4> list;
    const int nchunks = (n_coltiles + CH - 1) / CH;
    for (int c = 0; c < nchunks; ++c) {
        const int c0 = c * CH, c1 = std::min(n_coltiles, c0 + CH);
        for (int rb = 0; rb < n_rowblocks; ++rb) {
            int lo = c0;
            if (self_join) lo = std::max(lo, rb * rb_tiles);
            if (lo >= c1) continue;
            list.push_back(make_int4(rb, lo, c1, 0));
        }
    }
    if (remap) xcd_permute(list);
    ws.work_host.swap(list);
    RNNG_TRY(ws.work.reserve(std::max<size_t>(1, ws.work_host.size()) * sizeof(int4)));
    RNNG_HIP_TRY(hipMemcpyAsync(ws.work.ptr, ws.work_host.data(), ws.work_host.size() * sizeof(int4),
                                hipMemcpyHostToDevice, stream()));
    ws.work_key[0] = n_rowblocks;
    ws.work_key[1] = n_coltiles;
    ws.work_key[2] = self_join;
    ws.work_key[3] = CH * 2 + remap;
    ws.work_key[4] = rb_tiles;
    return RNNG_OK;
}

template <int KS, int RT, bool LDS, int CT = 1>
void launch_filter(const FilterArgs& a) {
    k_l2_mfma_filter<KS, RT, LDS, CT><<<a.n_items, 256, 0, stream()>>>(
        a.xa, a.yb, a.xn, a.ycn, a.work, a.n_rows, a.n_cols, a.self_join, a.prm, a.cand, a.cap, a.counter);
}

template <int KS, int G>
void launch_filter_g(const FilterArgs& a) {
    k_l2_mfma_filter_g<KS, G><<<a.n_items, 256, 0, stream()>>>(
        a.xa, a.yb, a.xn, a.ycn, a.work, a.n_rows, a.n_cols, a.self_join, a.prm, a.cand, a.cap, a.counter);
}

template <int KS, int G, int S = 3, int NW = 4, bool L16 = false>
void launch_filter_ring(const FilterArgs& a) {
    k_l2_mfma_filter_ring<MfBF16, KS, G, 1, S, NW, false, false, MRT, L16><<<a.n_items, 64 * NW, 0, stream()>>>(
        a.xa, a.yb, a.xn, a.ycn, a.work, a.n_rows, a.n_cols, a.self_join, a.cand, a.cap, a.counter, nullptr,
        a.n_items, nullptr);
}

template <int KS>
void launch_filter_v(const FilterArgs& a, int variant) {
    if constexpr (KS % 2 == 0) {
        if (variant == 18) return launch_filter_ring<KS, (KS >= 9 ? 6 : 8), 2, 12, true>(a);
        if (variant == 19) return launch_filter_ring<KS, 2, 2, 4, true>(a);
    }
    switch (variant) {
        case 11: launch_filter_ring<KS, 1>(a); return;
        case 12: launch_filter_ring<KS, 2>(a); return;
        case 13: launch_filter_ring<KS, 2, 2>(a); return;
        case 14: launch_filter_ring<KS, 3, 2>(a); return;
        case 15: launch_filter_ring<KS, 4, 2, 8>(a); return;
        case 16: launch_filter_ring<KS, 4, 2, 12>(a); return;
        case 17: launch_filter_ring<KS, (KS >= 9 ? 6 : 8), 2, 12>(a); return;
        case 7: launch_filter_g<KS, 1>(a); return;
        case 8: launch_filter_g<KS, 2>(a); return;
        case 9: launch_filter_g<KS, 3>(a); return;
        case 10: launch_filter_g<KS, 4>(a); return;
        default: break;
    }
    switch (variant) {
        case 0: launch_filter<KS, 2, false>(a); break;
        case 1: launch_filter<KS, 2, true>(a); break;
        case 3: launch_filter<KS, 3, true>(a); break;
        case 4: launch_filter<KS, 2, true, 2>(a); break;
        case 5: launch_filter<KS, 1, true, 4>(a); break;
        case 6: launch_filter<KS, 1, true, 2>(a); break;
        default: launch_filter<KS, 4, true>(a); break;
    }
}

bool g_int8_filter = true;   // integer data may use the exact int8 filter

float ord2f_host(unsigned u) {
    const unsigned v = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &v, sizeof(f));
    return f;
}

// int8 filter kernels by code: 0 = k_l2_mfma_filter_i8 (column tiles staged
// through registers); 1, 2 = the direct-to-LDS ring with 1 or 2 tiles per
// stage and three stages; 3, 4 = 3 or 4 tiles per stage and two stages; 5, 6,
// 7 = two stages of 4, 6 or 8 tiles shared by eight waves (512 threads,
// 16 row tiles per work item; 6 tiles at KS8 = 5 for two workgroups per CU);
// 8 = four waves of four row tiles; 9, 10 = the 16x16 forms of 7 and 3 (the
// default: code 9); 11, 12 = 16x16 forms with four row tiles per wave (four
// waves, two per SIMD) and three (twelve waves, three per SIMD).
template <int KS8, int CODE>
void launch_filter_i8(const I8Args& a) {
    constexpr int MINW = KS8 <= 4 ? 4 : 1;
    if constexpr (CODE == 0)
        k_l2_mfma_filter_i8<KS8><<<a.n_items, 256, 0, stream()>>>(a.xa, a.yb, a.rterm, a.cterm, a.work, a.n_rows,
                                                                  a.n_cols, a.self_join, a.cand, a.cap, a.counter);
    else if constexpr (CODE <= 2)
        k_l2_mfma_filter_ring<MfI8, KS8, CODE, MINW, 3><<<a.n_items, 256, 0, stream()>>>(
            a.xa, a.yb, a.rterm, a.cterm, a.work, a.n_rows, a.n_cols, a.self_join, a.cand, a.cap, a.counter,
            nullptr, a.n_items, nullptr);
    else if constexpr (CODE <= 4)
        k_l2_mfma_filter_ring<MfI8, KS8, CODE, MINW, 2><<<a.n_items, 256, 0, stream()>>>(
            a.xa, a.yb, a.rterm, a.cterm, a.work, a.n_rows, a.n_cols, a.self_join, a.cand, a.cap, a.counter,
            nullptr, a.n_items, nullptr);
    else if constexpr (CODE <= 7) {
        constexpr int G = CODE == 5 ? 4 : (CODE == 6 || KS8 >= 5) ? 6 : 8;
        k_l2_mfma_filter_ring<MfI8, KS8, G, MINW, 2, 8><<<a.n_items, 512, 0, stream()>>>(
            a.xa, a.yb, a.rterm, a.cterm, a.work, a.n_rows, a.n_cols, a.self_join, a.cand, a.cap, a.counter,
            nullptr, a.n_items, nullptr);
    } else if constexpr (CODE == 8) {
        // four waves of four row tiles each (the same 16 row tiles per item),
        // two waves per SIMD with twice the registers
        constexpr int G = KS8 >= 5 ? 6 : 8;
        k_l2_mfma_filter_ring<MfI8, KS8, G, 2, 2, 4, false, false, 4><<<a.n_items, 256, 0, stream()>>>(
            a.xa, a.yb, a.rterm, a.cterm, a.work, a.n_rows, a.n_cols, a.self_join, a.cand, a.cap, a.counter,
            nullptr, a.n_items, nullptr);
    } else if constexpr (KS8 % 2 == 0) {
        if constexpr (CODE == 9) {   // 16x16 form of code 7
            constexpr int G = KS8 >= 5 ? 6 : 8;
            k_l2_mfma_filter_ring<MfI8, KS8, G, MINW, 2, 8, false, false, MRT, true><<<a.n_items, 512, 0, stream()>>>(
                a.xa, a.yb, a.rterm, a.cterm, a.work, a.n_rows, a.n_cols, a.self_join, a.cand, a.cap, a.counter,
                nullptr, a.n_items, nullptr);
        } else if constexpr (CODE == 10) {   // 16x16 form of code 3
            k_l2_mfma_filter_ring<MfI8, KS8, 3, MINW, 2, 4, false, false, MRT, true><<<a.n_items, 256, 0, stream()>>>(
                a.xa, a.yb, a.rterm, a.cterm, a.work, a.n_rows, a.n_cols, a.self_join, a.cand, a.cap, a.counter,
                nullptr, a.n_items, nullptr);
        } else if constexpr (CODE == 11) {   // 16x16, four row tiles per wave, four waves (two per SIMD)
            k_l2_mfma_filter_ring<MfI8, KS8, 8, 2, 2, 4, false, false, 4, true><<<a.n_items, 256, 0, stream()>>>(
                a.xa, a.yb, a.rterm, a.cterm, a.work, a.n_rows, a.n_cols, a.self_join, a.cand, a.cap, a.counter,
                nullptr, a.n_items, nullptr);
        } else {                             // 16x16, three row tiles per wave, twelve waves (three per SIMD)
            k_l2_mfma_filter_ring<MfI8, KS8, 8, 3, 2, 12, false, false, 3, true><<<a.n_items, 768, 0, stream()>>>(
                a.xa, a.yb, a.rterm, a.cterm, a.work, a.n_rows, a.n_cols, a.self_join, a.cand, a.cap, a.counter,
                nullptr, a.n_items, nullptr);
        }
    }
}

constexpr int kI8Codes = 13;

// row tiles per work item of an int8 kernel code (codes 5-9 and 11: 16 row
// tiles, 12: 36)
int i8_row_tiles(int code) { return code == 12 ? 36 : (code >= 5 && code <= 9) || code == 11 ? 16 : 8; }

int g_i8_group = -1;   // int8 filter kernel code (launch_filter_i8); -1: not read yet

int i8_group() {
    if (g_i8_group < 0) {
        g_i8_group = 9;
        if (const char* e = getenv("RNNG_L2_I8_GROUP")) {
            const int v = atoi(e);
            if (v >= 0 && v < kI8Codes) g_i8_group = v;
        }
    }
    return g_i8_group;
}

I8Args i8_args(const Workspace& ws, int nx, int ny, int self_join) {
    I8Args a;
    a.xa = ws.xa.as<i32x4_t>();
    a.yb = ws.yb.as<i32x4_t>();
    a.rterm = ws.xn.as<int>();
    a.cterm = ws.ycn.as<int>();
    a.work = ws.work.as<int4>();
    a.n_items = (unsigned)ws.work_host.size();
    a.n_rows = nx;
    a.n_cols = ny;
    a.self_join = self_join;
    a.cand = ws.cand.as<int2>();
    a.cap = ws.cand_cap;
    a.counter = ws.counters.as<unsigned long long>();
    return a;
}

}  // namespace

int filter_dispatch_i8(const I8Args& a, int KS8, int code) {
    if (a.n_items == 0) return RNNG_OK;
    if (a.l16 != (code >= 9 && code <= 12) || (a.l16 && KS8 % 2 != 0)) {
        set_error("int8 filter: the kernel does not match the fragment layout");
        return RNNG_ERR_ARG;
    }
    switch (KS8 * 16 + code) {
        case 2 * 16 + 9: launch_filter_i8<2, 9>(a); break;
        case 2 * 16 + 10: launch_filter_i8<2, 10>(a); break;
        case 2 * 16 + 11: launch_filter_i8<2, 11>(a); break;
        case 2 * 16 + 12: launch_filter_i8<2, 12>(a); break;
        case 4 * 16 + 9: launch_filter_i8<4, 9>(a); break;
        case 4 * 16 + 10: launch_filter_i8<4, 10>(a); break;
        case 4 * 16 + 11: launch_filter_i8<4, 11>(a); break;
        case 4 * 16 + 12: launch_filter_i8<4, 12>(a); break;
        case 1 * 16 + 0: launch_filter_i8<1, 0>(a); break;
        case 1 * 16 + 1: launch_filter_i8<1, 1>(a); break;
        case 1 * 16 + 2: launch_filter_i8<1, 2>(a); break;
        case 1 * 16 + 3: launch_filter_i8<1, 3>(a); break;
        case 1 * 16 + 4: launch_filter_i8<1, 4>(a); break;
        case 1 * 16 + 5: launch_filter_i8<1, 5>(a); break;
        case 1 * 16 + 6: launch_filter_i8<1, 6>(a); break;
        case 1 * 16 + 7: launch_filter_i8<1, 7>(a); break;
        case 1 * 16 + 8: launch_filter_i8<1, 8>(a); break;
        case 2 * 16 + 0: launch_filter_i8<2, 0>(a); break;
        case 2 * 16 + 1: launch_filter_i8<2, 1>(a); break;
        case 2 * 16 + 2: launch_filter_i8<2, 2>(a); break;
        case 2 * 16 + 3: launch_filter_i8<2, 3>(a); break;
        case 2 * 16 + 4: launch_filter_i8<2, 4>(a); break;
        case 2 * 16 + 5: launch_filter_i8<2, 5>(a); break;
        case 2 * 16 + 6: launch_filter_i8<2, 6>(a); break;
        case 2 * 16 + 7: launch_filter_i8<2, 7>(a); break;
        case 2 * 16 + 8: launch_filter_i8<2, 8>(a); break;
        case 3 * 16 + 0: launch_filter_i8<3, 0>(a); break;
        case 3 * 16 + 1: launch_filter_i8<3, 1>(a); break;
        case 3 * 16 + 2: launch_filter_i8<3, 2>(a); break;
        case 3 * 16 + 3: launch_filter_i8<3, 3>(a); break;
        case 3 * 16 + 4: launch_filter_i8<3, 4>(a); break;
        case 3 * 16 + 5: launch_filter_i8<3, 5>(a); break;
        case 3 * 16 + 6: launch_filter_i8<3, 6>(a); break;
        case 3 * 16 + 7: launch_filter_i8<3, 7>(a); break;
        case 3 * 16 + 8: launch_filter_i8<3, 8>(a); break;
        case 4 * 16 + 0: launch_filter_i8<4, 0>(a); break;
        case 4 * 16 + 1: launch_filter_i8<4, 1>(a); break;
        case 4 * 16 + 2: launch_filter_i8<4, 2>(a); break;
        case 4 * 16 + 3: launch_filter_i8<4, 3>(a); break;
        case 4 * 16 + 4: launch_filter_i8<4, 4>(a); break;
        case 4 * 16 + 5: launch_filter_i8<4, 5>(a); break;
        case 4 * 16 + 6: launch_filter_i8<4, 6>(a); break;
        case 4 * 16 + 7: launch_filter_i8<4, 7>(a); break;
        case 4 * 16 + 8: launch_filter_i8<4, 8>(a); break;
        case 5 * 16 + 0: launch_filter_i8<5, 0>(a); break;
        case 5 * 16 + 1: launch_filter_i8<5, 1>(a); break;
        case 5 * 16 + 2: launch_filter_i8<5, 2>(a); break;
        case 5 * 16 + 3: launch_filter_i8<5, 3>(a); break;
        case 5 * 16 + 4: launch_filter_i8<5, 4>(a); break;
        case 5 * 16 + 5: launch_filter_i8<5, 5>(a); break;
        case 5 * 16 + 6: launch_filter_i8<5, 6>(a); break;
        case 5 * 16 + 7: launch_filter_i8<5, 7>(a); break;
        case 5 * 16 + 8: launch_filter_i8<5, 8>(a); break;
        default:
            set_error("int8 filter supports d <= 160");
            return RNNG_ERR_UNSUPPORTED;
    }
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

namespace {

template <class T, int KS, int G, int MINW, int NW, bool PQ, bool L16 = false>
int launch_tiles(const TileFilterArgs& a) {
    auto kern = k_l2_mfma_filter_ring<T, KS, G, MINW, 2, NW, true, PQ, MRT, L16>;
    unsigned grid = a.n_items;
    if constexpr (PQ) {
        // resident workgroups of this shape on the whole device (once)
        static int slots = 0;
        if (slots == 0) {
            int per_cu = 0, cus = 0, dev = 0;
            RNNG_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * NW, 0));
            RNNG_HIP_TRY(hipGetDevice(&dev));
            RNNG_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            slots = std::max(1, per_cu * cus);
        }
        grid = std::min<unsigned>(grid, (unsigned)slots);
        if (a.dry_run) return RNNG_OK;
        RNNG_TRY(launch_fill_u32(a.queue, 1, 0u));
    }
    if (a.dry_run) return RNNG_OK;
    using frag = typename T::frag;
    using term = typename T::term;
    kern<<<grid, 64 * NW, 0, stream()>>>(
        static_cast<const frag*>(a.xa), static_cast<const frag*>(a.yb), static_cast<const term*>(a.rterm),
        static_cast<const term*>(a.cterm), a.work, a.n_rows, a.n_cols, a.self_join, a.cand, a.cap, a.counter, a.ctl,
        a.n_items, a.queue);
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

template <int KS, bool PQ>
int launch_tiles_bf16(const TileFilterArgs& a) {
    if constexpr (KS % 2 == 0) {
        if (a.variant == 19) return launch_tiles<MfBF16, KS, 2, 1, 4, PQ, true>(a);
        if (a.variant == 18) return launch_tiles<MfBF16, KS, (KS >= 9 ? 6 : 8), 1, 12, PQ, true>(a);
    }
    if (a.variant == 13) return launch_tiles<MfBF16, KS, 2, 1, 4, PQ>(a);
    return launch_tiles<MfBF16, KS, (KS >= 9 ? 6 : 8), 1, 12, PQ>(a);
}

template <bool PQ>
int dispatch_tiles(const TileFilterArgs& a) {
    if (a.i8 && a.l16) {   // the 16x16 form of the four-wave shape
        if (a.ks == 2) return launch_tiles<MfI8, 2, 3, 4, 4, PQ, true>(a);
        if (a.ks == 4) return launch_tiles<MfI8, 4, 3, 4, 4, PQ, true>(a);
    }
    if (a.i8) {
        switch (a.ks) {
            case 1: return launch_tiles<MfI8, 1, 3, 4, 4, PQ>(a);
            case 2: return launch_tiles<MfI8, 2, 3, 4, 4, PQ>(a);
            case 3: return launch_tiles<MfI8, 3, 3, 4, 4, PQ>(a);
            case 4: return launch_tiles<MfI8, 4, 3, 4, 4, PQ>(a);
            case 5: return launch_tiles<MfI8, 5, 3, 1, 4, PQ>(a);
            default: break;
        }
        set_error("int8 tile-list filter supports d <= 160");
        return RNNG_ERR_UNSUPPORTED;
    }
    switch (a.ks) {
        case 1: return launch_tiles_bf16<1, PQ>(a);
        case 2: return launch_tiles_bf16<2, PQ>(a);
        case 3: return launch_tiles_bf16<3, PQ>(a);
        case 4: return launch_tiles_bf16<4, PQ>(a);
        case 5: return launch_tiles_bf16<5, PQ>(a);
        case 6: return launch_tiles_bf16<6, PQ>(a);
        case 7: return launch_tiles_bf16<7, PQ>(a);
        case 8: return launch_tiles_bf16<8, PQ>(a);
        case 9: return launch_tiles_bf16<9, PQ>(a);
        default: break;
    }
    set_error("BF16 tile-list filter supports d <= 144");
    return RNNG_ERR_UNSUPPORTED;
}

}  // namespace

int filter_dispatch_tiles(const TileFilterArgs& a) {
    if (a.n_items == 0 && !a.dry_run) return RNNG_OK;
    const bool bf_l32 = a.variant == 17 || a.variant == 13, bf_l16 = a.variant == 18 || a.variant == 19;
    if (!a.ctl || (a.persistent && !a.queue) || (!a.i8 && !(a.l16 ? bf_l16 : bf_l32)) ||
        (a.l16 && a.ks % 2 != 0)) {
        set_error("filter_dispatch_tiles: bad arguments");
        return RNNG_ERR_ARG;
    }
    return a.persistent ? dispatch_tiles<true>(a) : dispatch_tiles<false>(a);
}

int filter_i8_code() { return i8_group(); }
int filter_i8_row_tiles(int code) { return i8_row_tiles(code); }

namespace {

FilterArgs ws_args(Workspace& ws, int n_rows, int n_cols, int self_join) {
    FilterArgs a;
    a.xa = ws.xa.as<bf16x8_t>();
    a.yb = ws.yb.as<bf16x8_t>();
    a.xn = ws.xn.as<float>();
    a.ycn = ws.ycn.as<float>();
    a.work = ws.work.as<int4>();
    a.n_items = (unsigned)ws.work_host.size();
    a.n_rows = n_rows;
    a.n_cols = n_cols;
    a.self_join = self_join;
    a.prm = ws.params.as<MfmaParams>();
    a.cand = ws.cand.as<int2>();
    a.cap = ws.cand_cap;
    a.counter = ws.counters.as<unsigned long long>();
    return a;
}

int launch_recheck_i8(const Workspace& ws, int KS8, bool self_join, unsigned long long ncand, int Tq, float eps,
                      uint8_t* flag, float* cw) {
    if (ncand == 0) return RNNG_OK;
    const unsigned grid = (unsigned)std::min<unsigned long long>(8192, (ncand + 31) / 32);
    k_recheck_i8<<<grid, 256, 0, stream()>>>(ws.xr8.as<i32x4_t>(), self_join ? ws.xr8.as<i32x4_t>() : ws.yr8.as<i32x4_t>(),
                                             ws.hx.as<int>(), self_join ? ws.hx.as<int>() : ws.hy.as<int>(), KS8,
                                             ws.cand.as<int2>(), ws.cand_cap, Tq, eps, flag, cw,
                                             ws.counters.as<unsigned long long>());
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

// The MFMA filters bound |x - y|^2 from centered BF16 or int8 forms; that
// needs finite coordinates, and FP32 sums of d squared differences of at most
// the widest per-dimension range that stay far from overflow.
bool mfma_filter_applies(const unsigned* hs, float range, int d) {
    return hs[ST_NONFIN] == 0 && range >= 0.0f && (double)d * range * range <= 1e36;
}

int run_plain(const rnng_dataset_s* A, const rnng_dataset_s* B, bool self_join, float eps,
              rnng_graph_s* g, rnng_stats* st) {
    Workspace& ws = g_ws;
    const int nx = (int)A->n, ny = (int)B->n, d = A->d;
    const float T = sq_threshold_host(eps);
    const int nbx = (nx + PB - 1) / PB, nby = (ny + PB - 1) / PB;
    const long long nblocks = self_join ? (long long)nbx * (nbx + 1) / 2 : (long long)nbx * nby;
    if (g->cap == 0) {
        const int64_t guess = g_edge_hint > 0 ? g_edge_hint
                                              : std::max<int64_t>(1 << 20, 32 * (int64_t)std::max(nx, ny));
        RNNG_TRY(edge_buffers(g, guess));
    }
    Events ev;
    RNNG_TRY(ev.create());
    for (int attempt = 0; attempt < 2; ++attempt) {
        RNNG_TRY(launch_fill_u64(ws.counters.as<unsigned long long>(), 8, 0ull));
        RNNG_HIP_TRY(hipEventRecord(ev.e[0], stream()));
        // at most 2^23 blocks (2^31 work-items) per dispatch
        const long long chunk = 1ll << 23;
        for (long long b0 = 0; b0 < nblocks; b0 += chunk) {
            const long long nb = std::min(chunk, nblocks - b0);
            k_l2_plain<<<(unsigned)nb, 256, 0, stream()>>>(
                A->X, nx, B->X, ny, d, T, self_join ? 1 : 0, nbx, nby, b0, g->ij.as<int2>(),
                g->w.as<float>(), (unsigned long long)g->cap, ws.counters.as<unsigned long long>() + 1);
            RNNG_HIP_TRY(hipGetLastError());
        }
        RNNG_HIP_TRY(hipEventRecord(ev.e[1], stream()));
        unsigned long long cnt[8];
        RNNG_HIP_TRY(hipMemcpyAsync(cnt, ws.counters.ptr, sizeof(cnt), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        const int64_t ne = (int64_t)cnt[1];
        if (ne <= g->cap) {
            g->n_edges = ne;
            if (st) {
                st->edges += ne;
                st->filter_ms += ev.ms(0, 1);
                st->total_ms += ev.ms(0, 1);
            }
            return RNNG_OK;
        }
        if (st) st->refills++;
        RNNG_TRY(edge_buffers(g, ne + ne / 8 + 1024));
    }
    set_error("edge buffer refill did not converge");
    return RNNG_ERR_HIP;
}

int run_mfma(const rnng_dataset_s* A, const rnng_dataset_s* B, bool self_join, float eps,
             rnng_graph_s* g, rnng_stats* st) {
    Workspace& ws = g_ws;
    const int nx = (int)A->n, ny = (int)B->n, d = A->d;
    const int KS = mfma_ksteps(d);
    const float T = sq_threshold_host(eps);
    const float eps_band = (float)((double)eps * (1.0 + 1e-5));
    float Tb = sq_threshold_host(std::nextafter(eps_band, INFINITY));
    if (Tb < T) Tb = T;

    RNNG_TRY(ws.colsum.reserve(sizeof(double) * d));
    RNNG_TRY(ws.colmin.reserve(sizeof(unsigned) * d));
    RNNG_TRY(ws.colmax.reserve(sizeof(unsigned) * d));
    RNNG_TRY(ws.stat.reserve(sizeof(unsigned) * NSTAT));
    RNNG_TRY(ws.range.reserve(sizeof(unsigned)));
    RNNG_TRY(ws.mu.reserve(sizeof(float) * d));
    RNNG_TRY(ws.params.reserve(sizeof(MfmaParams)));

    Events ev;
    RNNG_TRY(ev.create());
    RNNG_HIP_TRY(hipEventRecord(ev.e[0], stream()));

    // ---- column statistics: integer flag, ranges, centering ----
    {
        // colmin initialised to f2ord(+inf) = 0xff800000, colmax to f2ord(-inf) = 0x007fffff
        k_fill_u64<<<blocks_for(d, 256), 256, 0, stream()>>>((unsigned long long*)ws.colsum.ptr, d, 0ull);
        RNNG_TRY(launch_fill_u32(ws.colmin.as<unsigned>(), d, 0xff800000u));
        RNNG_TRY(launch_fill_u32(ws.colmax.as<unsigned>(), d, 0x007fffffu));
        RNNG_TRY(launch_fill_u32(ws.stat.as<unsigned>(), NSTAT, 0u));
        RNNG_TRY(launch_fill_u32(ws.stat.as<unsigned>() + ST_INT, 1, 1u));
        RNNG_TRY(launch_fill_u32(ws.stat.as<unsigned>() + ST_MAXDEV, 1, 0x80000000u));
        RNNG_TRY(launch_fill_u32(ws.range.as<unsigned>(), 1, 0x80000000u));
    }
    const rnng_dataset_s* sets[2] = {A, B};
    const int nsets = self_join ? 1 : 2;
    int ntot = 0;
    for (int q = 0; q < nsets; ++q) {
        const rnng_dataset_s* S = sets[q];
        if (S->n > 0)
            k_colstats<<<blocks_for(S->n, COLSTAT_ROWS), 256, 0, stream()>>>(
                S->X, (int)S->n, d, ws.colsum.as<double>(), ws.colmin.as<unsigned>(),
                ws.colmax.as<unsigned>(), ws.stat.as<unsigned>());
        ntot += (int)S->n;
    }
    RNNG_HIP_TRY(hipGetLastError());
    k_center<<<blocks_for(d, 64), 64, 0, stream()>>>(ws.colsum.as<double>(), ws.colmin.as<unsigned>(),
                                                      ws.colmax.as<unsigned>(), ntot, d,
                                                      ws.mu.as<float>(), ws.stat.as<unsigned>(),
                                                      ws.range.as<unsigned>());
    RNNG_HIP_TRY(hipGetLastError());
    // integer coordinates spanning at most 255 per dimension: the exact int8
    // filter; coordinates no filter can bound (NaN, infinities, ranges whose
    // squares approach FP32 overflow) or d > 144 (row fragments held in
    // registers): the plain kernel, exact for any input
    bool i8 = false;
    {
        unsigned hs[NSTAT], hr = 0;
        RNNG_HIP_TRY(hipMemcpyAsync(hs, ws.stat.ptr, sizeof(hs), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipMemcpyAsync(&hr, ws.range.ptr, sizeof(hr), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        const float range = ord2f_host(hr);
        if (!mfma_filter_applies(hs, range, d) || KS > kMaxKS) {
            if (st) ++st->plain_fallbacks;
            return run_plain(A, B, self_join, eps, g, st);
        }
        i8 = g_int8_filter && hs[ST_INT] != 0 && range <= 255.0f;
    }
    const int KS8 = (d + 31) / 32;

    // the kernel that runs, and whether its fragments take the 16-row layout
    int variant = 0, i8_code = 0;
    const bool l16 = i8 ? l16_i8_code(i8_group(), KS8, &i8_code) : l16_variant(filter_variant(), KS, &variant);
    const int rb_tiles = i8 ? i8_row_tiles(i8_code) : filter_block_tiles(variant);   // row tiles per workgroup
    const int rb_rows = 32 * rb_tiles;
    const int n_rowtiles = ((nx + rb_rows - 1) / rb_rows) * rb_tiles;
    const int n_coltiles = (ny + 31) / 32;
    const int n_rowblocks = n_rowtiles / rb_tiles;
    const int npad_x = n_rowtiles * 32;
    const int npad_y = std::max(n_coltiles * 32, self_join ? npad_x : 0);
    const int npad = std::max(npad_x, npad_y);
    const size_t frag_bytes_x = (size_t)(self_join ? npad : npad_x) * KS * 16 * 2;
    const size_t frag_bytes_y = (size_t)npad_y * KS * 16 * 2;

    // For a self join the x-form and the y-form come from the same points.
    // (The int8 forms need half the bytes: KS8 * 32 per point.)
    RNNG_TRY(ws.xa.reserve(frag_bytes_x));
    RNNG_TRY(ws.yb.reserve(frag_bytes_y));
    RNNG_TRY(ws.xn.reserve((size_t)npad * sizeof(float)));
    RNNG_TRY(ws.yn.reserve((size_t)npad * sizeof(float)));
    RNNG_TRY(ws.ycn.reserve((size_t)npad * sizeof(float)));
    RNNG_TRY(build_work(ws, n_rowblocks, rb_tiles, n_coltiles, self_join ? 1 : 0));
    {
        const unsigned long long guess =
            g_cand_hint > 0 ? (unsigned long long)g_cand_hint
                            : std::max<unsigned long long>(1ull << 22,
                                                           48ull * (unsigned long long)std::max(nx, ny));
        if (ws.cand_cap < guess || g_cand_hint > 0) {
            RNNG_TRY(ws.cand.reserve(guess * sizeof(int2)));
            ws.cand_cap = guess;
        }
    }
    if (g->cap == 0) {
        const int64_t guess = g_edge_hint > 0 ? g_edge_hint
                                              : std::max<int64_t>(1 << 20, 32 * (int64_t)std::max(nx, ny));
        RNNG_TRY(edge_buffers(g, guess));
    }

    // T' = floor(T), clamped: squared distances of int8-eligible data stay below 2^24
    const int Tq = T >= 1073741823.0f ? 1073741823 : (T < 0.0f ? -1 : (int)std::floor(T));
    if (i8) {
        RNNG_TRY(ws.hx.reserve((size_t)npad * sizeof(int)));
        RNNG_TRY(ws.xr8.reserve((size_t)npad * KS8 * 32));
        if (!self_join) {
            RNNG_TRY(ws.hy.reserve((size_t)npad_y * sizeof(int)));
            RNNG_TRY(ws.yr8.reserve((size_t)npad_y * KS8 * 32));
        }
        auto vec_ok = [&](const float* X) { return (d % 4 == 0 && ((uintptr_t)X % 16) == 0) ? 1 : 0; };
        if (self_join) {
            k_prep_i8<<<npad / 32, 256, 0, stream()>>>(A->X, nx, npad, d, KS8, nullptr, vec_ok(A->X),
                                                       ws.colmin.as<unsigned>(),
                                                       Tq, ws.xa.as<i32x4_t>(), ws.yb.as<i32x4_t>(), ws.xn.as<int>(),
                                                       ws.ycn.as<int>(), ws.hx.as<int>(), ws.xr8.as<i32x4_t>(),
                                                       l16 ? 1 : 0);
        } else {
            k_prep_i8<<<npad_x / 32, 256, 0, stream()>>>(A->X, nx, npad_x, d, KS8, nullptr, vec_ok(A->X),
                                                         ws.colmin.as<unsigned>(), Tq, ws.xa.as<i32x4_t>(), nullptr,
                                                         ws.xn.as<int>(), nullptr, ws.hx.as<int>(), ws.xr8.as<i32x4_t>(),
                                                         l16 ? 1 : 0);
            k_prep_i8<<<npad_y / 32, 256, 0, stream()>>>(B->X, ny, npad_y, d, KS8, nullptr, vec_ok(B->X),
                                                         ws.colmin.as<unsigned>(), Tq, nullptr, ws.yb.as<i32x4_t>(),
                                                         nullptr, ws.ycn.as<int>(), ws.hy.as<int>(), ws.yr8.as<i32x4_t>(),
                                                         l16 ? 1 : 0);
        }
        RNNG_HIP_TRY(hipGetLastError());
    } else if (self_join) {
        // one pass writes both forms for every padded row of the join (the
        // row norms cover npad_x rows, the column norms every padded row)
        RNNG_TRY(launch_prep(A->X, nx, npad, d, KS, ws.mu.as<float>(), nullptr, ws.xa.as<uint16_t>(),
                             ws.yb.as<uint16_t>(), npad == npad_x ? ws.xn.as<float>() : nullptr,
                             ws.yn.as<float>(), ws.stat.as<unsigned>(), l16));
        if (npad != npad_x)
            RNNG_HIP_TRY(hipMemcpyAsync(ws.xn.ptr, ws.yn.ptr, (size_t)npad_x * sizeof(float),
                                        hipMemcpyDeviceToDevice, stream()));
    } else {
        RNNG_TRY(launch_prep(A->X, nx, npad_x, d, KS, ws.mu.as<float>(), nullptr, ws.xa.as<uint16_t>(),
                             nullptr, ws.xn.as<float>(), nullptr, ws.stat.as<unsigned>(), l16));
        RNNG_TRY(launch_prep(B->X, ny, npad_y, d, KS, ws.mu.as<float>(), nullptr, nullptr,
                             ws.yb.as<uint16_t>(), nullptr, ws.yn.as<float>(), ws.stat.as<unsigned>(), l16));
    }
    RNNG_HIP_TRY(hipGetLastError());
    if (!i8) {
        k_finalize_c0<<<1, 1, 0, stream()>>>(ws.stat.as<unsigned>(), ws.range.as<unsigned>(), d, KS, Tb,
                                             ws.params.as<MfmaParams>());
        RNNG_HIP_TRY(hipGetLastError());
        const int npy = self_join ? npad : npad_y;
        k_column_bias<<<blocks_for(npy, 256), 256, 0, stream()>>>(ws.yn.as<float>(), ws.ycn.as<float>(), npy,
                                                                   ws.params.as<MfmaParams>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    RNNG_HIP_TRY(hipEventRecord(ev.e[1], stream()));

    for (int attempt = 0; attempt < 3; ++attempt) {
        RNNG_TRY(launch_fill_u64(ws.counters.as<unsigned long long>(), 8, 0ull));
        RNNG_HIP_TRY(hipEventRecord(ev.e[2], stream()));
        if (i8) {
            I8Args ia = i8_args(ws, nx, ny, self_join ? 1 : 0);
            ia.l16 = l16;
            RNNG_TRY(filter_dispatch_i8(ia, KS8, i8_code));
        } else {
            FilterArgs fa = ws_args(ws, nx, ny, self_join ? 1 : 0);
            fa.l16 = l16;
            RNNG_TRY(filter_dispatch(fa, KS, variant));
        }
        RNNG_HIP_TRY(hipEventRecord(ev.e[3], stream()));
        unsigned long long ncand = 0;
        RNNG_HIP_TRY(hipMemcpyAsync(&ncand, ws.counters.ptr, sizeof(ncand), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        if (ncand > ws.cand_cap) {
            // count-then-fill: grow to the exact count and run the filter again
            if (st) st->refills++;
            const unsigned long long nc = ncand + ncand / 8 + 1024;
            RNNG_TRY(ws.cand.reserve(nc * sizeof(int2)));
            ws.cand_cap = nc;
            continue;
        }
        RNNG_TRY(ws.cflag.reserve(std::max<size_t>(1, ncand)));
        RNNG_TRY(ws.cw.reserve(std::max<size_t>(1, ncand) * sizeof(float)));
        if ((int64_t)ncand > g->cap) RNNG_TRY(edge_buffers(g, (int64_t)ncand));
        RNNG_HIP_TRY(hipEventRecord(ev.e[4], stream()));
        if (i8)
            RNNG_TRY(launch_recheck_i8(ws, KS8, self_join, ncand, Tq, eps, ws.cflag.as<uint8_t>(), ws.cw.as<float>()));
        else
            RNNG_TRY(launch_recheck(A->X, B->X, d, ws.cand.as<int2>(), ws.cand_cap, ncand, T, eps,
                                    ws.cflag.as<uint8_t>(), ws.cw.as<float>(), ws.counters.as<unsigned long long>()));
        int64_t ne = 0;
        RNNG_TRY(compact_edges(ws.cand.as<int2>(), ws.cw.as<float>(), ws.cflag.as<uint8_t>(), (int64_t)ncand,
                               g->ij.as<int2>(), g->w.as<float>(), &ne));
        RNNG_HIP_TRY(hipEventRecord(ev.e[5], stream()));
        unsigned long long cnt[8];
        RNNG_HIP_TRY(hipMemcpyAsync(cnt, ws.counters.ptr, sizeof(cnt), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        g->n_edges = ne;
        if (st) {
            st->candidates += (int64_t)ncand;
            st->edges += ne;
            st->int8_filter += i8 ? 1 : 0;
            st->band_pairs += (int64_t)cnt[2];
            st->band_flips += (int64_t)cnt[3];
            st->prep_ms += ev.ms(0, 1);
            st->filter_ms += ev.ms(2, 3);
            st->recheck_ms += ev.ms(4, 5);
            st->total_ms += ev.ms(0, 1) + ev.ms(2, 3) + ev.ms(4, 5);
        }
        return RNNG_OK;
    }
    set_error("candidate/edge buffer refill did not converge");
    return RNNG_ERR_HIP;
}

int check_dataset(const rnng_dataset_s* ds) {
    if (!ds) {
        set_error("null dataset");
        return RNNG_ERR_ARG;
    }
    if (ds->n > 0x7fffffff - 4096) {
        set_error("datasets are limited to 2^31 - 4096 points");
        return RNNG_ERR_UNSUPPORTED;
    }
    return RNNG_OK;
}

}  // namespace

int launch_plain_self(const float* X, int n, int d, float T, int2* ij, float* w, unsigned long long cap,
                      unsigned long long* counter) {
    if (n <= 1) return RNNG_OK;
    const int nb = (n + PB - 1) / PB;
    const long long nblocks = (long long)nb * (nb + 1) / 2;
    const long long chunk = 1ll << 23;   // at most 2^31 work-items per dispatch
    for (long long b0 = 0; b0 < nblocks; b0 += chunk) {
        const long long nbk = std::min(chunk, nblocks - b0);
        k_l2_plain<<<(unsigned)nbk, 256, 0, stream()>>>(X, n, X, n, d, T, 1, nb, nb, b0, ij, w, cap, counter);
        RNNG_HIP_TRY(hipGetLastError());
    }
    return RNNG_OK;
}

int filter_dispatch(const FilterArgs& a, int KS, int variant) {
    if (variant < 0 || variant >= kNumVariants) {
        set_error("unknown filter variant");
        return RNNG_ERR_ARG;
    }
    if (a.l16 != (variant == 18 || variant == 19) || (a.l16 && KS % 2 != 0)) {
        set_error("BF16 filter: the kernel does not match the fragment layout");
        return RNNG_ERR_ARG;
    }
    if (a.n_items == 0) return RNNG_OK;
    switch (KS) {
        case 1: launch_filter_v<1>(a, variant); break;
        case 2: launch_filter_v<2>(a, variant); break;
        case 3: launch_filter_v<3>(a, variant); break;
        case 4: launch_filter_v<4>(a, variant); break;
        case 5: launch_filter_v<5>(a, variant); break;
        case 6: launch_filter_v<6>(a, variant); break;
        case 7: launch_filter_v<7>(a, variant); break;
        case 8: launch_filter_v<8>(a, variant); break;
        case 9: launch_filter_v<9>(a, variant); break;
        default:
            set_error("MFMA mode supports d <= 144");
            return RNNG_ERR_UNSUPPORTED;
    }
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

int filter_block_tiles(int variant) { return kVariants[variant].nw * kVariants[variant].rt; }

bool l16_i8_code(int code, int ks8, int* run_code) {
    const bool l16 = code >= 9 && code <= 12;
    *run_code = l16 && ks8 % 2 != 0 ? (code == 10 ? 3 : 7) : code;
    return l16 && ks8 % 2 == 0;
}

bool l16_variant(int variant, int ks, int* run_variant) {
    const bool l16 = variant == 18 || variant == 19;
    *run_variant = l16 && ks % 2 != 0 ? (variant == 18 ? 17 : 13) : variant;
    return l16 && ks % 2 == 0;
}
int filter_default_variant() { return filter_variant(); }
int filter_num_variants() { return kNumVariants; }

void xcd_permute(std::vector<int4>& list) {
    const size_t per = (list.size() + 7) / 8;
    std::vector<int4> out(per * 8, make_int4(0, 0, 0, 0));   // empty items run no tile
    for (size_t b = 0; b < per * 8; ++b) {
        const size_t src = (b % 8) * per + b / 8;
        if (src < list.size()) out[b] = list[src];
    }
    list.swap(out);
}

int launch_fill_u64(unsigned long long* p, int n, unsigned long long v) {
    k_fill_u64<<<blocks_for(n, 256), 256, 0, stream()>>>(p, n, v);
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

int launch_fill_u32(unsigned* p, int n, unsigned v) {
    k_fill_u32<<<blocks_for(n, 256), 256, 0, stream()>>>(p, n, v);
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

int launch_recheck(const float* X, const float* Y, int d, const int2* cand, unsigned long long ccap,
                   unsigned long long ncand, float T, float eps, uint8_t* flag, float* cw,
                   unsigned long long* counters) {
    if (ncand == 0) return RNNG_OK;
    const unsigned rgrid = (unsigned)std::min<unsigned long long>(2048, (ncand + 63) / 64);
    const size_t rlds = (size_t)4 * 2 * RC * recheck_ld(d) * sizeof(float);
    if ((d % 4) == 0)
        k_l2_recheck<true><<<rgrid, 256, rlds, stream()>>>(X, Y, d, cand, ccap, T, eps, flag, cw, counters);
    else
        k_l2_recheck<false><<<rgrid, 256, rlds, stream()>>>(X, Y, d, cand, ccap, T, eps, flag, cw, counters);
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

int mfma_input_check(MfmaPrep& P, const float* X, int n, int d) {
    P.stats_ready = false;
    RNNG_TRY(P.colsum.reserve(sizeof(double) * d));
    RNNG_TRY(P.colmin.reserve(sizeof(unsigned) * d));
    RNNG_TRY(P.colmax.reserve(sizeof(unsigned) * d));
    RNNG_TRY(P.stat.reserve(sizeof(unsigned) * NSTAT));
    RNNG_TRY(P.range.reserve(sizeof(unsigned)));
    RNNG_TRY(P.mu.reserve(sizeof(float) * d));
    RNNG_TRY(P.params.reserve(sizeof(MfmaParams)));
    k_fill_u64<<<blocks_for(d, 256), 256, 0, stream()>>>((unsigned long long*)P.colsum.ptr, d, 0ull);
    RNNG_TRY(launch_fill_u32(P.colmin.as<unsigned>(), d, 0xff800000u));
    RNNG_TRY(launch_fill_u32(P.colmax.as<unsigned>(), d, 0x007fffffu));
    RNNG_TRY(launch_fill_u32(P.stat.as<unsigned>(), NSTAT, 0u));
    RNNG_TRY(launch_fill_u32(P.stat.as<unsigned>() + ST_INT, 1, 1u));
    RNNG_TRY(launch_fill_u32(P.stat.as<unsigned>() + ST_MAXDEV, 1, 0x80000000u));
    RNNG_TRY(launch_fill_u32(P.range.as<unsigned>(), 1, 0x80000000u));
    if (n > 0)
        k_colstats<<<blocks_for(n, COLSTAT_ROWS), 256, 0, stream()>>>(X, n, d, P.colsum.as<double>(),
                                                             P.colmin.as<unsigned>(), P.colmax.as<unsigned>(),
                                                             P.stat.as<unsigned>());
    k_center<<<blocks_for(d, 64), 64, 0, stream()>>>(P.colsum.as<double>(), P.colmin.as<unsigned>(),
                                                      P.colmax.as<unsigned>(), n, d, P.mu.as<float>(),
                                                      P.stat.as<unsigned>(), P.range.as<unsigned>());
    RNNG_HIP_TRY(hipGetLastError());
    P.i8 = false;
    if (n > 0) {
        unsigned hs[NSTAT], hr = 0;
        RNNG_HIP_TRY(hipMemcpyAsync(hs, P.stat.ptr, sizeof(hs), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipMemcpyAsync(&hr, P.range.ptr, sizeof(hr), hipMemcpyDeviceToHost, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        const float range = ord2f_host(hr);
        if (!mfma_filter_applies(hs, range, d)) {
            set_error("the cell filter needs finite coordinates with d * (widest range)^2 <= 1e36 "
                      "(the all-pairs graph takes any input)");
            return RNNG_ERR_UNSUPPORTED;
        }
        P.i8 = g_int8_filter && hs[ST_INT] != 0 && range <= 255.0f;
    }
    P.stats_ready = true;
    return RNNG_OK;
}

int mfma_i8_rows(const MfmaPrep& P, const float* X, int n, int d, const int* idmap, int npad, int* hsq,
                 i32x4_t* xrow) {
    if (!P.i8 || npad <= 0) return RNNG_OK;
    if (npad % 32 != 0) {
        set_error("int8 rows: positions must be a multiple of 32");
        return RNNG_ERR_ARG;
    }
    const int vec4 = (d % 4 == 0 && ((uintptr_t)X % 16) == 0) ? 1 : 0;
    k_prep_i8<<<npad / 32, 256, 0, stream()>>>(X, n, npad, d, P.KS8, idmap, vec4, P.colmin.as<unsigned>(), 0,
                                               nullptr, nullptr, nullptr, nullptr, hsq, xrow, 0);
    RNNG_HIP_TRY(hipGetLastError());
    return RNNG_OK;
}

int mfma_prepare_mapped(MfmaPrep& P, const float* X, int n, int d, float Tb, const int* rowmap,
                        int n_rowpad, const int* colmap, int n_colpad) {
    const int KS = mfma_ksteps(d);
    if (KS > kMaxKS) {
        set_error("MFMA mode supports d <= 144");
        return RNNG_ERR_UNSUPPORTED;
    }
    if (!P.stats_ready) RNNG_TRY(mfma_input_check(P, X, n, d));
    P.stats_ready = false;
    P.KS = KS;
    {
        // the 16-row layout when the kernels in use are the 16x16 forms
        int run = 0;
        P.l16 = P.i8 ? l16_i8_code(i8_group(), (d + 31) / 32, &run) : l16_variant(filter_variant(), KS, &run);
    }
    RNNG_TRY(P.xa.reserve(std::max<size_t>(16, (size_t)n_rowpad * KS * 16 * 2)));
    RNNG_TRY(P.yb.reserve(std::max<size_t>(16, (size_t)n_colpad * KS * 16 * 2)));
    RNNG_TRY(P.xn.reserve(std::max<size_t>(16, (size_t)n_rowpad * sizeof(float))));
    RNNG_TRY(P.yn.reserve(std::max<size_t>(16, (size_t)n_colpad * sizeof(float))));
    RNNG_TRY(P.ycn.reserve(std::max<size_t>(16, (size_t)n_colpad * sizeof(float))));
    if (P.i8) {
        // the int8 forms, candidates for squared distances <= floor(Tb)
        P.KS8 = (d + 31) / 32;
        const int Tq = Tb >= 1073741823.0f ? 1073741823 : (Tb < 0.0f ? -1 : (int)std::floor(Tb));
        const int vec4 = (d % 4 == 0 && ((uintptr_t)X % 16) == 0) ? 1 : 0;
        if (rowmap == colmap && n_rowpad == n_colpad) {
            RNNG_TRY(P.xrow.reserve(std::max<size_t>(16, (size_t)n_rowpad * P.KS8 * 32)));
            RNNG_TRY(P.hsq.reserve(std::max<size_t>(16, (size_t)n_rowpad * sizeof(int))));
            if (n_rowpad > 0)
                k_prep_i8<<<n_rowpad / 32, 256, 0, stream()>>>(X, n, n_rowpad, d, P.KS8, rowmap, vec4,
                                                               P.colmin.as<unsigned>(), Tq, P.xa.as<i32x4_t>(),
                                                               P.yb.as<i32x4_t>(), P.xn.as<int>(), P.ycn.as<int>(),
                                                               P.hsq.as<int>(), P.xrow.as<i32x4_t>(),
                                                               P.l16 ? 1 : 0);
        } else {
            if (n_rowpad > 0)
                k_prep_i8<<<n_rowpad / 32, 256, 0, stream()>>>(X, n, n_rowpad, d, P.KS8, rowmap, vec4,
                                                               P.colmin.as<unsigned>(), Tq, P.xa.as<i32x4_t>(),
                                                               nullptr, P.xn.as<int>(), nullptr, nullptr, nullptr,
                                                               P.l16 ? 1 : 0);
            if (n_colpad > 0)
                k_prep_i8<<<n_colpad / 32, 256, 0, stream()>>>(X, n, n_colpad, d, P.KS8, colmap, vec4,
                                                               P.colmin.as<unsigned>(), Tq, nullptr,
                                                               P.yb.as<i32x4_t>(), nullptr, P.ycn.as<int>(), nullptr,
                                                               nullptr, P.l16 ? 1 : 0);
        }
        RNNG_HIP_TRY(hipGetLastError());
        return RNNG_OK;
    }
    if (rowmap == colmap && n_rowpad == n_colpad) {
        // one pass writes both forms
        RNNG_TRY(launch_prep(X, n, n_rowpad, d, KS, P.mu.as<float>(), rowmap, P.xa.as<uint16_t>(),
                             P.yb.as<uint16_t>(), P.xn.as<float>(), P.yn.as<float>(), P.stat.as<unsigned>(), P.l16));
    } else {
        RNNG_TRY(launch_prep(X, n, n_rowpad, d, KS, P.mu.as<float>(), rowmap, P.xa.as<uint16_t>(), nullptr,
                             P.xn.as<float>(), nullptr, P.stat.as<unsigned>(), P.l16));
        RNNG_TRY(launch_prep(X, n, n_colpad, d, KS, P.mu.as<float>(), colmap, nullptr, P.yb.as<uint16_t>(),
                             nullptr, P.yn.as<float>(), P.stat.as<unsigned>(), P.l16));
    }
    k_finalize_c0<<<1, 1, 0, stream()>>>(P.stat.as<unsigned>(), P.range.as<unsigned>(), d, KS, Tb,
                                         P.params.as<MfmaParams>());
    RNNG_HIP_TRY(hipGetLastError());
    if (n_colpad > 0) {
        k_column_bias<<<blocks_for(n_colpad, 256), 256, 0, stream()>>>(P.yn.as<float>(), P.ycn.as<float>(), n_colpad,
                                                                       P.params.as<MfmaParams>());
        RNNG_HIP_TRY(hipGetLastError());
    }
    return RNNG_OK;
}

int l2_graph_build(const rnng_dataset_s* A, const rnng_dataset_s* B, bool self_join, float eps,
                   int32_t mode, rnng_graph_s* g, rnng_stats* st) {
    RNNG_TRY(ensure_init());
    RNNG_TRY(check_dataset(A));
    RNNG_TRY(check_dataset(B));
    if (A->d != B->d || A->d <= 0) {
        set_error("dimension mismatch or d <= 0");
        return RNNG_ERR_ARG;
    }
    if (mode != RNNG_MODE_PLAIN && mode != RNNG_MODE_MFMA && mode != RNNG_MODE_GRID) {
        set_error("unknown mode");
        return RNNG_ERR_ARG;
    }
    if (mode == RNNG_MODE_GRID && (!self_join || A->d > 3)) {
        set_error("grid mode is a self join of points with d <= 3");
        return RNNG_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(g_ws_mu);
    RNNG_TRY(g_ws.counters.reserve(8 * sizeof(unsigned long long)));
    // an infinite eps makes every pair an edge: the grid has nothing to prune
    if (mode == RNNG_MODE_GRID && std::isfinite(eps)) return grid_graph_build(A, eps, g, st);
    if (st) {
        const long long nx = A->n, ny = B->n;
        st->pair_evals += self_join ? nx * (nx - 1) / 2 : nx * ny;
    }
    if (A->n == 0 || B->n == 0) {
        RNNG_TRY(edge_buffers(g, 1));
        g->n_edges = 0;
        return RNNG_OK;
    }
    if (mode == RNNG_MODE_PLAIN) return run_plain(A, B, self_join, eps, g, st);
    return run_mfma(A, B, self_join, eps, g, st);
}

namespace {
int graph_build(const rnng_dataset_s* A, const rnng_dataset_s* B, bool self_join, float eps,
                int32_t mode, rnng_graph_t* out, rnng_stats* st) {
    if (!out) {
        set_error("null output handle");
        return RNNG_ERR_ARG;
    }
    *out = nullptr;
    rnng_graph_s* g = new rnng_graph_s();
    const int rc = l2_graph_build(A, B, self_join, eps, mode, g, st);
    if (rc != RNNG_OK) {
        g->ij.release();
        g->w.release();
        delete g;
        return rc;
    }
    *out = g;
    return RNNG_OK;
}

}  // namespace

int debug_mfma_layout(int32_t* rows, int32_t* cols) {
    RNNG_TRY(ensure_init());
    LocalBuf r, c;
    RNNG_TRY(r.reserve(64 * 16 * sizeof(int)));
    RNNG_TRY(c.reserve(64 * 16 * sizeof(int)));
    k_mfma_layout<<<1, 64, 0, stream()>>>(r.as<int>(), c.as<int>());
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_HIP_TRY(hipMemcpyAsync(rows, r.ptr, 64 * 16 * sizeof(int), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipMemcpyAsync(cols, c.ptr, 64 * 16 * sizeof(int), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    r.release();
    c.release();
    return RNNG_OK;
}

}  // namespace rnng

// ===========================================================================
// C ABI (L2 path)
// ===========================================================================
using namespace rnng;

extern "C" {

float rnng_sq_threshold(float eps) { return sq_threshold_host(eps); }

int rnng_set_int8_filter(int32_t enable) {
    g_int8_filter = enable != 0;
    return RNNG_OK;
}

int rnng_set_filter_kernels(int32_t int8_code, int32_t bf16_variant) {
    if (int8_code >= kI8Codes || bf16_variant >= kNumVariants) {
        set_error("rnng_set_filter_kernels: int8 code 0..12, BF16 variant 0..19 (negative: unchanged)");
        return RNNG_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_ws_mu);   // not in the middle of a build
    if (int8_code >= 0) g_i8_group = int8_code;
    if (bf16_variant >= 0) g_variant = bf16_variant;
    return RNNG_OK;
}

int rnng_get_filter_kernels(int32_t* int8_code, int32_t* bf16_variant) {
    std::lock_guard<std::mutex> lock(g_ws_mu);
    if (int8_code) *int8_code = i8_group();
    if (bf16_variant) *bf16_variant = filter_variant();
    return RNNG_OK;
}

int rnng_dataset_create_f32(const float* points, int64_t n, int32_t d, rnng_dataset_t* out) {
    if (!out || (n > 0 && !points) || n < 0 || d <= 0) {
        set_error("rnng_dataset_create_f32: bad arguments");
        return RNNG_ERR_ARG;
    }
    *out = nullptr;
    RNNG_TRY(ensure_init());
    rnng_dataset_s* ds = new rnng_dataset_s();
    ds->n = n;
    ds->cap = std::max<int64_t>(n, 1);
    ds->d = d;
    ds->owned = true;
    const size_t bytes = (size_t)std::max<int64_t>(n, 1) * d * sizeof(float);
    hipError_t e = hipMalloc(&ds->X, bytes);
    if (e != hipSuccess) {
        delete ds;
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return RNNG_ERR_HIP;
    }
    if (n > 0) {
        e = hipMemcpy(ds->X, points, (size_t)n * d * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(ds->X);
            delete ds;
            set_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
            return RNNG_ERR_HIP;
        }
    }
    *out = ds;
    return RNNG_OK;
}

int rnng_dataset_upload_f32(rnng_dataset_t ds, const float* points, int64_t n) {
    if (!ds || n < 0 || (n > 0 && !points)) {
        set_error("rnng_dataset_upload_f32: bad arguments");
        return RNNG_ERR_ARG;
    }
    if (!ds->owned || n > ds->cap) {
        set_error("rnng_dataset_upload_f32: the dataset is not the library's or holds fewer points");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    if (n > 0)
        RNNG_HIP_TRY(hipMemcpyAsync(ds->X, points, (size_t)n * ds->d * sizeof(float), hipMemcpyHostToDevice,
                                    stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    ds->n = n;
    return RNNG_OK;
}

int rnng_dataset_wrap_device_f32(const float* dev_points, int64_t n, int32_t d, rnng_dataset_t* out) {
    if (!out || (n > 0 && !dev_points) || n < 0 || d <= 0) {
        set_error("rnng_dataset_wrap_device_f32: bad arguments");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    rnng_dataset_s* ds = new rnng_dataset_s();
    ds->X = const_cast<float*>(dev_points);
    ds->n = n;
    ds->d = d;
    ds->owned = false;
    *out = ds;
    return RNNG_OK;
}

int rnng_dataset_destroy(rnng_dataset_t ds) {
    if (!ds) return RNNG_OK;
    if (ds->owned && ds->X) (void)hipFree(ds->X);
    delete ds;
    return RNNG_OK;
}

int rnng_graph_build(rnng_dataset_t ds, float eps, int32_t mode, rnng_graph_t* out, rnng_stats* st) {
    const auto t0 = std::chrono::steady_clock::now();
    const int s = graph_build(ds, ds, true, eps, mode, out, st);
    if (st) st->call_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return s;
}

int rnng_graph_build_bipartite(rnng_dataset_t a, rnng_dataset_t b, float eps, int32_t mode,
                               rnng_graph_t* out, rnng_stats* st) {
    const auto t0 = std::chrono::steady_clock::now();
    const int s = graph_build(a, b, false, eps, mode, out, st);
    if (st) st->call_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return s;
}

int rnng_graph_num_edges(rnng_graph_t g, int64_t* n_edges) {
    if (!g || !n_edges) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    *n_edges = g->n_edges;
    return RNNG_OK;
}

int rnng_graph_copy(rnng_graph_t g, int64_t* ij, float* w, int64_t cap, int32_t sorted) {
    if (!g || cap < 0) {
        set_error("null graph");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    const int64_t ne = std::min<int64_t>(g->n_edges, cap);
    if (ne <= 0) return RNNG_OK;
    const int64_t E = g->n_edges;
    if (sorted) {
        // sort by (i, j) on the device: key = i << 32 | j
        LocalBuf keys;
        RNNG_TRY(keys.reserve((size_t)E * sizeof(uint64_t)));
        k_pack_keys<<<blocks_for(E, 256), 256, 0, stream()>>>(g->ij.as<int2>(), keys.as<uint64_t>(), E);
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(sort_pairs_u64_f32(keys.as<uint64_t>(), g->w.as<float>(), E));
        k_unpack_keys<<<blocks_for(E, 256), 256, 0, stream()>>>(keys.as<uint64_t>(), g->ij.as<int2>(), E);
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    }
    std::vector<int2> hij((size_t)ne);
    RNNG_HIP_TRY(hipMemcpyAsync(hij.data(), g->ij.ptr, (size_t)ne * sizeof(int2), hipMemcpyDeviceToHost, stream()));
    if (w) RNNG_HIP_TRY(hipMemcpyAsync(w, g->w.ptr, (size_t)ne * sizeof(float), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    if (ij)
        for (int64_t e = 0; e < ne; ++e) {
            ij[2 * e] = hij[(size_t)e].x;
            ij[2 * e + 1] = hij[(size_t)e].y;
        }
    return RNNG_OK;
}

int rnng_graph_copy_directed(rnng_graph_t g, int64_t n_vertices, int64_t* uv, float* w, int64_t cap,
                             int64_t* needed) {
    if (!g || cap < 0 || n_vertices < 0 || n_vertices >= INT32_MAX) {
        set_error("graph_copy_directed: bad arguments");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    const int64_t E = g->n_edges;
    const int64_t total = 2 * E + n_vertices;
    if (needed) *needed = total;
    if (cap < total) {
        if (uv || w) {
            set_error("graph_copy_directed: output holds fewer than 2 E + n edges");
            return RNNG_ERR_CAPACITY;
        }
        return RNNG_OK;
    }
    if (total == 0) return RNNG_OK;
    LocalBuf keys, ww;
    RNNG_TRY(keys.reserve((size_t)total * sizeof(uint64_t)));
    RNNG_TRY(ww.reserve((size_t)total * sizeof(float)));
    k_directed_keys<<<blocks_for(total, 256), 256, 0, stream()>>>(g->ij.as<int2>(), g->w.as<float>(), E, n_vertices,
                                                                  keys.as<uint64_t>(), ww.as<float>());
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_TRY(sort_pairs_u64_f32(keys.as<uint64_t>(), ww.as<float>(), total));
    std::vector<uint64_t> hk((size_t)total);
    RNNG_HIP_TRY(hipMemcpyAsync(hk.data(), keys.ptr, (size_t)total * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                stream()));
    if (w) RNNG_HIP_TRY(hipMemcpyAsync(w, ww.ptr, (size_t)total * sizeof(float), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    if (uv)
        for (int64_t e = 0; e < total; ++e) {
            uv[2 * e] = (int64_t)(hk[(size_t)e] >> 32);
            uv[2 * e + 1] = (int64_t)(hk[(size_t)e] & 0xffffffffull);
        }
    return RNNG_OK;
}

int rnng_graph_copy_csr(rnng_graph_t g, int64_t n_vertices, int32_t self_loops, int64_t* rowptr, int32_t* col,
                        float* w, int64_t cap, int64_t* needed) {
    if (!g || cap < 0 || n_vertices < 0 || n_vertices >= INT32_MAX) {
        set_error("graph_copy_csr: bad arguments");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    const int64_t E = g->n_edges;
    const int64_t loops = self_loops ? n_vertices : 0;
    const int64_t total = 2 * E + loops;
    if (needed) *needed = total;
    if (cap < total && (col || w)) {
        set_error("graph_copy_csr: output holds fewer entries than needed");
        return RNNG_ERR_CAPACITY;
    }
    if (!rowptr && !col && !w) return RNNG_OK;
    LocalBuf keys, ww, rp, cl;
    RNNG_TRY(keys.reserve((size_t)std::max<int64_t>(total, 1) * sizeof(uint64_t)));
    RNNG_TRY(ww.reserve((size_t)std::max<int64_t>(total, 1) * sizeof(float)));
    RNNG_TRY(rp.reserve(((size_t)n_vertices + 1) * sizeof(int64_t)));
    RNNG_TRY(cl.reserve((size_t)std::max<int64_t>(total, 1) * sizeof(int)));
    if (total > 0) {
        k_directed_keys<<<blocks_for(total, 256), 256, 0, stream()>>>(g->ij.as<int2>(), g->w.as<float>(), E, loops,
                                                                      keys.as<uint64_t>(), ww.as<float>());
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_TRY(sort_pairs_u64_f32(keys.as<uint64_t>(), ww.as<float>(), total));
        uint64_t last = 0;
        RNNG_HIP_TRY(hipMemcpyAsync(&last, keys.as<uint64_t>() + (total - 1), sizeof(last), hipMemcpyDeviceToHost,
                                    stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
        if ((int64_t)(last >> 32) >= n_vertices) {
            set_error("graph_copy_csr: an edge names a vertex >= n_vertices");
            return RNNG_ERR_ARG;
        }
    }
    k_csr_from_keys<<<blocks_for(total + 1, 256), 256, 0, stream()>>>(keys.as<uint64_t>(), total, n_vertices,
                                                                      rp.as<int64_t>(), cl.as<int>());
    RNNG_HIP_TRY(hipGetLastError());
    if (rowptr)
        RNNG_HIP_TRY(hipMemcpyAsync(rowptr, rp.ptr, ((size_t)n_vertices + 1) * sizeof(int64_t),
                                    hipMemcpyDeviceToHost, stream()));
    if (col && total)
        RNNG_HIP_TRY(hipMemcpyAsync(col, cl.ptr, (size_t)total * sizeof(int), hipMemcpyDeviceToHost, stream()));
    if (w && total)
        RNNG_HIP_TRY(hipMemcpyAsync(w, ww.ptr, (size_t)total * sizeof(float), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

int rnng_graph_destroy(rnng_graph_t g) {
    if (!g) return RNNG_OK;
    g->ij.release();
    g->w.release();
    delete g;
    return RNNG_OK;
}

static int one_shot(const float* a, int64_t na, const float* b, int64_t nb, int32_t d, float eps,
                    int32_t mode, bool self_join, int64_t* ij, float* w, int64_t cap,
                    int64_t* needed, rnng_stats* st) {
    if (!needed || cap < 0 || (cap > 0 && !ij && !w)) {
        set_error("bad output arguments");
        return RNNG_ERR_ARG;
    }
    rnng_dataset_t A = nullptr, B = nullptr;
    RNNG_TRY(rnng_dataset_create_f32(a, na, d, &A));
    int rc = RNNG_OK;
    if (!self_join) rc = rnng_dataset_create_f32(b, nb, d, &B);
    rnng_graph_t g = nullptr;
    if (rc == RNNG_OK) rc = graph_build(A, self_join ? A : B, self_join, eps, mode, &g, st);
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
    rnng_dataset_destroy(A);
    rnng_dataset_destroy(B);
    return rc;
}

int rnng_l2_eps_pairs_f32(const float* points, int64_t n, int32_t d, float eps, int32_t mode,
                          int64_t* ij, float* w, int64_t cap, int64_t* needed, rnng_stats* st) {
    return one_shot(points, n, nullptr, 0, d, eps, mode, true, ij, w, cap, needed, st);
}

int rnng_l2_eps_bipartite_f32(const float* a, int64_t na, const float* b, int64_t nb, int32_t d,
                              float eps, int32_t mode, int64_t* ij, float* w, int64_t cap,
                              int64_t* needed, rnng_stats* st) {
    return one_shot(a, na, b, nb, d, eps, mode, false, ij, w, cap, needed, st);
}

int rnng_debug_sqrt(const float* in, float* out, int64_t n) {
    if (n < 0 || (n > 0 && (!in || !out))) {
        set_error("bad arguments");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    if (n == 0) return RNNG_OK;
    LocalBuf a, b;
    RNNG_TRY(a.reserve((size_t)n * 4));
    RNNG_TRY(b.reserve((size_t)n * 4));
    RNNG_HIP_TRY(hipMemcpyAsync(a.ptr, in, (size_t)n * 4, hipMemcpyHostToDevice, stream()));
    k_debug_sqrt<<<blocks_for(n, 256), 256, 0, stream()>>>(a.as<float>(), b.as<float>(), n);
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    RNNG_HIP_TRY(hipMemcpyAsync(out, b.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

int rnng_debug_copy(int64_t nbytes, double* ms) {
    if (nbytes <= 0 || (nbytes & 15)) {
        set_error("nbytes must be a positive multiple of 16");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    LocalBuf a, b;
    RNNG_TRY(a.reserve((size_t)nbytes));
    RNNG_TRY(b.reserve((size_t)nbytes));
    Events ev;
    RNNG_TRY(ev.create());
    const int64_t n16 = nbytes / 16;
    k_debug_copy<<<4096, 256, 0, stream()>>>(b.as<uint4>(), a.as<uint4>(), n16);   // touch both
    RNNG_HIP_TRY(hipEventRecord(ev.e[0], stream()));
    k_debug_copy<<<4096, 256, 0, stream()>>>(a.as<uint4>(), b.as<uint4>(), n16);
    RNNG_HIP_TRY(hipEventRecord(ev.e[1], stream()));
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    if (ms) *ms = ev.ms(0, 1);
    return RNNG_OK;
}

int rnng_set_initial_capacity(int64_t edges, int64_t candidates) {
    if (edges < 0 || candidates < 0) {
        set_error("capacities must be >= 0");
        return RNNG_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_ws_mu);
    g_edge_hint = edges;
    g_cand_hint = candidates;
    g_ws.cand.release();
    g_ws.cand_cap = 0;
    return RNNG_OK;
}

int rnng_debug_xcd_permute(const int32_t* items, int64_t n, int32_t* out, int64_t cap, int64_t* n_out) {
    if (n < 0 || cap < 0 || !n_out || (n > 0 && !items) || (cap > 0 && !out)) {
        set_error("rnng_debug_xcd_permute: bad arguments");
        return RNNG_ERR_ARG;
    }
    std::vector<int4> list((size_t)n);
    for (int64_t i = 0; i < n; ++i)
        list[(size_t)i] = make_int4(items[4 * i], items[4 * i + 1], items[4 * i + 2], items[4 * i + 3]);
    xcd_permute(list);
    *n_out = (int64_t)list.size();
    if ((int64_t)list.size() > cap) {
        set_error("rnng_debug_xcd_permute: output capacity too small");
        return RNNG_ERR_CAPACITY;
    }
    for (size_t i = 0; i < list.size(); ++i) {
        out[4 * i] = list[i].x;
        out[4 * i + 1] = list[i].y;
        out[4 * i + 2] = list[i].z;
        out[4 * i + 3] = list[i].w;
    }
    return RNNG_OK;
}

int rnng_debug_mfma_layout(int32_t* rows, int32_t* cols) {
    if (!rows || !cols) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    return debug_mfma_layout(rows, cols);
}

}  // extern "C"
