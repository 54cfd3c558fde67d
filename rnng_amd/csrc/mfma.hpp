// Shared definitions of the MFMA candidate filters (l2_eps.hip, the batched
// Voronoi-cell path in voronoi.hip, and the edit-distance filter in
// edit_eps.hip, which shares the per-wave candidate stage).
#pragma once

#include <vector>

#include "internal.hpp"

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef int i32x16_t __attribute__((ext_vector_type(16)));

namespace rnng {

// A wave's own slice of the LDS stage (st, cap_st >= 64 entries), its fill n
// counted in a scalar register: a round needs no LDS reservation, and one
// that would not fit first writes the slice out under one global reservation
// (flush, also at the end of a work item).  No other wave touches the slice.
struct WaveSink {
    int2* st;
    unsigned n;
    unsigned cap_st;
    int2* __restrict__ cand;
    unsigned long long cap;
    unsigned long long* __restrict__ counter;
    __device__ __forceinline__ void flush(int lane) {
        if (n == 0) return;
        unsigned long long gb = 0;
        if (lane == 0) gb = atomicAdd(counter, (unsigned long long)n);
        gb = __shfl(gb, 0, 64);
        // the slice's entries were written by other lanes of this wave: its
        // LDS operations complete in order; the fence keeps the compiler's too
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        for (unsigned i = lane; i < n; i += 64)
            if (gb + i < cap) cand[gb + i] = st[i];
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        n = 0;
    }
    // one round: p = this lane has candidate (row, col), m = the wave's ballot of p
    __device__ __forceinline__ void put(bool p, unsigned long long m, int row, int col, int lane) {
        const unsigned c = (unsigned)__popcll(m);
        if (n + c > cap_st) flush(lane);
        const unsigned slot =
            n + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (p) st[slot] = make_int2(row, col);
        n += c;
    }
};

struct MfmaParams {
    float negc0;      // -C0 as float (|negc0| >= C0)
    int int_mode;     // 1: exact integer filter
    double c0;        // C0 in double, for reporting
};

// Fragment buffers, start values and the statistics that size C0.
struct MfmaPrep {
    LocalBuf xa, yb;                                   // bf16 fragments (x-form rows, y-form columns)
    LocalBuf xn, yn, ycn;                              // row norms, column norms, column norms - C0
    LocalBuf colsum, colmin, colmax, stat, range, mu;  // centering and error statistics
    LocalBuf params;                                   // MfmaParams
    int KS = 0;                                      // 16-wide k-steps per fragment
    // integer data spanning <= 255 per dimension: int8 fragments (xa, yb) and
    // int32 start terms (xn rows, ycn columns) for the exact int8 filter
    bool i8 = false;
    int KS8 = 0;                                     // 32-wide int8 k-steps
    // int8 forms of a mapped self join: x' rows (2 KS8 16-byte words) and the
    // exact |x'|^2 by padded position, for an exact integer recheck
    LocalBuf xrow, hsq;
    bool stats_ready = false;                        // mfma_input_check ran for the next prepare
    // fragments in the 16-row layout of the 16x16 MFMA kernels (set before
    // mfma_prepare_mapped; only kernels of that layout may then run)
    bool l16 = false;
};

// Number of 16-wide k-steps for dimension d.
inline int mfma_ksteps(int d) { return (d + 15) / 16; }
constexpr int kMaxKS = 9;     // d <= 144

// Column statistics of the n points of X (centering, integer flag, ranges),
// read back to decide the filter: RNNG_ERR_UNSUPPORTED when no filter can bound
// the distances (a NaN or an infinity, or d * (widest range)^2 > 1e36), else
// sets P.i8 and marks the statistics ready for the next mfma_prepare_mapped.
int mfma_input_check(MfmaPrep& P, const float* X, int n, int d);

// Builds the x-form fragments and row norms of the padded row positions
// (rowmap[p] = point index or -1) and the y-form fragments and column start
// values (norm - C0, C0 for the squared band radius Tb) of the padded column
// positions (colmap), centred over all n points of X.  On integer data whose
// dimensions each span at most 255 (and unless rnng_set_int8_filter(0)) it
// builds the int8 forms instead and sets P.i8 (threshold floor(Tb)).
int mfma_prepare_mapped(MfmaPrep& P, const float* X, int n, int d, float Tb, const int* rowmap,
                        int n_rowpad, const int* colmap, int n_colpad);

// After an int8 mfma_prepare_mapped: x' rows (2 P.KS8 16-byte words each) and
// exact |x'|^2 of the npad positions of idmap (npad a multiple of 32, -1:
// padding), with the offsets of that preparation.
int mfma_i8_rows(const MfmaPrep& P, const float* X, int n, int d, const int* idmap, int npad, int* hsq,
                 i32x4_t* xrow);

constexpr int LSTAGE = 1024;   // candidates staged in LDS per work item

// One launch of the MFMA filter over a work list of (row block, first column
// tile, end column tile, 0) items; a row block is filter_block_tiles(variant)
// row tiles of 32.  With self_join only pairs row < column are candidates.
struct FilterArgs {
    const bf16x8_t* xa = nullptr;
    const bf16x8_t* yb = nullptr;
    const float* xn = nullptr;     // row norms (accumulator start, row part)
    const float* ycn = nullptr;    // column norms - C0 (accumulator start, column part)
    const int4* work = nullptr;
    unsigned n_items = 0;
    int n_rows = 0, n_cols = 0, self_join = 0;
    const MfmaParams* prm = nullptr;
    int2* cand = nullptr;
    unsigned long long cap = 0;
    unsigned long long* counter = nullptr;
    bool l16 = false;   // fragments in the 16-row layout (variants 18, 19)
};
int filter_dispatch(const FilterArgs& a, int KS, int variant);

// The exact int8 filter (integer data, every dimension spanning <= 255):
// int8 fragments, int32 start terms (rows r_x, columns c_y), candidates as for
// FilterArgs.  code: the kernel shape (0..4, filter_i8_code() = default).
struct I8Args {
    const i32x4_t* xa = nullptr;
    const i32x4_t* yb = nullptr;
    const int* rterm = nullptr;
    const int* cterm = nullptr;
    const int4* work = nullptr;
    unsigned n_items = 0;
    int n_rows = 0, n_cols = 0, self_join = 0;
    int2* cand = nullptr;
    unsigned long long cap = 0;
    unsigned long long* counter = nullptr;
    bool l16 = false;   // fragments in the 16-row layout (codes 9, 10)
};
int filter_dispatch_i8(const I8Args& a, int KS8, int code);

// The ring filter over tile lists (the cell path): work items (row block,
// first, end, 0) index ctl, the column tiles of each row block in order.
// int8 (i8): the four-wave shape with two stages of 3 tiles (code 3, or its
// 16x16 form code 10 with l16); BF16: variant 17 or 13 (18 or 19 with l16).  persistent: a grid of resident workgroups takes items
// from *queue (cleared by the launch) in list order; otherwise one workgroup
// per item.  xa/yb/rterm/cterm are the int8 or BF16 forms.
struct TileFilterArgs {
    bool i8 = false;
    const void* xa = nullptr;
    const void* yb = nullptr;
    const void* rterm = nullptr;
    const void* cterm = nullptr;
    const int4* work = nullptr;
    unsigned n_items = 0;
    const int* ctl = nullptr;
    bool persistent = false;
    unsigned* queue = nullptr;
    int n_rows = 0, n_cols = 0, self_join = 0;
    int2* cand = nullptr;
    unsigned long long cap = 0;
    unsigned long long* counter = nullptr;
    int ks = 0;        // k-steps (KS8 for int8)
    int variant = 17;  // BF16 variant (17 or 13)
    bool dry_run = false;   // size the persistent grid (occupancy query) without launching
    bool l16 = false;  // fragments in the 16-row layout: the 16x16 MFMA forms of the same shapes
};
int filter_dispatch_tiles(const TileFilterArgs& a);
int filter_i8_code();
// The 16x16 MFMA kernels (16-row fragment layout): int8 codes 9 and 10 are
// the 16x16 forms of codes 7 and 3, BF16 variants 18 and 19 those of 17 and
// 13; they need an even number of k-steps.  l16_i8_code / l16_variant give
// the kernel that runs for k-step count ks (the 32x32 form for odd ks), and
// whether its fragments take the 16-row layout.
bool l16_i8_code(int code, int ks8, int* run_code);
bool l16_variant(int variant, int ks, int* run_variant);
int filter_block_tiles(int variant);
int filter_i8_row_tiles(int code);   // row tiles per work item of an int8 kernel code
int filter_default_variant();
int filter_num_variants();
constexpr int kFilterChunk = 192;   // column tiles per chunk of the work list
constexpr int kMaxListItem = 256;   // column tiles per tile-list work item (four registers per lane)
constexpr int kMaxCellWords = 16384;   // cell bitset words the device work-list builder holds in LDS (x2)

// Reorders a work list so that XCD x (workgroups are dealt to the 8 XCDs
// round-robin) runs the contiguous x-th eighth; pads with empty items.
void xcd_permute(std::vector<int4>& list);

// Exact FP32 recheck of candidate pairs (ids into X and Y): flag[c] = 1 for an
// edge (reference-order squared distance <= T), cw[c] = its distance; pairs in
// the |w - eps| <= 1e-5 eps band are counted in counters[2] and re-decided in
// FP64 (disagreements in counters[3]).  The candidate count is read from
// counters[0] on the device (capped at ccap); ncand only sizes the grid.
int launch_recheck(const float* X, const float* Y, int d, const int2* cand, unsigned long long ccap,
                   unsigned long long ncand, float T, float eps, uint8_t* flag, float* cw,
                   unsigned long long* counters);

// Fill kernels usable under stream capture (no memset nodes).
int launch_fill_u64(unsigned long long* p, int n, unsigned long long v);
int launch_fill_u32(unsigned* p, int n, unsigned v);

}  // namespace rnng
