// Internal types shared by the library's translation units.
#pragma once

#include "common.hpp"

struct rnng_dataset_s {
    float* X = nullptr;       // n*d floats in HBM, row-major
    int64_t n = 0;
    int64_t cap = 0;          // points the owned buffer holds
    int32_t d = 0;
    bool owned = false;
};

// Edge list in HBM: (i, j) int32 pairs and float weights.
struct rnng_graph_s {
    rnng::DevBuf ij;          // int2 (i, j)
    rnng::DevBuf w;           // float distance
    int64_t n_edges = 0;
    int64_t cap = 0;
};

namespace rnng {

extern int64_t g_edge_hint;   // initial edge-buffer size of new graphs (0: default)
int edge_buffers(rnng_graph_s* g, int64_t cap);
float sq_threshold_host(float eps);

// HIP events on the library stream for stage timing.
struct Events {
    hipEvent_t e[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int create() {
        for (auto& x : e) RNNG_HIP_TRY(hipEventCreate(&x));
        return RNNG_OK;
    }
    ~Events() {
        for (auto& x : e)
            if (x) (void)hipEventDestroy(x);
    }
    double ms(int a, int b) const {
        float t = 0.f;
        if (hipEventElapsedTime(&t, e[a], e[b]) != hipSuccess) return 0.0;
        return (double)t;
    }
};

inline unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Device sorts (sort.hip).  All run on the library stream and return after
// the stream has drained.
int sort_pairs_u64_f32(uint64_t* keys, float* vals, int64_t n);
int sort_pairs_u64_i64(uint64_t* keys, int64_t* vals, int64_t n);
// payload (may be null): one byte per element, moved with it.
int sort_by_value_then_id(float* value, int64_t* id, int64_t n, uint8_t* payload = nullptr);
// Stable LSD sorts on key bits [0, end_bit).
int sort_pairs_u32_i32(uint32_t* keys, int32_t* vals, int64_t n, unsigned end_bit);
int sort_pairs_u32_i64(uint32_t* keys, int64_t* vals, int64_t n, unsigned end_bit);
int sort_pairs_u32_u32(uint32_t* keys, uint32_t* vals, int64_t n, unsigned end_bit);
int sort_pairs_u64_f32_bits(uint64_t* keys, float* vals, int64_t n, unsigned end_bit);
// The same, leaving the sorted data in whichever of two buffers the sort
// ended in: keys and vals then own it (no copy back).
int sort_pairs_u32_i64_swap(DevBuf& keys, DevBuf& vals, int64_t n, unsigned end_bit);
int sort_pairs_u32_u32_swap(DevBuf& keys, DevBuf& vals, int64_t n, unsigned end_bit);
int sort_pairs_u64_f32_swap(DevBuf& keys, DevBuf& vals, int64_t n, unsigned end_bit);
int exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n);
// Stable compaction of (ij, w) by flag != 0 into (out_ij, out_w); *n_out gets
// the count.  The outputs must hold n elements.
int compact_edges(const int2* ij, const float* w, const uint8_t* flag, int64_t n, int2* out_ij,
                  float* out_w, int64_t* n_out);

// Builds the L2 epsilon graph of a self join (l2_eps.hip).
int l2_graph_build(const rnng_dataset_s* A, const rnng_dataset_s* B, bool self_join, float eps,
                   int32_t mode, rnng_graph_s* g, rnng_stats* st);

// The plain kernel over the upper triangle of one point set (i < j): writes the
// first `cap` edges as (i, j) and float distance, counts every edge in
// *counter (which the caller zeroes).  Enqueued on the library stream.
int launch_plain_self(const float* X, int n, int d, float T, int2* ij, float* w, unsigned long long cap,
                      unsigned long long* counter);

// Per-device address of the no-allocation entry point's static scratch array
// (noalloc.hip); called by device initialisation with that device current.
int noalloc_attach(int device);

// A whole file into device memory through two pinned host buffers, the next
// chunk read while the previous one is in flight (ingest.hip).  *bytes gets
// the file size; dst null: size query only; a file above max_bytes is an
// error.  Returns after the copies have landed.
int upload_file_bytes(const char* path, uint8_t* dst, int64_t* bytes, int64_t max_bytes);

// Uniform-grid self join for d <= 3 (grid.hip); eps must be finite.
int grid_graph_build(const rnng_dataset_s* A, float eps, rnng_graph_s* g, rnng_stats* st);

}  // namespace rnng
