// Shared host/device helpers for the rnng_amd HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/rnng_amd.h"

namespace rnng {

// ---------------------------------------------------------------------------
// Error reporting: every C-ABI entry point returns an int status; the message
// of the last failure on the calling thread is kept for rnng_last_error().
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);

#define RNNG_HIP_TRY(expr)                                                        \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ::rnng::set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
            return RNNG_ERR_HIP;                                                  \
        }                                                                         \
    } while (0)

#define RNNG_TRY(expr)               \
    do {                             \
        int s_ = (expr);             \
        if (s_ != RNNG_OK) return s_; \
    } while (0)

// The library's stream on the current device (default priority, non-blocking).
hipStream_t stream();
int ensure_init();

// Device buffer that grows on demand and is released by the owner.  Objects
// with static storage use it as is (the process exit reclaims the memory);
// buffers local to a call use LocalBuf, which frees on scope exit.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t need);   // keeps contents only if no growth is needed
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

struct LocalBuf : DevBuf {
    LocalBuf() = default;
    LocalBuf(const LocalBuf&) = delete;
    LocalBuf& operator=(const LocalBuf&) = delete;
    ~LocalBuf() { release(); }
};

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Work items of the filter kernels are int4 (row block, first column tile,
// end column tile, flags).  An item whose first tile is not below its end
// covers nothing: the XCD-ordered lists are padded with such items.  Every
// filter kernel reads its item through this helper and, when it returns
// false, leaves before any other memory access (the whole workgroup takes the
// same branch, so no barrier is skipped by part of it).
__device__ __forceinline__ bool load_work_item(const int4* __restrict__ work, unsigned item, int4& wi) {
    wi = work[item];
    return wi.y < wi.z;
}
__device__ __forceinline__ bool load_work_item(const int4* __restrict__ work, int4& wi) {
    return load_work_item(work, blockIdx.x, wi);
}

// Exclusive prefix of per-lane counts over the 64-lane wave and one atomic
// reservation per wave.  Every lane of the wave must call it (converged).
// Returns the global base for this lane's first item.
__device__ __forceinline__ unsigned long long wave_reserve(unsigned cnt,
                                                           unsigned long long* counter) {
    const int lane = lane_id();
    unsigned incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        unsigned v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    const unsigned total = __shfl(incl, 63, 64);
    unsigned lo = 0, hi = 0;
    if (lane == 63 && total) {
        unsigned long long b = atomicAdd(counter, (unsigned long long)total);
        lo = (unsigned)b;
        hi = (unsigned)(b >> 32);
    }
    lo = __shfl(lo, 63, 64);
    hi = __shfl(hi, 63, 64);
    return (((unsigned long long)hi << 32) | lo) + (incl - cnt);
}

// Round-to-nearest-even float -> bf16 bits (inputs are finite).
__device__ __host__ __forceinline__ uint16_t f32_to_bf16_rn(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

__device__ __host__ __forceinline__ float bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

}  // namespace rnng
