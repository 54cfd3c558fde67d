// Epsilon pairs of a device-resident point set with no device allocation:
// rnng_l2_eps_pairs_f32_noalloc, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// It runs the PLAIN path of rnng_l2_eps_pairs_f32 (k_l2_plain in the
// reference's FP32 order, then a device radix sort by (i, j)) and keeps every
// intermediate in `scratch`, a fixed array in device memory that each GPU
// holds for the life of the process:
//
//   [0, 256)                 counters (edge count at word 0)
//   keys, keys_alt           uint64 (i << 32 | j), kMaxEdges each
//   vals, vals_alt           float weights, kMaxEdges each
//   rest                     the radix sort's temporary storage
//
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address comes from hipGetSymbolAddress, taken once per device when
// the library initialises that device, and only that address is handed to
// kernels and library calls.
#include <cstring>
#include <mutex>

#include <rocprim/rocprim.hpp>

#include "internal.hpp"

namespace rnng {
namespace {

constexpr size_t kScratchBytes = (size_t)RNNG_NOALLOC_SCRATCH_BYTES;
constexpr int64_t kMaxEdges = RNNG_NOALLOC_MAX_EDGES;
constexpr size_t kCounterBytes = 256;
constexpr size_t kKeysOff = kCounterBytes;
constexpr size_t kKeysAltOff = kKeysOff + (size_t)kMaxEdges * sizeof(uint64_t);
constexpr size_t kValsOff = kKeysAltOff + (size_t)kMaxEdges * sizeof(uint64_t);
constexpr size_t kValsAltOff = kValsOff + (size_t)kMaxEdges * sizeof(float);
constexpr size_t kTempOff = (kValsAltOff + (size_t)kMaxEdges * sizeof(float) + 255) / 256 * 256;
static_assert(kTempOff < kScratchBytes, "scratch layout exceeds the array");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;                          // serialises the entry point's calls
unsigned char* g_scratch[kMaxDevices];    // device address of `scratch`, per device

__global__ void k_zero_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// (i, j) staged as int2 by the plain kernel -> sort key i << 32 | j, in place
__global__ void k_pair_keys(uint64_t* __restrict__ kv, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) {
        const int2 p = reinterpret_cast<const int2*>(kv)[e];
        kv[e] = ((uint64_t)(uint32_t)p.x << 32) | (uint32_t)p.y;
    }
}

__global__ void k_emit_pairs(const uint64_t* __restrict__ keys, const float* __restrict__ vals, int64_t n,
                             int64_t* __restrict__ ij, float* __restrict__ w) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) {
        const uint64_t k = keys[e];
        if (ij) {
            ij[2 * e] = (int64_t)(k >> 32);
            ij[2 * e + 1] = (int64_t)(k & 0xffffffffull);
        }
        if (w) w[e] = vals[e];
    }
}

// Device memory (or managed memory) of the current device; anything else
// would be dereferenced by a kernel as a device address.
int check_device_ptr(const void* p, const char* what) {
    hipPointerAttribute_t a;
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error(std::string("rnng_l2_eps_pairs_f32_noalloc: ") + what + " is not device memory");
        return RNNG_ERR_ARG;
    }
    int dev = -1;
    RNNG_HIP_TRY(hipGetDevice(&dev));
    const bool ok = (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) && a.device == dev;
    if (!ok) {
        set_error(std::string("rnng_l2_eps_pairs_f32_noalloc: ") + what +
                  " is not device memory of the library's device");
        return RNNG_ERR_ARG;
    }
    return RNNG_OK;
}

}  // namespace

int noalloc_attach(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || device >= kMaxDevices) {
        set_error("noalloc_attach: device index beyond the supported range");
        return RNNG_ERR_ARG;
    }
    if (g_scratch[device]) return RNNG_OK;
    void* p = nullptr;
    RNNG_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p || ((uintptr_t)p & 255u)) {
        set_error("noalloc_attach: scratch array has no aligned device address");
        return RNNG_ERR_HIP;
    }
    // the sort's temporary storage for the largest staged list must fit behind
    // the four key/value arrays
    size_t tb = 0;
    rocprim::double_buffer<uint64_t> kb(nullptr, nullptr);
    rocprim::double_buffer<float> vb(nullptr, nullptr);
    RNNG_HIP_TRY(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)kMaxEdges, 0, 64, (hipStream_t)0));
    if (kTempOff + tb > kScratchBytes) {
        set_error("noalloc_attach: sort storage (" + std::to_string(tb) + " bytes) exceeds the scratch array");
        return RNNG_ERR_HIP;
    }
    g_scratch[device] = static_cast<unsigned char*>(p);
    return RNNG_OK;
}

}  // namespace rnng

using namespace rnng;

extern "C" int rnng_l2_eps_pairs_f32_noalloc(const float* dev_points, int64_t n, int32_t d, float eps,
                                             int64_t* dev_ij, float* dev_w, int64_t cap, int64_t* needed,
                                             rnng_stats* st) {
    if (!needed || n < 0 || d <= 0 || cap < 0 || (n > 0 && !dev_points) || (cap > 0 && !dev_ij && !dev_w)) {
        set_error("rnng_l2_eps_pairs_f32_noalloc: bad arguments");
        return RNNG_ERR_ARG;
    }
    if (n > 0x7fffffff - 4096) {
        set_error("rnng_l2_eps_pairs_f32_noalloc: at most 2^31 - 4096 points");
        return RNNG_ERR_UNSUPPORTED;
    }
    RNNG_TRY(ensure_init());
    *needed = 0;
    if (st) st->pair_evals += n * (n - 1) / 2;
    if (n <= 1) return RNNG_OK;
    RNNG_TRY(check_device_ptr(dev_points, "points"));
    if (cap > 0 && dev_ij) RNNG_TRY(check_device_ptr(dev_ij, "ij"));
    if (cap > 0 && dev_w) RNNG_TRY(check_device_ptr(dev_w, "w"));

    std::lock_guard<std::mutex> lk(g_mu);
    int dev = -1;
    RNNG_HIP_TRY(hipGetDevice(&dev));
    unsigned char* base = (dev >= 0 && dev < kMaxDevices) ? g_scratch[dev] : nullptr;
    if (!base) {
        set_error("rnng_l2_eps_pairs_f32_noalloc: scratch array not attached to this device");
        return RNNG_ERR_HIP;
    }
    unsigned long long* counter = reinterpret_cast<unsigned long long*>(base);
    uint64_t* keys = reinterpret_cast<uint64_t*>(base + kKeysOff);
    uint64_t* keys_alt = reinterpret_cast<uint64_t*>(base + kKeysAltOff);
    float* vals = reinterpret_cast<float*>(base + kValsOff);
    float* vals_alt = reinterpret_cast<float*>(base + kValsAltOff);
    void* temp = base + kTempOff;
    const size_t temp_bytes = kScratchBytes - kTempOff;

    // the array keeps the previous call's contents: clear the counters
    k_zero_u64<<<1, 64, 0, stream()>>>(counter, (int)(kCounterBytes / sizeof(unsigned long long)));
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_TRY(launch_plain_self(dev_points, (int)n, d, sq_threshold_host(eps), reinterpret_cast<int2*>(keys), vals,
                               (unsigned long long)kMaxEdges, counter));
    unsigned long long h = 0;
    RNNG_HIP_TRY(hipMemcpyAsync(&h, counter, sizeof(h), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    const int64_t ne = (int64_t)h;
    *needed = ne;
    if (st) st->edges += ne;
    if (ne > kMaxEdges) {
        set_error("rnng_l2_eps_pairs_f32_noalloc: " + std::to_string(ne) + " edges exceed the scratch limit of " +
                  std::to_string(kMaxEdges));
        return RNNG_ERR_UNSUPPORTED;
    }
    if (ne > cap) {
        set_error("output capacity too small");
        return RNNG_ERR_CAPACITY;
    }
    if (ne == 0) return RNNG_OK;
    k_pair_keys<<<blocks_for(ne, 256), 256, 0, stream()>>>(keys, ne);
    RNNG_HIP_TRY(hipGetLastError());
    rocprim::double_buffer<uint64_t> kb(keys, keys_alt);
    rocprim::double_buffer<float> vb(vals, vals_alt);
    size_t need = 0;
    RNNG_HIP_TRY(rocprim::radix_sort_pairs(nullptr, need, kb, vb, (size_t)ne, 0, 64, stream()));
    if (need > temp_bytes) {
        set_error("rnng_l2_eps_pairs_f32_noalloc: sort storage exceeds the scratch array");
        return RNNG_ERR_HIP;
    }
    RNNG_HIP_TRY(rocprim::radix_sort_pairs(temp, need, kb, vb, (size_t)ne, 0, 64, stream()));
    k_emit_pairs<<<blocks_for(ne, 256), 256, 0, stream()>>>(kb.current(), vb.current(), ne, dev_ij, dev_w);
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}
