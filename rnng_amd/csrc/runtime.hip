// Runtime plumbing: device selection, the library stream, error strings,
// growable device buffers.
#include "common.hpp"

#include <map>
#include <mutex>

namespace rnng {

namespace {
thread_local std::string g_last_error;
std::mutex g_init_mu;
int g_device = -1;
hipStream_t g_stream = nullptr;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

int noalloc_attach(int device);

const char* last_error() { return g_last_error.c_str(); }

static int init_locked(int device) {
    if (g_device == device && g_stream) return RNNG_OK;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        set_error("no HIP device visible");
        return RNNG_ERR_NODEV;
    }
    if (device < 0 || device >= count) {
        set_error("device index out of range");
        return RNNG_ERR_ARG;
    }
    hipDeviceProp_t prop;
    RNNG_HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) {
        set_error(std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
        return RNNG_ERR_NODEV;
    }
    RNNG_HIP_TRY(hipSetDevice(device));
    // the no-allocation entry point's scratch array: its address on this device
    RNNG_TRY(noalloc_attach(device));
    if (g_stream) {
        (void)hipStreamDestroy(g_stream);
        g_stream = nullptr;
    }
    RNNG_HIP_TRY(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    g_device = device;
    return RNNG_OK;
}

int init(int device) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    return init_locked(device);
}

int ensure_init() {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_stream) {
        RNNG_HIP_TRY(hipSetDevice(g_device));
        return RNNG_OK;
    }
    return init_locked(0);
}

hipStream_t stream() { return g_stream; }

// Device blocks are cached after release and handed out again to requests
// that need between half of a block and all of it, so calls that allocate the
// same scratch sizes each time (graph builds, filtrations) reuse memory
// instead of paying hipMalloc/hipFree (which also synchronises the device).
// Every device access of the library is ordered on its one stream, so a block
// reused by a later request is only touched after the stream's earlier work.
// At most kPoolCap bytes stay cached; a failing hipMalloc empties the cache
// and retries.  The cache is never destroyed (buffers of static objects may
// still be released during process exit).
namespace {
struct BlockPool {
    std::mutex mu;
    std::multimap<size_t, void*> free_blocks;
    size_t cached = 0;
};
BlockPool& pool() {
    static BlockPool* p = new BlockPool();
    return *p;
}
constexpr size_t kPoolCap = (size_t)48 << 30;

size_t round_size(size_t need) {
    const size_t q = need >= ((size_t)2 << 20) ? ((size_t)2 << 20) : 256;
    return (need + q - 1) / q * q;
}

void empty_pool() {
    BlockPool& p = pool();
    std::lock_guard<std::mutex> lk(p.mu);
    for (auto& kv : p.free_blocks) (void)hipFree(kv.second);
    p.free_blocks.clear();
    p.cached = 0;
}
}  // namespace

int DevBuf::reserve(size_t need) {
    if (need <= bytes && ptr) return RNNG_OK;
    release();
    const size_t alloc = round_size(need ? need : 256);
    {
        BlockPool& p = pool();
        std::lock_guard<std::mutex> lk(p.mu);
        auto it = p.free_blocks.lower_bound(alloc);
        if (it != p.free_blocks.end() && it->first <= 2 * alloc) {
            ptr = it->second;
            bytes = it->first;
            p.cached -= it->first;
            p.free_blocks.erase(it);
            return RNNG_OK;
        }
    }
    hipError_t e = hipMalloc(&ptr, alloc);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        empty_pool();
        e = hipMalloc(&ptr, alloc);
    }
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        set_error(std::string("hipMalloc(") + std::to_string(alloc) + "): " + hipGetErrorString(e));
        return RNNG_ERR_HIP;
    }
    bytes = alloc;
    return RNNG_OK;
}

void DevBuf::release() {
    if (ptr) {
        BlockPool& p = pool();
        std::lock_guard<std::mutex> lk(p.mu);
        if (p.cached + bytes <= kPoolCap) {
            p.free_blocks.emplace(bytes, ptr);
            p.cached += bytes;
        } else {
            (void)hipFree(ptr);
        }
    }
    ptr = nullptr;
    bytes = 0;
}

}  // namespace rnng
