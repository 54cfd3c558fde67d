// Runtime plumbing: device selection, the library stream, error strings,
// growable device buffers.
#include "common.hpp"

#include <mutex>

namespace rnng {

namespace {
thread_local std::string g_last_error;
std::mutex g_init_mu;
int g_device = -1;
hipStream_t g_stream = nullptr;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

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

int DevBuf::reserve(size_t need) {
    if (need <= bytes && ptr) return RNNG_OK;
    release();
    size_t alloc = need ? need : 256;
    hipError_t e = hipMalloc(&ptr, alloc);
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
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

}  // namespace rnng
