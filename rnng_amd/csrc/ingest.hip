// Point files straight into HBM, for gfx950.
//
// Formats (gabe-raulet/RNNG):
//   .fvecs  records of int32 d + d float32 (include/point.hpp:112-168,
//           dataset_io.py:62-70)
//   .fbin   uint32 n, uint32 d, then n*d float32 (dataset_io.py:55-60)
//   .u8bin  uint32 n, uint32 d, then n*d uint8 (dataset_io.py:48-53)
// start/count select points [start, start + count) as dataset_io.py's
// read_file does (count < 0: to the end of the file).
//
// The file is read in chunks of whole records into two pinned host buffers;
// while one chunk is in flight to the device (asynchronous DMA on the library
// stream) the host reads the next into the other buffer.  fbin chunks land in
// the dataset directly; u8bin chunks are uploaded as bytes (a quarter of the
// PCIe traffic) and widened to float32 on the device; fvecs chunks are
// uploaded as records and the per-record dimension headers are checked and
// stripped on the device.
#include "internal.hpp"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

namespace rnng {
namespace {

enum { KIND_F32 = 0, KIND_U8 = 1 };
enum { FMT_FVECS = 0, FMT_FBIN = 1, FMT_U8BIN = 2 };

struct FileInfo {
    int fmt = -1;
    int64_t n = 0;
    int32_t d = 0;
    int64_t header = 0;   // bytes before the first record
    int64_t record = 0;   // bytes per record
};

bool ends_with(const std::string& s, const char* suf) {
    const size_t m = std::strlen(suf);
    return s.size() >= m && s.compare(s.size() - m, m, suf) == 0;
}

int read_exact(int fd, void* dst, size_t bytes, int64_t off) {
    char* p = static_cast<char*>(dst);
    while (bytes) {
        const ssize_t r = pread(fd, p, bytes, (off_t)off);
        if (r <= 0) {
            set_error("short read");
            return RNNG_ERR_FORMAT;
        }
        p += r;
        bytes -= (size_t)r;
        off += r;
    }
    return RNNG_OK;
}

int probe(int fd, const std::string& path, FileInfo* fi) {
    struct stat st;
    if (fstat(fd, &st) != 0) {
        set_error("cannot stat " + path);
        return RNNG_ERR_ARG;
    }
    const int64_t size = (int64_t)st.st_size;
    if (ends_with(path, ".fvecs")) {
        fi->fmt = FMT_FVECS;
        int32_t d = 0;
        if (size > 0) RNNG_TRY(read_exact(fd, &d, 4, 0));
        if (size > 0 && (d <= 0 || size % (4 * ((int64_t)d + 1)) != 0)) {
            set_error("malformed fvecs file " + path);
            return RNNG_ERR_FORMAT;
        }
        fi->d = d;
        fi->record = 4 * ((int64_t)d + 1);
        fi->n = size > 0 ? size / fi->record : 0;
        fi->header = 0;
        return RNNG_OK;
    }
    if (ends_with(path, ".fbin") || ends_with(path, ".u8bin")) {
        const bool u8 = ends_with(path, ".u8bin");
        fi->fmt = u8 ? FMT_U8BIN : FMT_FBIN;
        uint32_t h[2] = {0, 0};
        if (size < 8) {
            set_error("file too short for a header: " + path);
            return RNNG_ERR_FORMAT;
        }
        RNNG_TRY(read_exact(fd, h, 8, 0));
        fi->n = h[0];
        fi->d = (int32_t)h[1];
        fi->header = 8;
        fi->record = (int64_t)h[1] * (u8 ? 1 : 4);
        if (fi->d <= 0 || (int64_t)h[1] > INT32_MAX || size < 8 + fi->n * fi->record) {
            set_error("header does not match the file size: " + path);
            return RNNG_ERR_FORMAT;
        }
        return RNNG_OK;
    }
    set_error("cannot read file '" + path + "': unknown extension");
    return RNNG_ERR_FORMAT;
}

// u8 rows -> float rows
__global__ void k_widen_u8(const uint8_t* __restrict__ src, float* __restrict__ dst, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = (float)src[i];
}

// fvecs records (int32 d + d floats) -> rows; bad[0] set if a header differs
__global__ void k_strip_fvecs(const uint32_t* __restrict__ src, float* __restrict__ dst, int64_t rows, int d,
                              int* __restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)d + 1;
    if (i >= rows * per) return;
    const int64_t r = i / per, c = i % per;
    const uint32_t v = src[i];
    if (c == 0) {
        if ((int)v != d) atomicOr(bad, 1);
    } else {
        dst[r * d + c - 1] = __uint_as_float(v);
    }
}

__global__ void k_clear_flag(int* f) { *f = 0; }

// Two pinned staging buffers, allocated on first use and kept for the
// process (pinning costs milliseconds per call otherwise); loads are
// serialised on them.
constexpr int64_t kChunk = (int64_t)32 << 20;   // bytes per pinned buffer
std::mutex g_pin_mu;
void* g_pinned[2] = {nullptr, nullptr};

struct Fd {
    int fd = -1;
    ~Fd() {
        if (fd >= 0) close(fd);
    }
};

int load(const char* path, int64_t start, int64_t count, rnng_dataset_s* ds) {
    Fd f;
    f.fd = open(path, O_RDONLY);
    if (f.fd < 0) {
        set_error(std::string("cannot open ") + path);
        return RNNG_ERR_ARG;
    }
    FileInfo fi;
    RNNG_TRY(probe(f.fd, path, &fi));
    if (start < 0 || start > fi.n) {
        set_error("start beyond the end of the file");
        return RNNG_ERR_ARG;
    }
    if (count < 0) count = fi.n - start;
    if (start + count > fi.n) {
        set_error("start + count beyond the end of the file");
        return RNNG_ERR_ARG;
    }
    if (fi.d == 0 && count > 0) {
        set_error("empty records");
        return RNNG_ERR_FORMAT;
    }
    ds->n = count;
    ds->cap = std::max<int64_t>(count, 1);
    ds->d = std::max(fi.d, 1);
    ds->owned = true;
    const size_t bytes = (size_t)std::max<int64_t>(count, 1) * ds->d * sizeof(float);
    RNNG_HIP_TRY(hipMalloc(&ds->X, bytes));
    if (count == 0) return RNNG_OK;

    if (fi.record > kChunk) {
        set_error("records above 32 MiB are not supported");
        return RNNG_ERR_UNSUPPORTED;
    }
    const int64_t rows_per = kChunk / fi.record;
    const int64_t chunk_bytes = rows_per * fi.record;
    std::lock_guard<std::mutex> pin_lock(g_pin_mu);
    void* host[2];
    LocalBuf raw[2], bad;
    hipEvent_t done[2] = {nullptr, nullptr};
    struct EvGuard {
        hipEvent_t* e;
        ~EvGuard() {
            for (int b = 0; b < 2; ++b)
                if (e[b]) (void)hipEventDestroy(e[b]);
        }
    } evg{done};
    // runs on every return: no staging buffer is reused or released while a
    // copy from it may still be in flight
    struct Drain {
        ~Drain() { (void)hipStreamSynchronize(stream()); }
    } drain;
    for (int b = 0; b < 2; ++b) {
        if (!g_pinned[b]) RNNG_HIP_TRY(hipHostMalloc(&g_pinned[b], (size_t)kChunk, hipHostMallocDefault));
        host[b] = g_pinned[b];
        RNNG_HIP_TRY(hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
        if (fi.fmt != FMT_FBIN) RNNG_TRY(raw[b].reserve((size_t)chunk_bytes));
    }
    RNNG_TRY(bad.reserve(sizeof(int)));
    k_clear_flag<<<1, 1, 0, stream()>>>(bad.as<int>());
    bool used[2] = {false, false};
    int64_t r0 = 0;
    for (int it = 0; r0 < count; ++it) {
        const int b = it & 1;
        const int64_t rows = std::min(rows_per, count - r0);
        const int64_t nb = rows * fi.record;
        if (used[b]) RNNG_HIP_TRY(hipEventSynchronize(done[b]));   // buffer b's last DMA has finished
        RNNG_TRY(read_exact(f.fd, host[b], (size_t)nb, fi.header + (start + r0) * fi.record));
        float* dst = ds->X + r0 * ds->d;
        if (fi.fmt == FMT_FBIN) {
            RNNG_HIP_TRY(hipMemcpyAsync(dst, host[b], (size_t)nb, hipMemcpyHostToDevice, stream()));
        } else {
            RNNG_HIP_TRY(hipMemcpyAsync(raw[b].ptr, host[b], (size_t)nb, hipMemcpyHostToDevice, stream()));
            if (fi.fmt == FMT_U8BIN) {
                const int64_t m = rows * fi.d;
                k_widen_u8<<<blocks_for(m, 256), 256, 0, stream()>>>(raw[b].as<uint8_t>(), dst, m);
            } else {
                k_strip_fvecs<<<blocks_for(rows * ((int64_t)fi.d + 1), 256), 256, 0, stream()>>>(
                    raw[b].as<uint32_t>(), dst, rows, fi.d, bad.as<int>());
            }
            RNNG_HIP_TRY(hipGetLastError());
        }
        RNNG_HIP_TRY(hipEventRecord(done[b], stream()));
        used[b] = true;
        r0 += rows;
    }
    int hbad = 0;
    RNNG_HIP_TRY(hipMemcpyAsync(&hbad, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    if (hbad) {
        set_error(std::string("mixed dimensions in ") + path);
        return RNNG_ERR_FORMAT;
    }
    return RNNG_OK;
}

}  // namespace

int upload_file_bytes(const char* path, uint8_t* dst, int64_t* bytes_out, int64_t max_bytes) {
    Fd f;
    f.fd = open(path, O_RDONLY);
    if (f.fd < 0) {
        set_error(std::string("cannot open ") + path);
        return RNNG_ERR_ARG;
    }
    struct stat st;
    if (fstat(f.fd, &st) != 0) {
        set_error(std::string("cannot stat ") + path);
        return RNNG_ERR_ARG;
    }
    const int64_t size = (int64_t)st.st_size;
    *bytes_out = size;
    if (!dst) return RNNG_OK;   // size query
    if (size > max_bytes) {
        set_error("upload_file_bytes: destination too small");
        return RNNG_ERR_ARG;
    }
    std::lock_guard<std::mutex> pin_lock(g_pin_mu);
    hipEvent_t done[2] = {nullptr, nullptr};
    struct EvGuard {
        hipEvent_t* e;
        ~EvGuard() {
            for (int b = 0; b < 2; ++b)
                if (e[b]) (void)hipEventDestroy(e[b]);
        }
    } evg{done};
    struct Drain {
        ~Drain() { (void)hipStreamSynchronize(stream()); }
    } drain;
    for (int b = 0; b < 2; ++b) {
        if (!g_pinned[b]) RNNG_HIP_TRY(hipHostMalloc(&g_pinned[b], (size_t)kChunk, hipHostMallocDefault));
        RNNG_HIP_TRY(hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
    }
    bool used[2] = {false, false};
    for (int64_t off = 0, it = 0; off < size; off += kChunk, ++it) {
        const int b = (int)(it & 1);
        const int64_t nb = std::min(kChunk, size - off);
        if (used[b]) RNNG_HIP_TRY(hipEventSynchronize(done[b]));   // buffer b's last DMA has finished
        RNNG_TRY(read_exact(f.fd, g_pinned[b], (size_t)nb, off));
        RNNG_HIP_TRY(hipMemcpyAsync(dst + off, g_pinned[b], (size_t)nb, hipMemcpyHostToDevice, stream()));
        RNNG_HIP_TRY(hipEventRecord(done[b], stream()));
        used[b] = true;
    }
    return RNNG_OK;
}

}  // namespace rnng

using namespace rnng;

extern "C" {

int rnng_file_info(const char* path, int64_t* n, int32_t* d, int32_t* kind) {
    if (!path || !n || !d || !kind) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    Fd f;
    f.fd = open(path, O_RDONLY);
    if (f.fd < 0) {
        set_error(std::string("cannot open ") + path);
        return RNNG_ERR_ARG;
    }
    FileInfo fi;
    RNNG_TRY(probe(f.fd, path, &fi));
    *n = fi.n;
    *d = fi.d;
    *kind = fi.fmt == FMT_U8BIN ? KIND_U8 : KIND_F32;
    return RNNG_OK;
}

int rnng_dataset_read_file(const char* path, int64_t start, int64_t count, rnng_dataset_t* out) {
    if (!path || !out) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    *out = nullptr;
    RNNG_TRY(ensure_init());
    rnng_dataset_s* ds = new rnng_dataset_s();
    const int rc = load(path, start, count, ds);
    if (rc != RNNG_OK) {
        if (ds->X) (void)hipFree(ds->X);
        delete ds;
        return rc;
    }
    *out = ds;
    return RNNG_OK;
}

int rnng_dataset_shape(rnng_dataset_t ds, int64_t* n, int32_t* d) {
    if (!ds || !n || !d) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    *n = ds->n;
    *d = ds->d;
    return RNNG_OK;
}

int rnng_dataset_copy_f32(rnng_dataset_t ds, float* out) {
    if (!ds || (!out && ds->n > 0)) {
        set_error("null argument");
        return RNNG_ERR_ARG;
    }
    RNNG_TRY(ensure_init());
    if (ds->n == 0) return RNNG_OK;
    RNNG_HIP_TRY(hipMemcpyAsync(out, ds->X, (size_t)ds->n * ds->d * sizeof(float), hipMemcpyDeviceToHost,
                                stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

}  // extern "C"
