// Device-wide sorts and scans (rocPRIM radix sort, LSD and stable) used to
// canonicalise edge lists and filtrations on the GPU.
#include <cstring>
#include <utility>

#include <rocprim/rocprim.hpp>

#include "internal.hpp"

namespace rnng {
namespace {

template <class K, class V>
int sort_pairs_impl(K* keys, V* vals, int64_t n, unsigned end_bit) {
    if (n <= 1) return RNNG_OK;
    LocalBuf kalt, valt, tmp;
    RNNG_TRY(kalt.reserve((size_t)n * sizeof(K)));
    RNNG_TRY(valt.reserve((size_t)n * sizeof(V)));
    size_t tb = 0;
    RNNG_HIP_TRY(rocprim::radix_sort_pairs(nullptr, tb, keys, kalt.as<K>(), vals, valt.as<V>(), (size_t)n,
                                           0, end_bit, stream()));
    RNNG_TRY(tmp.reserve(tb));
    RNNG_HIP_TRY(rocprim::radix_sort_pairs(tmp.ptr, tb, keys, kalt.as<K>(), vals, valt.as<V>(), (size_t)n,
                                           0, end_bit, stream()));
    RNNG_HIP_TRY(hipMemcpyAsync(keys, kalt.ptr, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice, stream()));
    RNNG_HIP_TRY(hipMemcpyAsync(vals, valt.ptr, (size_t)n * sizeof(V), hipMemcpyDeviceToDevice, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

template <class K, class V>
int sort_pairs_swap_impl(DevBuf& keys, DevBuf& vals, int64_t n, unsigned end_bit) {
    if (n <= 1) return RNNG_OK;
    LocalBuf kalt, valt, tmp;
    RNNG_TRY(kalt.reserve((size_t)n * sizeof(K)));
    RNNG_TRY(valt.reserve((size_t)n * sizeof(V)));
    rocprim::double_buffer<K> kb(keys.as<K>(), kalt.as<K>());
    rocprim::double_buffer<V> vb(vals.as<V>(), valt.as<V>());
    size_t tb = 0;
    RNNG_HIP_TRY(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)n, 0, end_bit, stream()));
    RNNG_TRY(tmp.reserve(tb));
    RNNG_HIP_TRY(rocprim::radix_sort_pairs(tmp.ptr, tb, kb, vb, (size_t)n, 0, end_bit, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    if (kb.current() != keys.as<K>()) {
        std::swap(keys.ptr, kalt.ptr);
        std::swap(keys.bytes, kalt.bytes);
    }
    if (vb.current() != vals.as<V>()) {
        std::swap(vals.ptr, valt.ptr);
        std::swap(vals.bytes, valt.bytes);
    }
    return RNNG_OK;
}

__global__ void k_float_bits(const float* __restrict__ v, uint32_t* __restrict__ bits, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        // order-preserving map of a float to uint32 (negative values reversed)
        const uint32_t u = __float_as_uint(v[i]);
        bits[i] = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    }
}

__global__ void k_gather_f32(const float* __restrict__ src, const int64_t* __restrict__ idx,
                             float* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

__global__ void k_gather_i64(const int64_t* __restrict__ src, const int64_t* __restrict__ idx,
                             int64_t* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

__global__ void k_iota_i64(int64_t* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

}  // namespace

int sort_pairs_u64_f32(uint64_t* keys, float* vals, int64_t n) {
    return sort_pairs_impl<uint64_t, float>(keys, vals, n, 64);
}

int sort_pairs_u64_i64(uint64_t* keys, int64_t* vals, int64_t n) {
    return sort_pairs_impl<uint64_t, int64_t>(keys, vals, n, 64);
}

int sort_pairs_u32_i32(uint32_t* keys, int32_t* vals, int64_t n, unsigned end_bit) {
    return sort_pairs_impl<uint32_t, int32_t>(keys, vals, n, end_bit);
}

int sort_pairs_u32_i64(uint32_t* keys, int64_t* vals, int64_t n, unsigned end_bit) {
    return sort_pairs_impl<uint32_t, int64_t>(keys, vals, n, end_bit);
}

int sort_pairs_u32_u32(uint32_t* keys, uint32_t* vals, int64_t n, unsigned end_bit) {
    return sort_pairs_impl<uint32_t, uint32_t>(keys, vals, n, end_bit);
}

int sort_pairs_u64_f32_bits(uint64_t* keys, float* vals, int64_t n, unsigned end_bit) {
    return sort_pairs_impl<uint64_t, float>(keys, vals, n, end_bit);
}

int sort_pairs_u32_i64_swap(DevBuf& keys, DevBuf& vals, int64_t n, unsigned end_bit) {
    return sort_pairs_swap_impl<uint32_t, int64_t>(keys, vals, n, end_bit);
}

int sort_pairs_u32_u32_swap(DevBuf& keys, DevBuf& vals, int64_t n, unsigned end_bit) {
    return sort_pairs_swap_impl<uint32_t, uint32_t>(keys, vals, n, end_bit);
}

int sort_pairs_u64_f32_swap(DevBuf& keys, DevBuf& vals, int64_t n, unsigned end_bit) {
    return sort_pairs_swap_impl<uint64_t, float>(keys, vals, n, end_bit);
}

namespace {

__global__ void k_gather_u8(const uint8_t* __restrict__ src, const int64_t* __restrict__ idx,
                            uint8_t* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

}  // namespace

// Sorts (value, id) lexicographically: stable LSD passes by id, then by value;
// payload (may be null) follows its element.
int sort_by_value_then_id(float* value, int64_t* id, int64_t n, uint8_t* payload) {
    if (n <= 1) return RNNG_OK;
    LocalBuf perm, keys64, bits, tmpv, tmpi;
    RNNG_TRY(perm.reserve((size_t)n * sizeof(int64_t)));
    RNNG_TRY(keys64.reserve((size_t)n * sizeof(uint64_t)));
    RNNG_TRY(bits.reserve((size_t)n * sizeof(uint32_t)));
    k_iota_i64<<<blocks_for(n, 256), 256, 0, stream()>>>(perm.as<int64_t>(), n);
    RNNG_HIP_TRY(hipMemcpyAsync(keys64.ptr, id, (size_t)n * sizeof(int64_t), hipMemcpyDeviceToDevice,
                                stream()));
    // ids are non-negative, so their int64 order is the uint64 order
    RNNG_TRY((sort_pairs_impl<uint64_t, int64_t>(keys64.as<uint64_t>(), perm.as<int64_t>(), n, 64)));
    RNNG_TRY(tmpv.reserve((size_t)n * sizeof(float)));
    k_gather_f32<<<blocks_for(n, 256), 256, 0, stream()>>>(value, perm.as<int64_t>(), tmpv.as<float>(), n);
    k_float_bits<<<blocks_for(n, 256), 256, 0, stream()>>>(tmpv.as<float>(), bits.as<uint32_t>(), n);
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_TRY((sort_pairs_impl<uint32_t, int64_t>(bits.as<uint32_t>(), perm.as<int64_t>(), n, 32)));
    keys64.release();
    bits.release();
    RNNG_TRY(tmpi.reserve((size_t)n * sizeof(int64_t)));
    k_gather_f32<<<blocks_for(n, 256), 256, 0, stream()>>>(value, perm.as<int64_t>(), tmpv.as<float>(), n);
    k_gather_i64<<<blocks_for(n, 256), 256, 0, stream()>>>(id, perm.as<int64_t>(), tmpi.as<int64_t>(), n);
    RNNG_HIP_TRY(hipGetLastError());
    RNNG_HIP_TRY(hipMemcpyAsync(value, tmpv.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, stream()));
    RNNG_HIP_TRY(hipMemcpyAsync(id, tmpi.ptr, (size_t)n * sizeof(int64_t), hipMemcpyDeviceToDevice, stream()));
    if (payload) {
        LocalBuf tmpp;
        RNNG_TRY(tmpp.reserve((size_t)n));
        k_gather_u8<<<blocks_for(n, 256), 256, 0, stream()>>>(payload, perm.as<int64_t>(), tmpp.as<uint8_t>(), n);
        RNNG_HIP_TRY(hipGetLastError());
        RNNG_HIP_TRY(hipMemcpyAsync(payload, tmpp.ptr, (size_t)n, hipMemcpyDeviceToDevice, stream()));
        RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    }
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

int exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n) {
    if (n <= 0) return RNNG_OK;
    size_t tb = 0;
    RNNG_HIP_TRY(rocprim::exclusive_scan(nullptr, tb, in, out, (int64_t)0, (size_t)n,
                                         rocprim::plus<int64_t>(), stream()));
    LocalBuf tmp;
    RNNG_TRY(tmp.reserve(tb));
    RNNG_HIP_TRY(rocprim::exclusive_scan(tmp.ptr, tb, in, out, (int64_t)0, (size_t)n,
                                         rocprim::plus<int64_t>(), stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    return RNNG_OK;
}

int compact_edges(const int2* ij, const float* w, const uint8_t* flag, int64_t n, int2* out_ij,
                  float* out_w, int64_t* n_out) {
    *n_out = 0;
    if (n <= 0) return RNNG_OK;
    LocalBuf cnt, tmp;
    RNNG_TRY(cnt.reserve(2 * sizeof(unsigned long long)));
    size_t tb1 = 0, tb2 = 0;
    unsigned long long* c = cnt.as<unsigned long long>();
    RNNG_HIP_TRY(rocprim::select(nullptr, tb1, ij, flag, out_ij, c, (size_t)n, stream()));
    RNNG_HIP_TRY(rocprim::select(nullptr, tb2, w, flag, out_w, c + 1, (size_t)n, stream()));
    RNNG_TRY(tmp.reserve(std::max(tb1, tb2)));
    RNNG_HIP_TRY(rocprim::select(tmp.ptr, tb1, ij, flag, out_ij, c, (size_t)n, stream()));
    RNNG_HIP_TRY(rocprim::select(tmp.ptr, tb2, w, flag, out_w, c + 1, (size_t)n, stream()));
    unsigned long long h[2] = {0, 0};
    RNNG_HIP_TRY(hipMemcpyAsync(h, c, sizeof(h), hipMemcpyDeviceToHost, stream()));
    RNNG_HIP_TRY(hipStreamSynchronize(stream()));
    if (h[0] != h[1]) {
        set_error("compact_edges: inconsistent select counts");
        return RNNG_ERR_HIP;
    }
    *n_out = (int64_t)h[0];
    return RNNG_OK;
}

}  // namespace rnng
