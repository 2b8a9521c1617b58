// ffbf_sort.hip -- device sort of the verified match list by (pattern id, offset).
//
// Each match packs into one 64-bit key, id << 40 | offset (offsets below
// 2^40, ids below 2^24), which rocPRIM's radix sort orders on the bits in use.
#include <rocprim/device/device_radix_sort.hpp>

#include "ffbf_device.h"

namespace ffbf {

namespace {

constexpr int kOffsetBits = 40;

__global__ void k_pack(const MatchRec *m, uint64_t n, uint64_t *keys)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        keys[i] = ((uint64_t)m[i].id << kOffsetBits) | m[i].offset;
}

__global__ void k_unpack(const uint64_t *keys, uint64_t n, MatchRec *m)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        MatchRec r;
        r.id = (uint32_t)(keys[i] >> kOffsetBits);
        r.reserved = 0u;
        r.offset = keys[i] & ((1ull << kOffsetBits) - 1ull);
        m[i] = r;
    }
}

__global__ void k_pack_batch(const BatchMatchRec *m, uint64_t n, int id_shift, int text_shift, uint64_t *keys)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        keys[i] = ((uint64_t)m[i].text << text_shift) | ((uint64_t)m[i].id << id_shift) | m[i].offset;
}

__global__ void k_unpack_batch(const uint64_t *keys, uint64_t n, int id_shift, int text_shift, BatchMatchRec *m)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i];
        BatchMatchRec r;
        r.text = (uint32_t)(k >> text_shift);
        r.id = (uint32_t)((k >> id_shift) & ((1ull << (text_shift - id_shift)) - 1ull));
        r.offset = k & ((1ull << id_shift) - 1ull);
        m[i] = r;
    }
}

uint32_t grid_for(uint64_t n)
{
    uint64_t b = (n + 255u) / 256u;
    if (b > 4096u) b = 4096u;
    return b < 1u ? 1u : (uint32_t)b;
}

}  // namespace

bool sort_packable(uint64_t n_patterns, uint64_t text_bytes)
{
    return n_patterns <= (1ull << 24) && text_bytes <= (1ull << kOffsetBits);
}

int sort_end_bit(uint64_t n_patterns)
{
    int b = 0;
    while (b < 24 && (1ull << b) < n_patterns) b++;
    return kOffsetBits + (b ? b : 1);
}

hipError_t sort_temp_bytes(uint64_t n, int end_bit, size_t *bytes)
{
    return rocprim::radix_sort_keys(nullptr, *bytes, (uint64_t *)nullptr, (uint64_t *)nullptr, (size_t)n, 0, end_bit);
}

hipError_t sort_matches(MatchRec *m, uint64_t n, uint64_t *keys_a, uint64_t *keys_b, void *temp, size_t temp_bytes,
                        int end_bit, hipStream_t s)
{
    if (n < 2) return hipSuccess;
    hipLaunchKernelGGL(k_pack, dim3(grid_for(n)), dim3(256), 0, s, m, n, keys_a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t tb = temp_bytes;
    e = rocprim::radix_sort_keys(temp, tb, keys_a, keys_b, (size_t)n, 0, end_bit, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_unpack, dim3(grid_for(n)), dim3(256), 0, s, keys_b, n, m);
    return hipGetLastError();
}

// Batched matches sorted by (text, id, offset): key = text << text_shift |
// id << id_shift | offset, the caller guaranteeing that the fields fit.
hipError_t sort_batch_matches(BatchMatchRec *m, uint64_t n, uint64_t *keys_a, uint64_t *keys_b, void *temp,
                              size_t temp_bytes, int id_shift, int text_shift, int end_bit, hipStream_t s)
{
    if (n < 2) return hipSuccess;
    hipLaunchKernelGGL(k_pack_batch, dim3(grid_for(n)), dim3(256), 0, s, m, n, id_shift, text_shift, keys_a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t tb = temp_bytes;
    e = rocprim::radix_sort_keys(temp, tb, keys_a, keys_b, (size_t)n, 0, end_bit, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_unpack_batch, dim3(grid_for(n)), dim3(256), 0, s, keys_b, n, id_shift, text_shift, m);
    return hipGetLastError();
}

}  // namespace ffbf
