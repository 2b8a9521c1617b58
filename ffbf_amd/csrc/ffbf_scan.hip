// ffbf_scan.hip -- text scan with feed-forward (GrepKernel, greptest.cu:125-251)
// for gfx950, plus the random-gather microbenchmark that bounds it.
//
// Work layout: the text is cut into 16-byte units aligned to 16 bytes in
// memory; one lane owns the 16 windows that start in one unit, so a wave
// covers 62 units (992 bytes) of window starts per iteration (its top two
// lanes load the units the windows run into) and persistent waves stride over
// the text.  The LDS-prefilter instances at W 19, 24 and 32 give a lane two
// adjacent units instead (32 windows, 126 units per wave-iteration, lane 63
// only completes lane 62's windows), paying the first window's full hash once
// per 32 windows.  A lane loads its unit with one 128-bit non-temporal load (the
// next iteration's unit is loaded while the current one is hashed), looks its
// bytes up in an LDS copy of the sbox, takes the sbox values of the following
// bytes from the lanes above, and computes its 16 hash pairs as one full
// window plus 15 rolling slides (hash_rot_sbox_2.h:34-51).
//
// Each window is first tested against a prefilter: P forward bits of either
// F itself or an OR-fold of F (a fold bit is the OR of all F bits congruent
// to it, so a window whose forward bits are all set in F has them set in the
// fold: the test never rejects a true candidate).  The prefilter lives in
// LDS (fold or whole filter) or in global memory (F, or a fold small enough
// to stay in L2), or both: an LDS fold in front of the L2-sized fold for a
// filter larger than L2.  Windows that pass are pushed onto a per-wave LDS queue;
// 64 at a time, one per lane, they finish their forward probes, feed their
// reverse bits forward into R and are appended to the candidate list.
#include "ffbf_scan_common.h"

namespace ffbf {

namespace {

template <int WT, bool LDSPRE, int BLOCK, int P, bool SPLIT = false, int UPL = 1>
__global__ __launch_bounds__(BLOCK) void k_scan(ScanArgs a)
{
    static_assert(UPL == 1 || (UPL == 2 && WT > 0 && !SPLIT), "two units per lane: compile-time W, pair method");
    constexpr uint32_t IU = UPL == 2 ? kIterUnits2 : kIterUnits;
    using QG = QueueGeom<BLOCK>;
    constexpr uint32_t kSboxWords = SPLIT ? 512u : 256u;  // split: 256 (lo, hi) pairs
    extern __shared__ __align__(16) uint32_t s_pre[];
    __shared__ __align__(8) uint32_t s_sbox[kSboxWords];
    __shared__ uint4 s_q[BLOCK / 64][QG::cap];

    for (uint32_t i = threadIdx.x; i < kSboxWords; i += BLOCK) s_sbox[i] = a.sbox[i];
    if constexpr (LDSPRE) {
        const uint4 *src = (const uint4 *)a.pre;
        uint4 *dst = (uint4 *)s_pre;
        for (uint32_t i = threadIdx.x; i < a.pre_words / 4u; i += BLOCK) dst[i] = src[i];
        if (a.r_lds) zero_lds(s_pre + a.pre_words, a.pre_words);
    }
    __syncthreads();

    const uint32_t W = WT ? (uint32_t)WT : a.W;
    const int wv = (int)(threadIdx.x >> 6);
    const uint64_t lead = (uint64_t)((uintptr_t)a.text & 15u);
    const uint8_t *base = a.text - lead;
    const uint64_t npos = a.n >= W ? a.n - W + 1 : 0;
    const uint64_t units = (lead + npos + 15u) >> 4;
    const uint64_t iters = (units + IU - 1u) / IU;
    const uint64_t n_waves = (uint64_t)gridDim.x * (BLOCK / 64);
    uint4 *q = s_q[wv];
    uint32_t qn = 0;  // wave-uniform queue depth
    PendingCand pend;

    uint4 cur = make_uint4(0u, 0u, 0u, 0u);
    uint64_t it = (uint64_t)blockIdx.x * (BLOCK / 64) + wv;
    if constexpr (UPL == 2) {
        // lane L owns units 2L and 2L+1 of the wave-iteration
        uint4 cur1 = make_uint4(0u, 0u, 0u, 0u);
        if (it < iters) {
            const uint64_t u = it * IU + 2u * (uint64_t)lane_id();
            cur = load_chunk(base, lead, a.n, u);
            cur1 = load_chunk(base, lead, a.n, u + 1u);
        }
        for (; it < iters; it += n_waves) {
            const uint64_t u = it * IU + 2u * (uint64_t)lane_id();
            const uint64_t nit = it + n_waves;
            uint4 nxt = make_uint4(0u, 0u, 0u, 0u), nxt1 = nxt;
            if (nit < iters) {
                nxt = load_chunk(base, lead, a.n, nit * IU + 2u * (uint64_t)lane_id());
                nxt1 = load_chunk(base, lead, a.n, nit * IU + 2u * (uint64_t)lane_id() + 1u);
            }
            HashPair32<(WT > 0 ? WT : 1)> hp;
            hp.load(cur, cur1, s_sbox);
            uint32_t h1[kUnit], h2[kUnit];
            hp.first(h1, h2);
            const int64_t p0 = (int64_t)(u * 16u) - (int64_t)lead;
            test_and_queue<LDSPRE, BLOCK, P, false, 63>(a, s_pre, h1, h2, p0, npos, a.off_base, q, qn, pend);
            hp.second(h1, h2);
            test_and_queue<LDSPRE, BLOCK, P, false, 63>(a, s_pre, h1, h2, p0 + 16, npos, a.off_base, q, qn, pend);
            cur = nxt;
            cur1 = nxt1;
        }
        if (qn) {
            wave_sync();
            drain<LDSPRE, P, false>(q, qn, a, s_pre, pend);
        }
        finish_pending(pend, a);
        if (LDSPRE && a.r_lds) {
            __syncthreads();
            flush_lds_r(s_pre + a.pre_words, a.pre_words, a.R);
        }
        return;
    }
    if constexpr (WT > 0) {
        if (it < iters) cur = load_chunk(base, lead, a.n, it * kIterUnits + (uint64_t)lane_id());
    }

    for (; it < iters; it += n_waves) {
        const uint64_t u = it * kIterUnits + (uint64_t)lane_id();
        uint32_t h1[kUnit], h2[kUnit];
        uint32_t line_ok = 0xFFFFu;
        if constexpr (WT > 0) {
            // issue the next iteration's load before hashing this one
            const uint64_t nit = it + n_waves;
            uint4 nxt = make_uint4(0u, 0u, 0u, 0u);
            if (nit < iters) nxt = load_chunk(base, lead, a.n, nit * kIterUnits + (uint64_t)lane_id());
            if constexpr (SPLIT)
                line_ok = hash_unit64<(WT > 0 ? WT : 1)>(cur, (const uint2 *)s_sbox, h1, h2);
            else
                hash_unit<(WT > 0 ? WT : 1)>(cur, s_sbox, h1, h2);
            cur = nxt;
        } else {
            if constexpr (SPLIT)
                line_ok = hash_unit_generic64(base, lead, a.n, u, W, (const uint2 *)s_sbox, h1, h2);
            else
                hash_unit_generic(base, lead, a.n, u, W, s_sbox, h1, h2);
        }

        // window offset of slot b is 16u + b - lead
        test_and_queue<LDSPRE, BLOCK, P, SPLIT>(a, s_pre, h1, h2, (int64_t)(u * 16u) - (int64_t)lead, npos,
                                                a.off_base, q, qn, pend, line_ok);
    }
    if (qn) {
        wave_sync();
        drain<LDSPRE, P, SPLIT>(q, qn, a, s_pre, pend);
    }
    finish_pending(pend, a);
    if (LDSPRE && !SPLIT && a.r_lds) {
        __syncthreads();
        flush_lds_r(s_pre + a.pre_words, a.pre_words, a.R);
    }
}

template <int WT, bool LDSPRE, int BLOCK, int P, bool SPLIT = false, int UPL = 1>
hipError_t launch_scan_t(const ScanArgs &a, int n_cu, hipStream_t s, uint32_t *waves)
{
    const size_t dyn = LDSPRE ? (size_t)a.pre_words * 4u * (a.r_lds ? 2u : 1u) : 0;
    if (dyn > 0) {
        hipError_t e = hipFuncSetAttribute((const void *)k_scan<WT, LDSPRE, BLOCK, P, SPLIT, UPL>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
    }
    int per_cu = 0;
    hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_scan<WT, LDSPRE, BLOCK, P, SPLIT, UPL>, BLOCK, dyn);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    const uint64_t W = WT ? (uint64_t)WT : a.W;
    const uint64_t npos = a.n >= W ? a.n - W + 1 : 0;
    const uint64_t IU = UPL == 2 ? kIterUnits2 : kIterUnits;
    const uint64_t iters = ((npos + 31u) / 16u + IU - 1u) / IU;
    uint64_t grid = (uint64_t)n_cu * (uint64_t)per_cu;
    const uint64_t need = (iters + (BLOCK / 64) - 1) / (BLOCK / 64);
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    if (waves) *waves = (uint32_t)grid * (uint32_t)(BLOCK / 64);
    hipLaunchKernelGGL((k_scan<WT, LDSPRE, BLOCK, P, SPLIT, UPL>), dim3((uint32_t)grid), dim3(BLOCK), dyn, s, a);
    return hipGetLastError();
}

// Split method: the cache-part fold in LDS (1024-thread blocks, P <= 3 bits
// up front) or F's first bit read directly (256-thread blocks).
template <int WT>
hipError_t launch_scan_split_w(const ScanArgs &a, uint32_t mode, int n_cu, hipStream_t s, uint32_t *waves)
{
    if (mode == MODE_FOLD) {
        const uint32_t P = a.P < 1u ? 1u : (a.P > 3u ? 3u : a.P);
        if (P == 1) return launch_scan_t<WT, true, 1024, 1, true>(a, n_cu, s, waves);
        if (P == 2) return launch_scan_t<WT, true, 1024, 2, true>(a, n_cu, s, waves);
        return launch_scan_t<WT, true, 1024, 3, true>(a, n_cu, s, waves);
    }
    return launch_scan_t<WT, false, 256, 1, true>(a, n_cu, s, waves);
}

// LDS prefilter: 1024-thread blocks (512 where a smaller LDS forces it), P
// bits tested up front; global prefilter: 256-thread blocks, one bit.
template <int WT>
hipError_t launch_scan_w(const ScanArgs &a, uint32_t mode, uint32_t block, int n_cu, hipStream_t s, uint32_t *waves)
{
    const bool lds = mode_in_lds(mode);
    if (lds) {
        const uint32_t P = a.P < 1u ? 1u : (a.P > 3u ? 3u : a.P);
        if (block == 512) {
            if (P == 1) return launch_scan_t<WT, true, 512, 1>(a, n_cu, s, waves);
            if (P == 2) return launch_scan_t<WT, true, 512, 2>(a, n_cu, s, waves);
            return launch_scan_t<WT, true, 512, 3>(a, n_cu, s, waves);
        }
        constexpr int U = WT > 0 ? 2 : 1;
        if (P == 1) return launch_scan_t<WT, true, 1024, 1, false, U>(a, n_cu, s, waves);
        if (P == 2) return launch_scan_t<WT, true, 1024, 2, false, U>(a, n_cu, s, waves);
        return launch_scan_t<WT, true, 1024, 3, false, U>(a, n_cu, s, waves);
    }
    return launch_scan_t<WT, false, 256, 1>(a, n_cu, s, waves);
}

// Random dword gathers into a 2^log2_words-word array: 16 independent loads
// per lane per step, addresses from a counter hash.
// Random-access microbenchmark: 16 independent accesses per lane per step at
// hashed word addresses of arr[0 .. mask]. KIND 0: plain loads; 1:
// non-temporal loads; 2: relaxed atomic loads at agent scope (they bypass
// the CU's L1); 3: the same at system scope; 4: device-scope atomicOr with
// the result unused (the feed-forward's and the build's operation).
// Streaming pass counterpart: `steps` passes over the whole array, each
// reading and writing back every 16-byte word (a thread keeps the same words
// in every pass).
__global__ __launch_bounds__(256) void k_stream(uint4 *a, uint64_t n16, uint32_t steps)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint32_t s = 0; s < steps; s++)
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
            uint4 v = a[i];
            v.x ^= s + 1u;
            a[i] = v;
        }
}

template <int KIND>
__global__ __launch_bounds__(256) void k_gather(uint32_t *arr, uint32_t mask, uint32_t steps, uint32_t *out)
{
    uint32_t x = mix32((uint32_t)(blockIdx.x * 256u + threadIdx.x) * 0x9E3779B9u + 1u);
    uint32_t acc = 0;
    for (uint32_t s = 0; s < steps; s++) {
        uint32_t v[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint32_t i = mix32(x + (uint32_t)j * 0x61C88647u) & mask;
            if constexpr (KIND == 0) v[j] = arr[i];
            else if constexpr (KIND == 1) v[j] = __builtin_nontemporal_load(arr + i);
            else if constexpr (KIND == 2) v[j] = __hip_atomic_load(arr + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else if constexpr (KIND == 3) v[j] = __hip_atomic_load(arr + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            else {
                atomicOr(arr + i, 1u << (x & 31u));
                v[j] = i;
            }
        }
#pragma unroll
        for (int j = 0; j < 16; j++) acc ^= v[j];
        x = mix32(x ^ acc);
    }
    out[blockIdx.x * 256u + threadIdx.x] = acc;
}

}  // namespace

bool scan_has_fast_path(uint32_t W) { return W == 19 || W == 24 || W == 32; }

uint32_t max_fold_words(int device, bool split)
{
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || lds <= 0)
        lds = 64 * 1024;
    // static LDS of the 512-thread scan: sbox 1 KiB + 8 queues of 3.75 KiB;
    // the split method's 1024-thread scans: sbox 2 KiB + 16 queues of 119
    // entries, plus the batched scan's chunk word (with room for alignment)
    const int static_lds = split ? 2048 + 16 * QueueGeom<1024>::cap * 16 + 64 : 1024 + 8 * QueueGeom<512>::cap * 16;
    uint32_t words = 1u << 15;  // 128 KiB
    while (words > 256u && (int)(words * 4u) + static_lds > lds) words >>= 1;
    return words;
}

hipError_t launch_scan(const ScanArgs &a, uint32_t mode, uint32_t block, int n_cu, hipStream_t s, uint32_t *waves)
{
    if (a.method == METHOD_SPLIT) {
        if (mode != MODE_FOLD && mode != MODE_GLOBAL) return hipErrorInvalidValue;
        return a.W == 19 ? launch_scan_split_w<19>(a, mode, n_cu, s, waves) : launch_scan_split_w<0>(a, mode, n_cu, s, waves);
    }
    switch (a.W) {
    case 19: return launch_scan_w<19>(a, mode, block, n_cu, s, waves);
    case 24: return launch_scan_w<24>(a, mode, block, n_cu, s, waves);
    case 32: return launch_scan_w<32>(a, mode, block, n_cu, s, waves);
    default: return launch_scan_w<0>(a, mode, block, n_cu, s, waves);
    }
}

hipError_t launch_gather(uint32_t *arr, uint32_t log2_words, uint32_t steps, uint32_t *out, uint32_t blocks,
                         uint32_t kind, hipStream_t s)
{
    const uint32_t mask = (1u << log2_words) - 1u;
    switch (kind) {
    case 0: hipLaunchKernelGGL(k_gather<0>, dim3(blocks), dim3(256), 0, s, arr, mask, steps, out); break;
    case 1: hipLaunchKernelGGL(k_gather<1>, dim3(blocks), dim3(256), 0, s, arr, mask, steps, out); break;
    case 2: hipLaunchKernelGGL(k_gather<2>, dim3(blocks), dim3(256), 0, s, arr, mask, steps, out); break;
    case 3: hipLaunchKernelGGL(k_gather<3>, dim3(blocks), dim3(256), 0, s, arr, mask, steps, out); break;
    case 4: hipLaunchKernelGGL(k_gather<4>, dim3(blocks), dim3(256), 0, s, arr, mask, steps, out); break;
    case 5:
        hipLaunchKernelGGL(k_stream, dim3(blocks), dim3(256), 0, s, (uint4 *)arr, ((uint64_t)mask + 1u) / 4u, steps);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ffbf
