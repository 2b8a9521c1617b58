// ffbf_batch.hip -- batched scan of many texts against many resident filters
// in one persistent launch, and the matching verification kernel.
//
// The reference scans one text against one filter per run
// (greptest.cu:505-576).  Here the texts of a batch sit back to back in one
// device buffer, each routed to one filter; the host cuts every text into
// chunks of wave-iterations (62 units of 16 bytes, or 126 with two units
// per lane), ordered by filter, and persistent workgroups
// pull chunks from a global counter.  A workgroup reloads its LDS prefilter
// only when the next chunk belongs to another filter, and drains its survivor
// queues before it does.  Filters whose prefilter stays in global memory are
// scanned by a second launch of small, many-wave blocks.  Candidates carry their text id in the top bits of
// the offset.  Per text, the windows evaluated, the hash pairs, the probe
// cascade and the feed-forward are exactly those of the single-text scan.
#include "ffbf_scan_common.h"

namespace ffbf {

namespace {

// A record read at a wave-uniform address, kept in scalar registers: every
// 32-bit word is taken from the first active lane.
template <class T>
__device__ __forceinline__ T wave_uniform(const T &v)
{
    static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
    T out;
    const uint32_t *s = reinterpret_cast<const uint32_t *>(&v);
    uint32_t *d = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); i++) d[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int)s[i]);
    return out;
}

// GLOBAL_ONLY: every filter of the launch keeps its prefilter in global
// memory (F, or an L2-sized fold); no LDS is reserved for one, so small
// blocks fit many waves per CU to hide the probes' round trips.
// SPLIT: filters of the CPU binary's method (64-bit hash from a (lo, hi)
// sbox, windows holding a newline never tested, cache-part fold).
// UPL: 16-byte units per lane (2: 32 window starts per lane, chunks counted
// in wave-iterations of kIterUnits2 units; compile-time W, pair method, LDS
// class only).
template <int WT, int BLOCK, int P, bool GLOBAL_ONLY, bool SPLIT = false, int UPL = 1>
__global__ __launch_bounds__(BLOCK) void k_scan_batch(BatchArgs b)
{
    static_assert(UPL == 1 || (UPL == 2 && WT > 0 && !SPLIT && !GLOBAL_ONLY), "two units per lane: LDS class, pair");
    using QG = QueueGeom<BLOCK>;
    constexpr uint32_t kSboxWords = SPLIT ? 512u : 256u;
    extern __shared__ __align__(16) uint32_t s_pre[];
    __shared__ __align__(8) uint32_t s_sbox[kSboxWords];
    __shared__ uint4 s_q[BLOCK / 64][QG::cap];
    __shared__ uint32_t s_next;

    for (uint32_t i = threadIdx.x; i < kSboxWords; i += BLOCK) s_sbox[i] = b.sbox[i];

    const uint32_t W = WT ? (uint32_t)WT : b.W;
    const int wv = (int)(threadIdx.x >> 6);
    const uint32_t n_chunks = (uint32_t)b.ctl[1];
    const uint32_t chunk0 = (uint32_t)b.ctl[2];
    uint4 *q = s_q[wv];
    uint32_t qn = 0;
    PendingCand pend;
    uint32_t loaded = 0xFFFFFFFFu;
    ScanArgs a{};
    // the LDS-class launch only gets chunks of filters whose prefilter is
    // held in LDS, the other launch only the rest: the probe path is fixed
    // at compile time (one path's registers, not both)
    constexpr bool lds = !GLOBAL_ONLY;

    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_next = (uint32_t)atomicAdd(b.ctl, 1ull);
        __syncthreads();
        const uint32_t ci = __builtin_amdgcn_readfirstlane(s_next);
        if (ci >= n_chunks) break;
        const BatchChunk c = wave_uniform(b.chunks[chunk0 + ci]);
        if (c.filter != loaded) {
            if (qn) {
                wave_sync();
                drain<lds, lds ? P : 1, SPLIT>(q, qn, a, s_pre, pend);
                wave_sync();
                qn = 0;
            }
            if (loaded != 0xFFFFFFFFu) flush_pending(pend, a);
            __syncthreads();
            // the previous filter's reverse bits gathered in LDS (every wave
            // has drained) go out before its prefilter is overwritten
            if (lds && !SPLIT && loaded != 0xFFFFFFFFu && a.r_lds) flush_lds_r(s_pre + a.pre_words, a.pre_words, a.R);
            a = wave_uniform(b.filters[c.filter]);
            __syncthreads();
            if constexpr (lds) {
                const uint4 *src = (const uint4 *)a.pre;
                uint4 *dst = (uint4 *)s_pre;
                for (uint32_t i = threadIdx.x; i < a.pre_words / 4u; i += BLOCK) dst[i] = src[i];
                if (!SPLIT && a.r_lds) zero_lds(s_pre + a.pre_words, a.pre_words);
            }
            __syncthreads();
            loaded = c.filter;
        }
        const BatchText t = wave_uniform(b.tdesc[c.text]);
        const uint8_t *text = b.texts + t.start;
        const uint64_t lead = (uint64_t)((uintptr_t)text & 15u);
        const uint8_t *base = text - lead;
        const uint64_t npos = t.n >= W ? t.n - W + 1 : 0;
        const uint64_t tag = (uint64_t)c.text << kBatchOffsetBits;
        const uint64_t it_end = (uint64_t)c.it0 + c.n_it;
        uint64_t it = (uint64_t)c.it0 + (uint64_t)wv;

        uint4 cur = make_uint4(0u, 0u, 0u, 0u);
        if constexpr (UPL == 2) {
            // no load ahead: the 8 registers it holds would make the P = 3
            // instances spill at 1024 threads
            for (; it < it_end; it += BLOCK / 64) {
                const uint64_t u = it * kIterUnits2 + 2u * (uint64_t)lane_id();
                const uint4 u0 = load_chunk(base, lead, t.n, u);
                const uint4 u1 = load_chunk(base, lead, t.n, u + 1u);
                HashPair32<(WT > 0 ? WT : 1)> hp;
                hp.load(u0, u1, s_sbox);
                uint32_t h1[kUnit], h2[kUnit];
                hp.first(h1, h2);
                const int64_t p0 = (int64_t)(u * 16u) - (int64_t)lead;
                test_and_queue<true, BLOCK, P, false, 63>(a, s_pre, h1, h2, p0, npos, tag, q, qn, pend);
                hp.second(h1, h2);
                test_and_queue<true, BLOCK, P, false, 63>(a, s_pre, h1, h2, p0 + 16, npos, tag, q, qn, pend);
            }
            continue;
        }
        if constexpr (WT > 0) {
            if (it < it_end) cur = load_chunk(base, lead, t.n, it * kIterUnits + (uint64_t)lane_id());
        }
        for (; it < it_end; it += BLOCK / 64) {
            const uint64_t u = it * kIterUnits + (uint64_t)lane_id();
            uint32_t h1[kUnit], h2[kUnit];
            uint32_t line_ok = 0xFFFFu;
            if constexpr (WT > 0) {
                uint4 nxt = make_uint4(0u, 0u, 0u, 0u);
                if (it + BLOCK / 64 < it_end)
                    nxt = load_chunk(base, lead, t.n, (it + BLOCK / 64) * kIterUnits + (uint64_t)lane_id());
                if constexpr (SPLIT)
                    line_ok = hash_unit64<(WT > 0 ? WT : 1)>(cur, (const uint2 *)s_sbox, h1, h2);
                else
                    hash_unit<(WT > 0 ? WT : 1)>(cur, s_sbox, h1, h2);
                cur = nxt;
            } else {
                if constexpr (SPLIT)
                    line_ok = hash_unit_generic64(base, lead, t.n, u, W, (const uint2 *)s_sbox, h1, h2);
                else
                    hash_unit_generic(base, lead, t.n, u, W, s_sbox, h1, h2);
            }
            const int64_t p0 = (int64_t)(u * 16u) - (int64_t)lead;
            test_and_queue<lds, BLOCK, lds ? P : 1, SPLIT>(a, s_pre, h1, h2, p0, npos, tag, q, qn, pend, line_ok);
        }
    }
    if (qn) {
        wave_sync();
        drain<lds, lds ? P : 1, SPLIT>(q, qn, a, s_pre, pend);
    }
    if (loaded != 0xFFFFFFFFu) finish_pending(pend, a);
    if (lds && !SPLIT && loaded != 0xFFFFFFFFu && a.r_lds) {
        __syncthreads();
        flush_lds_r(s_pre + a.pre_words, a.pre_words, a.R);
    }
}

// Exact verification of a batch: the candidate's text id selects the text
// and, through it, the filter whose survivor index and patterns are used.
__global__ __launch_bounds__(256) void k_verify_batch(BatchVerifyArgs v)
{
    const unsigned long long total = v.counters[C_CVALID];
    const uint64_t count = total < v.cand_cap ? total : v.cand_cap;
    const uint64_t n_texts = v.ctl[3];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < count; c += stride) {
        const Cand e = v.cand[c];
        const uint64_t tid = e.offset >> kBatchOffsetBits;
        const uint64_t off = e.offset & ((1ull << kBatchOffsetBits) - 1ull);
        if (tid >= n_texts) continue;  // never for a recorded candidate
        const BatchText t = v.tdesc[tid];
        const VerifyArgs &f = v.filters[t.filter];
        verify_candidate(f, v.texts + t.start, t.n, off, e.h1, e.h2, [&](uint32_t id) {
            const unsigned long long s = atomicAdd(v.counters + C_MATCH, 1ull);
            if (s < v.out_cap) {
                BatchMatchRec m;
                m.text = (uint32_t)tid;
                m.id = id;
                m.offset = off;
                v.out[s] = m;
            }
        });
    }
}

template <int WT, int BLOCK, int P, bool GLOBAL_ONLY = false, bool SPLIT = false, int UPL = 1>
hipError_t launch_batch_t(const BatchArgs &a, uint32_t pre_words_max, int n_cu, hipStream_t s)
{
    const size_t dyn = GLOBAL_ONLY ? 0u : (size_t)pre_words_max * 4u;
    if (dyn > 0) {
        hipError_t e = hipFuncSetAttribute((const void *)k_scan_batch<WT, BLOCK, P, GLOBAL_ONLY, SPLIT, UPL>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
    }
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, k_scan_batch<WT, BLOCK, P, GLOBAL_ONLY, SPLIT, UPL>, BLOCK, dyn);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((k_scan_batch<WT, BLOCK, P, GLOBAL_ONLY, SPLIT, UPL>), dim3((uint32_t)(n_cu * per_cu)),
                       dim3(BLOCK), dyn, s, a);
    return hipGetLastError();
}

template <int WT, int BLOCK, int UPL>
hipError_t launch_batch_p(const BatchArgs &a, uint32_t probes, uint32_t pre_words_max, int n_cu, hipStream_t s)
{
    if (probes <= 1u) return launch_batch_t<WT, BLOCK, 1, false, false, UPL>(a, pre_words_max, n_cu, s);
    if (probes == 2u) return launch_batch_t<WT, BLOCK, 2, false, false, UPL>(a, pre_words_max, n_cu, s);
    return launch_batch_t<WT, BLOCK, 3, false, false, UPL>(a, pre_words_max, n_cu, s);
}

template <int WT>
hipError_t launch_batch_w(const BatchArgs &a, uint32_t block, uint32_t probes, uint32_t upl, uint32_t pre_words_max,
                          int n_cu, hipStream_t s)
{
    if constexpr (WT > 0) {
        if (upl == 2u) {
            if (block == 512u) return launch_batch_p<WT, 512, 2>(a, probes, pre_words_max, n_cu, s);
            return launch_batch_p<WT, 1024, 2>(a, probes, pre_words_max, n_cu, s);
        }
    }
    if (block == 512u) return launch_batch_p<WT, 512, 1>(a, probes, pre_words_max, n_cu, s);
    return launch_batch_p<WT, 1024, 1>(a, probes, pre_words_max, n_cu, s);
}

// Split filters: 1024-thread blocks for the LDS class (any requested block
// size), the window length of the CPU binary (19) or a run-time one.
template <int WT>
hipError_t launch_batch_split(const BatchArgs &a, uint32_t probes, uint32_t pre_words_max, int n_cu, hipStream_t s)
{
    if (probes <= 1u) return launch_batch_t<WT, 1024, 1, false, true>(a, pre_words_max, n_cu, s);
    if (probes == 2u) return launch_batch_t<WT, 1024, 2, false, true>(a, pre_words_max, n_cu, s);
    return launch_batch_t<WT, 1024, 3, false, true>(a, pre_words_max, n_cu, s);
}

}  // namespace

hipError_t launch_scan_batch(const BatchArgs &a, uint32_t block, uint32_t probes, uint32_t units_per_lane,
                             uint32_t pre_words_max, int n_cu, hipStream_t s)
{
    if (a.method == METHOD_SPLIT)
        return a.W == 19 ? launch_batch_split<19>(a, probes, pre_words_max, n_cu, s)
                         : launch_batch_split<0>(a, probes, pre_words_max, n_cu, s);
    switch (a.W) {
    case 19: return launch_batch_w<19>(a, block, probes, units_per_lane, pre_words_max, n_cu, s);
    case 24: return launch_batch_w<24>(a, block, probes, units_per_lane, pre_words_max, n_cu, s);
    case 32: return launch_batch_w<32>(a, block, probes, units_per_lane, pre_words_max, n_cu, s);
    default: return launch_batch_w<0>(a, block, probes, units_per_lane, pre_words_max, n_cu, s);
    }
}

hipError_t launch_scan_batch_global(const BatchArgs &a, int n_cu, hipStream_t s)
{
    if (a.method == METHOD_SPLIT)
        return a.W == 19 ? launch_batch_t<19, 256, 1, true, true>(a, 0u, n_cu, s)
                         : launch_batch_t<0, 256, 1, true, true>(a, 0u, n_cu, s);
    switch (a.W) {
    case 19: return launch_batch_t<19, 256, 1, true>(a, 0u, n_cu, s);
    case 24: return launch_batch_t<24, 256, 1, true>(a, 0u, n_cu, s);
    case 32: return launch_batch_t<32, 256, 1, true>(a, 0u, n_cu, s);
    default: return launch_batch_t<0, 256, 1, true>(a, 0u, n_cu, s);
    }
}

hipError_t launch_verify_batch(const BatchVerifyArgs &a, int n_cu, hipStream_t s)
{
    hipLaunchKernelGGL(k_verify_batch, dim3((uint32_t)(n_cu * 4)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace ffbf
