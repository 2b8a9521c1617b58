// ffbf_kernels.hip -- gfx950 kernels of the feed-forward Bloom filter path
// other than the scan (which lives in ffbf_scan.hip).
//
// Stages and the reference code whose results they reproduce bit for bit
// (efficient/ffbf):
//   k_build         prefix hash pairs of grepSetup  src/gpgpu/greptest.cu:96-123
//   k_build_atomic  the bits grepSetup sets (and, split, Filter::operator<<)
//   k_pattern_pass  filterPatterns   src/gpgpu/greptest.cu:253-290
//   k_verify        the exact step the reference leaves to `grep -Ff`
//                   (src/rabin-karp/README:49-52)
//   k_fold          OR-fold of F used by the scan as an exact-superset prefilter
//   k_build_split   Filter::operator<<'s prefix hashes  src/rabin-karp/ffbf.cpp:268-283
//   k_build_bin, k_build_slices  the same bits for filters of several 2^20-bit
//                   slices, set without global atomics
#include "ffbf_common.h"

namespace ffbf {

namespace {

__global__ void k_fill32(uint32_t *p, uint64_t n, uint32_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n4 = n / 4u;
    uint4 *p4 = (uint4 *)p;
    const uint4 v4 = make_uint4(v, v, v, v);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) p4[i] = v4;
    for (uint64_t i = n4 * 4u + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Plain 16-byte-per-lane device copy (measurement reference for HBM counters).
__global__ void k_copy(uint4 *dst, const uint4 *src, uint64_t n16)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// Calls step(c) for the bytes c of bytes[a0 .. a0+W) in order, read with
// aligned 16-byte loads (the pattern buffer is allocated with 64 bytes of
// padding, so the last load stays inside it).
template <class Step>
__device__ __forceinline__ void prefix_bytes(const uint8_t *bytes, uint64_t a0, uint32_t W, Step step)
{
    const uint64_t end = a0 + W;
    for (uint64_t q = a0 & ~15ull; q < end; q += 16u) {
        const uint4 v = *(const uint4 *)(bytes + q);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint32_t w = j < 4 ? v.x : j < 8 ? v.y : j < 12 ? v.z : v.w;
            const uint64_t pos = q + (uint64_t)j;
            if (pos >= a0 && pos < end) step((w >> (8 * (j & 3))) & 0xFFu);
        }
    }
}

// Prefix hash pairs, once per pattern set (the hashing half of grepSetup,
// greptest.cu:96-123): one lane per pattern hashes its W-byte prefix; a
// pattern shorter than W keeps (0, 0) and sets no bits (ffbf.cpp:271-275).
// The pairs feed both builds, the pattern pass and verification.
__global__ __launch_bounds__(256) void k_build(const uint8_t *bytes, const uint64_t *off, uint64_t n,
                                               const uint32_t *sbox, uint32_t W, uint2 *pat_hash)
{
    __shared__ uint32_t sb[256];
    for (uint32_t i = threadIdx.x; i < 256u; i += blockDim.x) sb[i] = sbox[i];
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t a0 = off[i], a1 = off[i + 1];
        uint32_t h1 = 0, h2 = 0;
        if (a1 - a0 >= W) {
            prefix_bytes(bytes, a0, W, [&](uint32_t c) {
                const uint32_t s = sb[c];
                h1 = rotl(h1, 1) ^ s;
                h2 = rotl(h2, 3) ^ s;
            });
        }
        pat_hash[i] = make_uint2(h1, h2);
    }
}

// The split method's prefix hashes (the hashing half of Filter::operator<<,
// ffbf.cpp:268-283): one lane per pattern hashes its W-byte prefix with the
// 64-bit loop form h = rol64(h,1) ^ S[c] (the ring update of
// hash_rot_sbox_pre_2.h:41-50 before the window is full; the sbox has bytes
// >= 0x80 mapped to entry 0).  Hash halves are kept as (h1, h2) = (low, high).
__global__ __launch_bounds__(256) void k_build_split(const uint8_t *bytes, const uint64_t *off, uint64_t n,
                                                     const uint2 *sbox, uint32_t W, uint2 *pat_hash)
{
    __shared__ uint2 sb[256];
    for (uint32_t i = threadIdx.x; i < 256u; i += blockDim.x) sb[i] = sbox[i];
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t a0 = off[i], a1 = off[i + 1];
        uint32_t lo = 0, hi = 0;
        if (a1 - a0 >= W) {
            prefix_bytes(bytes, a0, W, [&](uint32_t c) {
                const uint2 v = sb[c];
                uint32_t rl, rh;
                rol64(lo, hi, 1u, rl, rh);
                lo = rl ^ v.x;
                hi = rh ^ v.y;
            });
        }
        pat_hash[i] = make_uint2(lo, hi);
    }
}

// The atomic build (grepSetup, greptest.cu:96-123; Filter::operator<< with
// setBitsInFilter, ffbf.cpp:196-203 and 268-283, for the split method): one
// lane per pattern at least W long sets its k forward bits in a cleared F
// with atomicOr, from the prefix hash pair computed when the patterns were
// set; shorter patterns set none (ffbf.cpp:271-275).  The scan's OR-folds of
// F, when given, get the same bits at x mod their size (the bits k_fold would
// give them from the finished F).
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_build_atomic(const uint64_t *off, const uint2 *pat_hash, uint64_t n, uint32_t W,
                                                      uint32_t k, uint32_t mask, SplitMap m, uint32_t *F, BuildFolds fo)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (off[i + 1] - off[i] < W) continue;
        const uint2 h = pat_hash[i];
        for (int j = 0; j < (SPLIT ? kSplitMaxK : 5); j++) {
            if (j >= (int)k) break;
            const uint32_t x = SPLIT ? split_place(j, split_fwd_fn(j, h.x, h.y), m) : (fwd_idx(j, h.x, h.y) & mask);
            const uint32_t bit = 1u << (x & 31u);
            atomicOr(&F[x >> 5], bit);
            if ((uint64_t)x < fo.src_bits) {
                if (fo.p[0]) atomicOr(&fo.p[0][(x >> 5) & fo.wmask[0]], bit);
                if (fo.p[1]) atomicOr(&fo.p[1][(x >> 5) & fo.wmask[1]], bit);
            }
        }
    }
}

// Atomic-free build of a filter of several 2^20-bit slices.  Device-scope
// atomics run at a fixed ~27 G/s on this part whatever the array's size, so
// setting k bits per pattern with atomicOr costs k n / 27e9 s (160 us at 1M
// patterns and k = 4).  Instead k_build_bin takes kBinEntries / k consecutive
// patterns per block, forms their forward bit indices from the prefix hash
// pairs computed when the patterns were set, counting-sorts them by slice in
// LDS and writes them out in slice order, with the block's slice offsets;
// k_build_slices then gives each slice one workgroup, which sets that slice's
// bits from every block in an LDS copy (LDS atomics) and writes the slice
// whole, clearing what no pattern sets.  Bit placement is the same as
// k_build_atomic's.
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_build_bin(const uint64_t *off, const uint2 *pat_hash, uint64_t n, uint32_t P,
                                                   uint32_t W, uint32_t k, uint32_t mask, SplitMap m, uint32_t ns,
                                                   uint32_t *staging, uint32_t *offs)
{
    extern __shared__ uint32_t s_bin[];  // raw[kBinEntries], sorted[kBinEntries], hist[ns]
    uint32_t *raw = s_bin, *sorted = s_bin + kBinEntries, *hist = s_bin + 2u * kBinEntries;
    __shared__ uint32_t s_ne, s_wave[4];
    const uint64_t p0 = (uint64_t)blockIdx.x * P;
    const uint32_t np = (uint32_t)(n - p0 < (uint64_t)P ? n - p0 : (uint64_t)P);
    for (uint32_t t = threadIdx.x; t < ns; t += blockDim.x) hist[t] = 0u;
    if (threadIdx.x == 0) s_ne = 0u;
    __syncthreads();
    // the k bit indices of every pattern at least W long (shorter ones set
    // nothing, ffbf.cpp:271-275), in any order
    for (uint32_t t = threadIdx.x; t < np; t += blockDim.x) {
        const uint64_t i = p0 + t;
        if (off[i + 1] - off[i] < W) continue;
        const uint2 h = pat_hash[i];
        const uint32_t at = atomicAdd(&s_ne, k);
#pragma nounroll
        for (int j = 0; j < (SPLIT ? kSplitMaxK : 5); j++) {
            if (j >= (int)k) break;
            const uint32_t x = SPLIT ? split_place(j, split_fwd_fn(j, h.x, h.y), m) : (fwd_idx(j, h.x, h.y) & mask);
            raw[at + (uint32_t)j] = x;
            atomicAdd(&hist[x >> kBinSliceLog2], 1u);
        }
    }
    __syncthreads();
    const uint32_t ne = s_ne;
    // exclusive prefix of the slice counts, in place: each thread sums a run
    // of consecutive slices, a block scan of the run sums gives its start
    const uint32_t per = (ns + blockDim.x - 1u) / blockDim.x;
    const uint32_t lo = threadIdx.x * per < ns ? threadIdx.x * per : ns;
    const uint32_t hi = lo + per < ns ? lo + per : ns;
    uint32_t run = 0;
    for (uint32_t t = lo; t < hi; t++) run += hist[t];
    const uint32_t incl = wave_incl_sum(run);
    if (lane_id() == 63) s_wave[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint32_t start = incl - run;
    for (uint32_t w = 0; w < (threadIdx.x >> 6); w++) start += s_wave[w];
    uint32_t *bo = offs + (uint64_t)blockIdx.x * (ns + 1u);
    for (uint32_t t = lo; t < hi; t++) {
        const uint32_t c = hist[t];
        hist[t] = start;
        bo[t] = start;
        start += c;
    }
    if (threadIdx.x == 0) bo[ns] = ne;
    __syncthreads();
    // counting sort by slice, then one contiguous write
    for (uint32_t e = threadIdx.x; e < ne; e += blockDim.x) {
        const uint32_t x = raw[e];
        sorted[atomicAdd(&hist[x >> kBinSliceLog2], 1u)] = x;
    }
    __syncthreads();
    uint32_t *dst = staging + (uint64_t)blockIdx.x * kBinEntries;
    for (uint32_t e = threadIdx.x; e < ne; e += blockDim.x) dst[e] = sorted[e];
}

// One workgroup per 2^20-bit slice of F: the slice's bit indices from every
// block of k_build_bin set in LDS, then the slice written out whole.
__global__ __launch_bounds__(1024) void k_build_slices(const uint32_t *staging, const uint32_t *offs, uint32_t nb,
                                                       uint32_t ns, uint32_t *F, uint64_t f_words)
{
    extern __shared__ uint32_t s_slice[];  // 2^15 words
    constexpr uint32_t kWords = 1u << (kBinSliceLog2 - 5);
    const uint32_t sl = blockIdx.x;
    const uint64_t w0 = (uint64_t)sl * kWords;
    const uint32_t nw = (uint32_t)(f_words - w0 < (uint64_t)kWords ? f_words - w0 : (uint64_t)kWords);
    for (uint32_t i = threadIdx.x; i < nw; i += blockDim.x) s_slice[i] = 0u;
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
        const uint32_t *bo = offs + (uint64_t)b * (ns + 1u);
        const uint32_t a = bo[sl], e = bo[sl + 1u];
        const uint32_t *src = staging + (uint64_t)b * kBinEntries;
        uint32_t x = a;
        for (; x + 8u <= e; x += 8u) {
            uint32_t v[8];
#pragma unroll
            for (int i = 0; i < 8; i++) v[i] = src[x + (uint32_t)i];
#pragma unroll
            for (int i = 0; i < 8; i++) atomicOr(&s_slice[(v[i] >> 5) & (kWords - 1u)], 1u << (v[i] & 31u));
        }
        for (; x < e; x++) {
            const uint32_t v = src[x];
            atomicOr(&s_slice[(v >> 5) & (kWords - 1u)], 1u << (v & 31u));
        }
    }
    __syncthreads();
    uint32_t *dst = F + w0;
    if ((nw & 3u) == 0u && ((uintptr_t)dst & 15u) == 0u) {
        for (uint32_t i = threadIdx.x; i < nw / 4u; i += blockDim.x)
            ((uint4 *)dst)[i] = make_uint4(s_slice[4u * i], s_slice[4u * i + 1u], s_slice[4u * i + 2u], s_slice[4u * i + 3u]);
    } else {
        for (uint32_t i = threadIdx.x; i < nw; i += blockDim.x) dst[i] = s_slice[i];
    }
}

// OR-fold of F onto fold_words words: fold bit x is set iff some F bit
// congruent to x mod fold_words*32 is set, so a window whose k forward bits are
// set in F has its k folded bits set too (exact superset test).
__global__ void k_fold(const uint32_t *F, uint64_t f_words, uint32_t *fold, uint32_t fold_words)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ((fold_words & 3u) == 0u && f_words % fold_words == 0u && (((uintptr_t)F | (uintptr_t)fold) & 15u) == 0u) {
        // 16-byte columns: a thread ORs the r = f_words / fold_words
        // congruent 16-byte pieces of its column, eight loads in flight at a
        // time (the fold of a 2^23-word F onto 2^15 words has r = 256)
        const uint64_t n4 = fold_words / 4u, r = f_words / fold_words;
        const uint4 *F4 = (const uint4 *)F;
        for (uint64_t j = t0; j < n4; j += stride) {
            uint32_t x = 0, y = 0, z = 0, w = 0;
            uint64_t t = 0;
            for (; t + 8u <= r; t += 8u) {
                uint4 v[8];
#pragma unroll
                for (int i = 0; i < 8; i++) v[i] = F4[j + (t + (uint64_t)i) * n4];
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    x |= v[i].x;
                    y |= v[i].y;
                    z |= v[i].z;
                    w |= v[i].w;
                }
            }
            for (; t < r; t++) {
                const uint4 v = F4[j + t * n4];
                x |= v.x;
                y |= v.y;
                z |= v.z;
                w |= v.w;
            }
            ((uint4 *)fold)[j] = make_uint4(x, y, z, w);
        }
        return;
    }
    for (uint64_t j = t0; j < fold_words; j += stride) {
        uint32_t v = 0;
        for (uint64_t t = j; t < f_words; t += fold_words) v |= F[t];
        fold[j] = v;
    }
}

// Pattern pass (filterPatterns, greptest.cu:253-290): each pattern's prefix
// hash pair against R.  A wave owns 64 consecutive ids (i0 a multiple of 64),
// so its ballot is two whole bitmap words; survivors are compacted with a
// wave prefix count and indexed by prefix-hash key for verification.
// The split method tests its k (<= 11) reverse functions, placed in the
// cache or external part (checkInFilter_2, ffbf.cpp:241-266).
template <bool SPLIT = false>
__device__ __forceinline__ uint32_t pass_wave(const PassArgs &a, uint64_t i0)
{
    const int lane = lane_id();
    const uint64_t n_words = (a.n + 31u) / 32u;
    const uint64_t i = i0 + (uint64_t)lane;
    bool ok = false;
    uint2 h = make_uint2(0u, 0u);
    if (i < a.n && SPLIT) {
        ok = a.off[i + 1] - a.off[i] >= a.W;
        h = a.pat_hash[i];
        for (int j = 0; j < kSplitMaxK; j++) {
            if (j >= (int)a.k || !ok) break;
            const uint32_t ix = split_place(j, split_rev_fn(j, h.x, h.y), a.split);
            ok = (a.R[ix >> 5] >> (ix & 31u)) & 1u;
        }
    } else if (i < a.n) {
        ok = a.off[i + 1] - a.off[i] >= a.W;
        h = a.pat_hash[i];
        // R is sparse: its first bit for every pattern, the others (loaded
        // together) only where that one is set
        if (ok) {
            const uint32_t i0 = rev_idx(0, h.x, h.y) & a.mask;
            ok = (a.R[i0 >> 5] >> (i0 & 31u)) & 1u;
        }
        if (ok && a.k > 1u) {
            uint32_t words[5], idx[5];
#pragma unroll
            for (int j = 1; j < 5; j++) {
                idx[j] = rev_idx(j, h.x, h.y) & a.mask;
                words[j] = j < (int)a.k ? a.R[idx[j] >> 5] : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int j = 1; j < 5; j++) ok = ok && ((words[j] >> (idx[j] & 31u)) & 1u);
        }
    }
    const uint64_t bal = __ballot(ok);
    const uint64_t w = i0 >> 5;
    if (lane == 0 && w < n_words) a.bitmap[w] = (uint32_t)bal;
    if (lane == 32 && w + 1 < n_words) a.bitmap[w + 1] = (uint32_t)(bal >> 32);
    if (!ok) return (uint32_t)__builtin_popcountll(bal);
    // index by key: the first id to claim a slot represents the key, the
    // rest chain from the slot
    uint32_t slot = key_slot(h.x, h.y) & a.table_mask;
    for (;;) {
        uint32_t cur = atomicCAS(&a.table[slot], 0u, (uint32_t)i + 1u);
        if (cur == 0u) break;
        const uint2 ch = a.pat_hash[cur - 1u];
        if (ch.x == h.x && ch.y == h.y) {
            a.chain_next[i] = atomicExch(&a.slot_chain[slot], (uint32_t)i + 1u);
            break;
        }
        slot = (slot + 1u) & a.table_mask;
    }
    return (uint32_t)__builtin_popcountll(bal);
}

// The pattern pass as workgroup b of nb (the launch may hold other work).
template <bool SPLIT>
__device__ __forceinline__ void pattern_pass_blocks(const PassArgs &a, uint32_t b, uint32_t nb)
{
    __shared__ uint32_t s_n;
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
    const uint64_t wave0 = ((uint64_t)b * blockDim.x + (threadIdx.x & ~63u));
    const uint64_t stride = (uint64_t)nb * blockDim.x;
    uint32_t n = 0;
    for (uint64_t i0 = wave0; i0 < a.n; i0 += stride) n += pass_wave<SPLIT>(a, i0);
    // one count per block: atomics on one global address serialise, and a
    // per-wave add costs thousands of them per launch
    if (n && lane_id() == 0) atomicAdd(&s_n, n);
    __syncthreads();
    if (threadIdx.x == 0 && s_n) atomicAdd(a.counters + C_SURV, (unsigned long long)s_n);
}

template <bool SPLIT>
__global__ __launch_bounds__(256) void k_pattern_pass(PassArgs a)
{
    pattern_pass_blocks<SPLIT>(a, blockIdx.x, gridDim.x);
}

// The pattern passes of several filters in one launch: wave_start[f] is the
// first global 64-id wave of filter f (prefix over ceil(n_f/64)).  A filter
// whose `active` flag is clear (no text of the batch routed to it) has no
// survivors: its bitmap words are cleared and nothing is indexed.  Waves
// sweep the ids in order (grid stride), so the filters whose R is probed at
// any moment are few and stay in L2.  Survivors are not counted: a batch
// reports matches, and each filter's bitmap holds its survivors (a count per
// wave, added on one counter per filter, costs the largest filter thousands
// of serialised atomics).
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_pattern_pass_multi(const PassArgs *filters, const uint64_t *wave_start,
                                                            uint32_t nf, const uint32_t *active)
{
    const uint64_t total = wave_start[nf];
    const uint64_t gw0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t stride = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t gw = gw0; gw < total; gw += stride) {
        uint32_t lo = 0, hi = nf;  // last f with wave_start[f] <= gw
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (wave_start[mid] <= gw) lo = mid;
            else hi = mid;
        }
        const PassArgs &a = filters[lo];
        const uint64_t i0 = (gw - wave_start[lo]) << 6;
        if (active[lo]) {
            (void)pass_wave<SPLIT>(a, i0);
        } else {
            const uint64_t n_words = (a.n + 31u) / 32u, w = i0 >> 5;
            const int lane = lane_id();
            if (lane < 2 && w + (uint64_t)lane < n_words) a.bitmap[w + (uint64_t)lane] = 0u;
        }
    }
}

// Zeroes a list of spans of u32 words, one span per workgroup at a time:
// 16-byte stores from the first 16-byte boundary, single words around it.
__global__ __launch_bounds__(256) void k_zero_spans(const ZeroSpan *spans, uint32_t n)
{
    for (uint32_t c = blockIdx.x; c < n; c += gridDim.x) {
        const ZeroSpan z = spans[c];
        const uint32_t head = (uint32_t)(((16u - ((uintptr_t)z.p & 15u)) & 15u) / 4u);
        const uint32_t h = head < z.words ? head : z.words;
        const uint32_t n4 = (z.words - h) / 4u;
        if (threadIdx.x < h) z.p[threadIdx.x] = 0u;
        uint4 *v = (uint4 *)(z.p + h);
        for (uint32_t i = threadIdx.x; i < n4; i += blockDim.x) v[i] = make_uint4(0u, 0u, 0u, 0u);
        const uint32_t tail0 = h + n4 * 4u;
        if (threadIdx.x < z.words - tail0) z.p[tail0 + threadIdx.x] = 0u;
    }
}

// Exact verification: each candidate window looks its hash pair up among the
// survivors' prefix pairs, then compares the whole pattern byte for byte.
__global__ __launch_bounds__(256) void k_verify(VerifyArgs a)
{
    const unsigned long long total = a.counters[C_CVALID];
    const uint64_t count = total < a.cand_cap ? total : a.cand_cap;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < count; c += stride) {
        const Cand e = a.cand[c];
        verify_candidate(a, a.text, a.n, e.offset, e.h1, e.h2, [&](uint32_t id) {
            const unsigned long long s = atomicAdd(a.counters + C_MATCH, 1ull);
            if (s < a.out_cap) {
                MatchRec m;
                m.id = id;
                m.reserved = 0u;
                m.offset = e.offset;
                a.out[s] = m;
            }
        });
    }
}

// Filter::operator[] (ffbf.cpp:285-308) for many lines: one wave per line,
// lanes over the line's window starts 64 at a time, each hashing its window
// from the W bytes (the line is hashed from a reset and never restarted
// inside it, so every window's hash is its W bytes' alone) and testing its k
// forward bits in F.  The wave stops at the first window that hits.
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_query_lines(QueryArgs a)
{
    __shared__ uint32_t s_sb[512];
    constexpr uint32_t words = SPLIT ? 512u : 256u;
    for (uint32_t i = threadIdx.x; i < words; i += 256u) s_sb[i] = a.sbox[i];
    __syncthreads();
    const uint64_t n_waves = (uint64_t)gridDim.x * 4u;
    const uint32_t lane = threadIdx.x & 63u;
    for (uint64_t l = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6); l < a.n_lines; l += n_waves) {
        const uint64_t p0 = a.off[l], len = a.off[l + 1] - p0;
        bool any = false;
        if (len >= a.W) {
            const uint64_t nwin = len - a.W + 1u;
            for (uint64_t base = 0; base < nwin && !any; base += 64u) {
                const uint64_t o = base + lane;
                bool hit = false;
                if (o < nwin) {
                    const uint8_t *w = a.bytes + p0 + o;
                    uint32_t h1 = 0, h2 = 0;
                    for (uint32_t j = 0; j < a.W; j++) {
                        const uint32_t c = w[j];
                        if constexpr (SPLIT) {
                            uint32_t rl, rh;
                            rol64(h1, h2, 1u, rl, rh);
                            h1 = rl ^ s_sb[2u * c];
                            h2 = rh ^ s_sb[2u * c + 1u];
                        } else {
                            h1 = rotl(h1, 1) ^ s_sb[c];
                            h2 = rotl(h2, 3) ^ s_sb[c];
                        }
                    }
                    hit = true;
                    for (uint32_t j = 0; j < a.k && hit; j++) {
                        const uint32_t x = SPLIT ? split_place((int)j, split_fwd_fn((int)j, h1, h2), a.split)
                                                 : fwd_idx((int)j, h1, h2) & a.mask;
                        hit = ((a.F[x >> 5] >> (x & 31u)) & 1u) != 0u;
                    }
                }
                any = __ballot(hit) != 0;
            }
        }
        if (lane == 0) a.hits[l] = any ? 1u : 0u;
    }
}

// Host text into device memory (see UploadArgs).  Lane L of a segment's
// wave owns destination chunk c = first chunk + 63 * segment + L and loads
// the 16-aligned source chunk that holds the source of c's first byte; the
// chunk after it is lane L + 1's, so lanes 0..62 each assemble one
// destination chunk from two loaded ones (source and destination differ by
// the same shift r within a piece) and lane 63 only loads.  Source reads
// are buffer loads over the piece's 16-aligned extent: a chunk with no byte
// of the piece is out of the buffer's range and reads as zeros without an
// access, so no read leaves the pages the piece lies in.  Each wave first
// takes a system-scope acquire (an invalidate of the host lines in its L1
// and its XCD's L2), so a pinned buffer the host rewrote after an earlier
// upload is read anew; the text is not written while its request is in
// flight, so plain loads after that read the host's bytes (system-scope
// loads instead cost ~11% of the upload rate).  Chunks cut by a piece's
// ends are written byte by byte, which leaves the neighbouring piece's
// bytes of a shared chunk alone.
__global__ __launch_bounds__(256) void k_upload(const UploadArgs a)
{
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // system scope: buffer_inv sc0 sc1
    const uint32_t lane = threadIdx.x & 63u;
    // wave-uniform: the segment search and every per-piece value stay scalar
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t g = wave; g < a.total_segs; g += n_waves) {
        uint32_t lo = 0, hi = a.n;  // the piece: seg_start[lo] <= g < seg_start[lo + 1]
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.seg_start[mid] <= g) lo = mid;
            else hi = mid;
        }
        const uint64_t sp = a.src[lo], dp = a.dst_off[lo], n = a.len[lo];
        const uint64_t base = sp & ~15ull;
        const uint32_t extent = (uint32_t)(((sp + n + 15u) & ~15ull) - base);
        const uint64_t delta = sp - dp;  // source address of destination offset x: x + delta
        const uint32_t r = (uint32_t)(delta & 15u);
        const uint64_t c = (dp >> 4) + 63ull * (g - a.seg_start[lo]) + lane;
        // 16-aligned source chunk of c's first byte, relative to base (a
        // chunk before base wraps to a large offset: out of range as well)
        const uint32_t off = (uint32_t)(16ull * c + (delta - r) - base);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)extent, 0x00020000);
        const u32x4 cur = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
        u32x4 nxt;
        nxt.x = __shfl_down(cur.x, 1);
        nxt.y = __shfl_down(cur.y, 1);
        nxt.z = __shfl_down(cur.z, 1);
        nxt.w = __shfl_down(cur.w, 1);
        const uint64_t d0 = 16ull * c;
        const uint64_t b_lo = d0 > dp ? d0 : dp, b_hi = d0 + 16u < dp + n ? d0 + 16u : dp + n;
        if (lane == 63u || b_lo >= b_hi) continue;
        // bytes r .. r + 15 of cur:nxt (r is the same for the whole wave)
        const uint32_t w0 = cur.x, w1 = cur.y, w2 = cur.z, w3 = cur.w, w4 = nxt.x, w5 = nxt.y, w6 = nxt.z, w7 = nxt.w;
        const uint32_t sh = r & 3u;
        u32x4 o;
        switch (r >> 2) {
        case 0:
            o = u32x4{__builtin_amdgcn_alignbyte(w1, w0, sh), __builtin_amdgcn_alignbyte(w2, w1, sh),
                      __builtin_amdgcn_alignbyte(w3, w2, sh), __builtin_amdgcn_alignbyte(w4, w3, sh)};
            break;
        case 1:
            o = u32x4{__builtin_amdgcn_alignbyte(w2, w1, sh), __builtin_amdgcn_alignbyte(w3, w2, sh),
                      __builtin_amdgcn_alignbyte(w4, w3, sh), __builtin_amdgcn_alignbyte(w5, w4, sh)};
            break;
        case 2:
            o = u32x4{__builtin_amdgcn_alignbyte(w3, w2, sh), __builtin_amdgcn_alignbyte(w4, w3, sh),
                      __builtin_amdgcn_alignbyte(w5, w4, sh), __builtin_amdgcn_alignbyte(w6, w5, sh)};
            break;
        default:
            o = u32x4{__builtin_amdgcn_alignbyte(w4, w3, sh), __builtin_amdgcn_alignbyte(w5, w4, sh),
                      __builtin_amdgcn_alignbyte(w6, w5, sh), __builtin_amdgcn_alignbyte(w7, w6, sh)};
            break;
        }
        if (b_lo == d0 && b_hi == d0 + 16u) {
            *(u32x4 *)(a.dst + d0) = o;
        } else {
#pragma unroll
            for (uint32_t j = 0; j < 16u; j++)
                if (d0 + j >= b_lo && d0 + j < b_hi) a.dst[d0 + j] = (uint8_t)(o[j >> 2] >> (8u * (j & 3u)));
        }
    }
}

}  // namespace

hipError_t launch_fill32(uint32_t *p, uint64_t n_words, uint32_t value, hipStream_t s)
{
    if (n_words == 0) return hipSuccess;
    uint64_t blocks = (n_words / 4u + 255u) / 256u;
    if (blocks > 4096u) blocks = 4096u;
    if (blocks < 1u) blocks = 1u;
    hipLaunchKernelGGL(k_fill32, dim3((uint32_t)blocks), dim3(256), 0, s, p, n_words, value);
    return hipGetLastError();
}

hipError_t launch_copy(void *dst, const void *src, uint64_t bytes, hipStream_t s)
{
    const uint64_t n16 = bytes / 16u;
    if (n16 == 0) return hipSuccess;
    uint64_t blocks = (n16 + 255u) / 256u;
    if (blocks > 8192u) blocks = 8192u;
    hipLaunchKernelGGL(k_copy, dim3((uint32_t)blocks), dim3(256), 0, s, (uint4 *)dst, (const uint4 *)src, n16);
    return hipGetLastError();
}

hipError_t launch_upload(const UploadArgs &a, hipStream_t s)
{
    if (a.n == 0 || a.n > kUploadMax || a.seg_start[a.n] != a.total_segs) return hipErrorInvalidValue;
    for (uint32_t p = 0; p < a.n; p++)  // one 32-bit buffer range per piece
        if (!a.len[p] || a.len[p] > (1ull << 31)) return hipErrorInvalidValue;
    // 64 workgroups read the link at ~55 GB/s alone (the DMA engine: 57);
    // fewer for small batches
    const uint32_t blocks = std::min<uint32_t>(64u, (a.total_segs + 3u) / 4u);
    hipLaunchKernelGGL(k_upload, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_prefix_hash(const uint8_t *bytes, const uint64_t *off, uint64_t n, const uint32_t *sbox, uint32_t W,
                              uint2 *pat_hash, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255u) / 256u;
    if (blocks > 8192u) blocks = 8192u;
    hipLaunchKernelGGL(k_build, dim3((uint32_t)blocks), dim3(256), 0, s, bytes, off, n, sbox, W, pat_hash);
    return hipGetLastError();
}

hipError_t launch_prefix_hash_split(const uint8_t *bytes, const uint64_t *off, uint64_t n, const uint32_t *sbox64,
                                    uint32_t W, uint2 *pat_hash, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255u) / 256u;
    if (blocks > 8192u) blocks = 8192u;
    hipLaunchKernelGGL(k_build_split, dim3((uint32_t)blocks), dim3(256), 0, s, bytes, off, n, (const uint2 *)sbox64,
                       W, pat_hash);
    return hipGetLastError();
}

hipError_t launch_build_atomic(const uint64_t *off, const uint2 *pat_hash, uint64_t n, uint32_t W, uint32_t k,
                               uint32_t mask, bool split, const SplitMap &m, uint32_t *F, const BuildFolds &folds,
                               hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255u) / 256u;
    if (blocks > 8192u) blocks = 8192u;
    if (split)
        hipLaunchKernelGGL(k_build_atomic<true>, dim3((uint32_t)blocks), dim3(256), 0, s, off, pat_hash, n, W, k, mask,
                           m, F, folds);
    else
        hipLaunchKernelGGL(k_build_atomic<false>, dim3((uint32_t)blocks), dim3(256), 0, s, off, pat_hash, n, W, k,
                           mask, m, F, folds);
    return hipGetLastError();
}

uint64_t build_bin_blocks(uint64_t n, uint32_t k) { return (n + build_bin_patterns(k) - 1u) / build_bin_patterns(k); }

hipError_t launch_build_binned(const uint64_t *off, const uint2 *pat_hash, uint64_t n, uint32_t W, uint32_t k,
                               uint32_t mask, bool split, SplitMap m, uint32_t *F, uint64_t f_words, uint32_t *staging,
                               uint32_t *offs, hipStream_t s)
{
    const uint64_t ns = (f_words + (1u << (kBinSliceLog2 - 5)) - 1u) >> (kBinSliceLog2 - 5);
    const uint64_t nb = build_bin_blocks(n, k);
    if (ns > 4096u || nb > 0xFFFFFFFFu) return hipErrorInvalidValue;
    const uint32_t P = build_bin_patterns(k);
    if (nb) {
        const size_t lds = (2u * kBinEntries + (size_t)ns) * 4u;
        hipError_t e = hipFuncSetAttribute(split ? (const void *)k_build_bin<true> : (const void *)k_build_bin<false>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        if (split)
            hipLaunchKernelGGL(k_build_bin<true>, dim3((uint32_t)nb), dim3(256), lds, s, off, pat_hash, n, P, W, k, mask,
                               m, (uint32_t)ns, staging, offs);
        else
            hipLaunchKernelGGL(k_build_bin<false>, dim3((uint32_t)nb), dim3(256), lds, s, off, pat_hash, n, P, W, k,
                               mask, m, (uint32_t)ns, staging, offs);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const size_t lds = (size_t)(1u << (kBinSliceLog2 - 5)) * 4u;
    hipError_t e = hipFuncSetAttribute((const void *)k_build_slices, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_build_slices, dim3((uint32_t)ns), dim3(1024), lds, s, staging, offs, (uint32_t)nb,
                       (uint32_t)ns, F, f_words);
    return hipGetLastError();
}

hipError_t launch_fold(const uint32_t *F, uint64_t f_words, uint32_t *fold, uint32_t fold_words, hipStream_t s)
{
    // one thread per 16-byte column of the fold (per word for folds of fewer
    // than 4 words)
    const uint64_t cols = (fold_words & 3u) == 0u ? fold_words / 4u : fold_words;
    uint64_t blocks = (cols + 255u) / 256u;
    if (blocks < 1u) blocks = 1u;
    hipLaunchKernelGGL(k_fold, dim3((uint32_t)blocks), dim3(256), 0, s, F, f_words, fold, fold_words);
    return hipGetLastError();
}

hipError_t launch_pattern_pass(const PassArgs &a, int n_cu, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    uint64_t blocks = (a.n + 255u) / 256u;
    const uint64_t cap = (uint64_t)n_cu * 8u;
    if (blocks > cap) blocks = cap;
    if (a.method == METHOD_SPLIT)
        hipLaunchKernelGGL(k_pattern_pass<true>, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_pattern_pass<false>, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pattern_pass_multi(const PassArgs *filters, const uint64_t *wave_start, uint32_t nf,
                                     const uint32_t *active, uint64_t total_waves, int n_cu, hipStream_t s, bool split)
{
    if (total_waves == 0) return hipSuccess;
    uint64_t blocks = (total_waves + 3u) / 4u;
    const uint64_t cap = (uint64_t)n_cu * 8u;
    if (blocks > cap) blocks = cap;
    if (split)
        hipLaunchKernelGGL(k_pattern_pass_multi<true>, dim3((uint32_t)blocks), dim3(256), 0, s, filters, wave_start, nf,
                           active);
    else
        hipLaunchKernelGGL(k_pattern_pass_multi<false>, dim3((uint32_t)blocks), dim3(256), 0, s, filters, wave_start,
                           nf, active);
    return hipGetLastError();
}

hipError_t launch_zero_spans(const ZeroSpan *spans, uint32_t n, int n_cu, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const uint32_t blocks = n < (uint32_t)n_cu * 4u ? n : (uint32_t)n_cu * 4u;
    hipLaunchKernelGGL(k_zero_spans, dim3(blocks), dim3(256), 0, s, spans, n);
    return hipGetLastError();
}

hipError_t launch_verify(const VerifyArgs &a, int n_cu, hipStream_t s)
{
    hipLaunchKernelGGL(k_verify, dim3((uint32_t)(n_cu * 4)), dim3(256), 0, s, a);
    return hipGetLastError();
}


namespace {

// Exclusive prefix of the region counts, by one workgroup of 16 waves:
// offsets[w] = list length before region w, offsets[n] = where the spill
// goes, offsets[n+1] = spill entries kept.  Each wave owns a contiguous run
// of the counts, a whole number of 256-count rounds (one 16-byte group per
// lane per round), loads it coalesced and sums it; after one barrier it
// writes its offsets from the totals of the waves before it, by a lane
// prefix per round.  Counts and the spill counter are zeroed for the next
// launch; C_CAND becomes the new length, or exceeds out_cap when entries
// were lost (the caller then grows and rescans).  Only thread 0 reads and
// writes the counters.  A region keeps at most 4096 entries, so a wave's
// sums fit 32 bits for n < 2^20.
constexpr uint32_t kOffWaves = 16;
constexpr uint32_t kOffHeld = 8;  // rounds a wave keeps in registers between its two passes

// counts[i..i+3] (zeros past n); i % 4 == 0 and counts is 256-byte aligned
__device__ __forceinline__ uint4 count_group(const uint32_t *counts, uint32_t i, uint32_t n)
{
    if (i + 3u < n) return *(const uint4 *)(counts + i);
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (i < n) v.x = counts[i];
    if (i + 1u < n) v.y = counts[i + 1u];
    if (i + 2u < n) v.z = counts[i + 2u];
    return v;
}

__global__ __launch_bounds__(1024) void k_cand_offsets(uint32_t *counts, uint64_t *offsets, uint32_t n, uint64_t spill_cap,
                                                      uint64_t out_cap, unsigned long long *counters)
{
    __shared__ uint32_t s_wave[kOffWaves];
    __shared__ uint64_t s_base;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = threadIdx.x >> 6;
    // the scan kernel before this one left the counters final: thread 0
    // loads all three up front, beside the counts
    uint64_t base = 0, spill = 0, cvalid = 0;
    if (threadIdx.x == 0) {
        base = counters[C_CAND];
        spill = counters[C_SPILL];
        cvalid = counters[C_CVALID];
        s_base = base;
    }
    const uint32_t rounds = (n + kOffWaves * 256u - 1u) / (kOffWaves * 256u);
    const uint32_t first = w * rounds * 256u + lane * 4u;
    uint4 held[kOffHeld];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t r = 0; r < kOffHeld; r++) {
        held[r] = r < rounds ? count_group(counts, first + r * 256u, n) : make_uint4(0u, 0u, 0u, 0u);
        sum += held[r].x + held[r].y + held[r].z + held[r].w;
    }
    for (uint32_t r = kOffHeld; r < rounds; r++) {
        const uint4 v = count_group(counts, first + r * 256u, n);
        sum += v.x + v.y + v.z + v.w;
    }
    for (uint32_t d = 32u; d; d >>= 1) sum += __shfl_xor(sum, (int)d);
    if (lane == 0) s_wave[w] = sum;
    __syncthreads();
    uint64_t run = s_base, total = 0;
    for (uint32_t j = 0; j < kOffWaves; j++) {
        const uint32_t t = s_wave[j];
        if (j < w) run += t;
        total += t;
    }
    // one round: the lanes' exclusive prefix over their groups, four offsets
    // per lane, the counts zeroed; run moves on by the round's total
    auto emit = [&](uint32_t i, const uint4 v) {
        const uint32_t g = v.x + v.y + v.z + v.w;
        uint32_t inc = g;
        for (uint32_t d = 1u; d < 64u; d <<= 1) {
            const uint32_t o = __shfl_up(inc, d);
            if (lane >= d) inc += o;
        }
        const uint64_t o0 = run + (inc - g), o1 = o0 + v.x, o2 = o1 + v.y, o3 = o2 + v.z;
        if (i + 3u < n) {
            *(ulonglong2 *)(offsets + i) = make_ulonglong2(o0, o1);
            *(ulonglong2 *)(offsets + i + 2u) = make_ulonglong2(o2, o3);
            *(uint4 *)(counts + i) = make_uint4(0u, 0u, 0u, 0u);
        } else {
            if (i < n) offsets[i] = o0, counts[i] = 0u;
            if (i + 1u < n) offsets[i + 1u] = o1, counts[i + 1u] = 0u;
            if (i + 2u < n) offsets[i + 2u] = o2, counts[i + 2u] = 0u;
        }
        run += __shfl(inc, 63);
    };
#pragma unroll
    for (uint32_t r = 0; r < kOffHeld; r++)
        if (r < rounds) emit(first + r * 256u, held[r]);
    for (uint32_t r = kOffHeld; r < rounds; r++) emit(first + r * 256u, count_group(counts, first + r * 256u, n));
    if (threadIdx.x == 0) {
        const uint64_t end_regions = base + total;
        const uint64_t kept = spill < spill_cap ? spill : spill_cap;
        offsets[n] = end_regions;
        offsets[n + 1] = kept;
        const uint64_t end = end_regions + spill;
        counters[C_CAND] = (spill > spill_cap && end <= out_cap) ? out_cap + 1u : end;
        counters[C_SPILL] = 0u;
        // the list is whole up to the last entry written, if it was whole
        // before this launch; a lost spill leaves a gap after it
        const uint64_t written = end_regions + kept < out_cap ? end_regions + kept : out_cap;
        if (cvalid == base) counters[C_CVALID] = written;
    }
}

// Appends each region (one wave per region) and then the spill to out, as
// workgroup b of nb.
__device__ __forceinline__ void cand_gather_blocks(const GatherArgs &g, uint32_t b, uint32_t nb)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = (b * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nw = (nb * blockDim.x) >> 6;
    for (uint32_t w = gw; w < g.n; w += nw) {
        const uint64_t o = g.offsets[w], cnt = g.offsets[w + 1] - o;
        for (uint64_t i = lane; i < cnt; i += 64u)
            if (o + i < g.out_cap) g.out[o + i] = g.raw[(uint64_t)w * g.region + i];
    }
    const uint64_t so = g.offsets[g.n], sc = g.offsets[g.n + 1];
    const uint64_t stride = (uint64_t)nb * blockDim.x;
    for (uint64_t i = (uint64_t)b * blockDim.x + threadIdx.x; i < sc; i += stride)
        if (so + i < g.out_cap) g.out[so + i] = g.raw[g.spill_base + i];
}

__global__ __launch_bounds__(256) void k_cand_gather(GatherArgs g) { cand_gather_blocks(g, blockIdx.x, gridDim.x); }

// The pattern pass and the candidate gather in one launch: both read only
// what the scan and the offsets left (R; the staged candidates), nothing of
// each other, so workgroups [0, pass_blocks) run the pass and the rest the
// gather.  One launch fewer on the fused run path.
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_pass_gather(PassArgs a, GatherArgs g, uint32_t pass_blocks)
{
    if (blockIdx.x < pass_blocks)
        pattern_pass_blocks<SPLIT>(a, blockIdx.x, pass_blocks);
    else
        cand_gather_blocks(g, blockIdx.x - pass_blocks, gridDim.x - pass_blocks);
}

}  // namespace

hipError_t launch_cand_offsets(const Cand *raw, uint64_t raw_cap, uint32_t *wave_counts, uint64_t *offsets,
                               uint32_t n_regions, uint32_t region, Cand *out, uint64_t out_cap,
                               unsigned long long *counters, hipStream_t s, GatherArgs *g, uint32_t n_used)
{
    const uint64_t spill_base = (uint64_t)n_regions * region;
    if (raw_cap < spill_base) return hipErrorInvalidValue;
    // offsets[n] and offsets[n + 1] describe the spill, so offsets needs n + 2
    // entries for any n <= n_regions
    const uint32_t n = n_used < n_regions ? n_used : n_regions;
    hipLaunchKernelGGL(k_cand_offsets, dim3(1), dim3(1024), 0, s, wave_counts, offsets, n, raw_cap - spill_base, out_cap,
                       counters);
    *g = GatherArgs{raw, offsets, n, region, spill_base, out, out_cap};
    return hipGetLastError();
}

hipError_t launch_cand_compact(const Cand *raw, uint64_t raw_cap, uint32_t *wave_counts, uint64_t *offsets,
                               uint32_t n_regions, uint32_t region, Cand *out, uint64_t out_cap,
                               unsigned long long *counters, int n_cu, hipStream_t s, uint32_t n_used)
{
    GatherArgs g;
    hipError_t e = launch_cand_offsets(raw, raw_cap, wave_counts, offsets, n_regions, region, out, out_cap, counters, s,
                                       &g, n_used);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cand_gather, dim3((uint32_t)(n_cu * 2)), dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t launch_pass_gather(const PassArgs &a, const GatherArgs &g, int n_cu, hipStream_t s)
{
    uint64_t pass_blocks = (a.n + 255u) / 256u;
    const uint64_t cap = (uint64_t)n_cu * 8u;
    if (pass_blocks > cap) pass_blocks = cap;
    const uint32_t grid = (uint32_t)pass_blocks + (uint32_t)(n_cu * 2);
    if (a.method == METHOD_SPLIT)
        hipLaunchKernelGGL(k_pass_gather<true>, dim3(grid), dim3(256), 0, s, a, g, (uint32_t)pass_blocks);
    else
        hipLaunchKernelGGL(k_pass_gather<false>, dim3(grid), dim3(256), 0, s, a, g, (uint32_t)pass_blocks);
    return hipGetLastError();
}

hipError_t launch_query_lines(const QueryArgs &a, int n_cu, hipStream_t s)
{
    if (a.n_lines == 0) return hipSuccess;
    uint64_t blocks = (a.n_lines + 3u) / 4u;
    const uint64_t cap = (uint64_t)n_cu * 8u;
    if (blocks > cap) blocks = cap;
    if (a.method == METHOD_SPLIT)
        hipLaunchKernelGGL(k_query_lines<true>, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_query_lines<false>, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace ffbf
