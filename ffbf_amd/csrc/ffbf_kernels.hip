// ffbf_kernels.hip -- gfx950 kernels of the feed-forward Bloom filter path.
//
// Stages and the reference code whose results they reproduce bit for bit
// (efficient/ffbf):
//   k_build         grepSetup        src/gpgpu/greptest.cu:96-123
//   k_scan          GrepKernel       src/gpgpu/greptest.cu:125-251
//   k_pattern_pass  filterPatterns   src/gpgpu/greptest.cu:253-290
//   k_verify        the exact step the reference leaves to `grep -Ff`
//                   (src/rabin-karp/README:49-52)
//
// Hash pair of the W-byte window at offset o (greptest.cu:106-114):
//   h1 = XOR_j rotl(S[t[o+j]], W-1-j),  h2 = XOR_j rotl(S[t[o+j]], 3(W-1-j))
// computed here in its rolling form (hash_rot_sbox_2.h:34-51): one full window
// per lane, then 15 one-byte slides.  Forward bits h1, h2, h1+h2, h1+5h2
// (+ (h1<<16)|(h2>>16) for k=5), reverse bits h1+7h2, h1+3h2, (h1<<1)+h2,
// (h1<<2)+h2 (+ 11h1+h2), all mod 2^32 then masked to the filter size.
#include "ffbf_device.h"

namespace ffbf {

namespace {

constexpr int kUnit = 16;   // consecutive window offsets per lane (one aligned 16-byte text chunk)
constexpr int kQCap = 128;  // per-wave LDS queue of first-level survivors

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t rotl(uint32_t x, uint32_t r) { return __builtin_rotateleft32(x, r); }

__device__ __forceinline__ uint32_t fwd_idx(int j, uint32_t h1, uint32_t h2)
{
    switch (j) {
    case 0: return h1;
    case 1: return h2;
    case 2: return h1 + h2;
    case 3: return h1 + 5u * h2;
    default: return (h1 << 16) | (h2 >> 16);
    }
}

__device__ __forceinline__ uint32_t rev_idx(int j, uint32_t h1, uint32_t h2)
{
    switch (j) {
    case 0: return h1 + 7u * h2;
    case 1: return h1 + 3u * h2;
    case 2: return (h1 << 1) + h2;
    case 3: return (h1 << 2) + h2;
    default: return h1 * 11u + h2;
    }
}

__device__ __forceinline__ bool test_bit(const uint32_t *v, uint32_t b) { return (v[b >> 5] >> (b & 31u)) & 1u; }

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l)
{
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// Orders this wave's LDS traffic across lanes (writes before the barrier are
// seen by reads after it, from any lane of the same wave).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t mix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

// Hash-table slot of a prefix-window hash pair (verification index).
__device__ __forceinline__ uint32_t key_slot(uint32_t h1, uint32_t h2) { return mix32(h1 ^ rotl(h2, 13) ^ 0x9E3779B9u); }

// Feed-forward: set the k reverse bits of every lane with `hit` in R.  Lanes
// of the wave that target the same word are merged into one atomicOr
// (the reference issues one atomicOr per bit, greptest.cu:86-89, 236-239).
__device__ __forceinline__ void feed_forward(bool hit, uint32_t h1, uint32_t h2, uint32_t k, uint32_t mask, uint32_t *R)
{
    const int lane = lane_id();
#pragma unroll
    for (int j = 0; j < 5; j++) {
        if (j >= (int)k) break;
        const uint32_t idx = rev_idx(j, h1, h2) & mask;
        const uint32_t w = idx >> 5, bit = 1u << (idx & 31u);
        uint64_t act = __ballot(hit);
        while (act) {
            const int leader = __builtin_ctzll(act);
            const uint32_t w0 = __builtin_amdgcn_readlane(w, leader);
            const uint64_t same = __ballot(hit && w == w0) & act;
            uint32_t bits = 0;
            for (uint64_t m = same; m; m &= m - 1) bits |= __builtin_amdgcn_readlane(bit, __builtin_ctzll(m));
            if (lane == leader) atomicOr(&R[w0], bits);
            act &= ~same;
        }
    }
}

// Appends (offset, h1, h2) of every lane with `hit`: one atomicAdd per wave,
// slots by lane rank.
__device__ __forceinline__ void append_cand(bool hit, uint64_t off, uint32_t h1, uint32_t h2, Cand *cand, uint64_t cap,
                                            unsigned long long *counter)
{
    const uint64_t hb = __ballot(hit);
    if (!hb) return;
    const int leader = __builtin_ctzll(hb);
    unsigned long long first = 0;
    if (lane_id() == leader) first = atomicAdd(counter, (unsigned long long)__builtin_popcountll(hb));
    first = readlane64(first, leader);
    if (hit) {
        const uint64_t slot = first + lanes_below(hb);
        if (slot < cap) {
            Cand c;
            c.offset = off;
            c.h1 = h1;
            c.h2 = h2;
            cand[slot] = c;
        }
    }
}

// One aligned 16-byte text chunk c (index relative to the 16-byte aligned base
// below the text); bytes outside [lead, lead+n) read as zero and are never
// part of an evaluated window.
__device__ __forceinline__ uint4 load_chunk(const uint8_t *base, uint64_t lead, uint64_t n, uint64_t c)
{
    const uint64_t lo = c * 16u;
    if (lo >= lead && lo + 16u <= lead + n) {
        const u32x4 v = __builtin_nontemporal_load((const u32x4 *)(base + lo));
        return make_uint4(v.x, v.y, v.z, v.w);
    }
    uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint64_t q = lo + (uint64_t)i;
        if (q >= lead && q < lead + n) w[i >> 2] |= (uint32_t)base[q] << (8 * (i & 3));
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// Hash pairs of the 16 windows starting at base-relative bytes 16u .. 16u+15,
// W a compile-time constant: 16 + W - 1 bytes in registers, one S lookup per
// byte, then one full window and 15 rolling slides.
template <int W>
__device__ __forceinline__ void hash_unit_fast(const uint8_t *base, uint64_t lead, uint64_t n, uint64_t u,
                                               const uint32_t *sb, uint32_t (&h1)[kUnit], uint32_t (&h2)[kUnit])
{
    constexpr int NB = kUnit + W - 1;
    constexpr int NCH = (NB + 15) / 16;
    uint32_t dw[NCH * 4];
#pragma unroll
    for (int c = 0; c < NCH; c++) {
        const uint4 v = load_chunk(base, lead, n, u + (uint64_t)c);
        dw[4 * c + 0] = v.x;
        dw[4 * c + 1] = v.y;
        dw[4 * c + 2] = v.z;
        dw[4 * c + 3] = v.w;
    }
    uint32_t s[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) s[j] = sb[(dw[j >> 2] >> (8 * (j & 3))) & 0xFFu];
    uint32_t x1 = 0, x2 = 0;
#pragma unroll
    for (int j = 0; j < W; j++) {
        x1 = rotl(x1, 1) ^ s[j];
        x2 = rotl(x2, 3) ^ s[j];
    }
    h1[0] = x1;
    h2[0] = x2;
#pragma unroll
    for (int b = 1; b < kUnit; b++) {
        x1 = rotl(x1, 1) ^ s[b + W - 1] ^ rotl(s[b - 1], (uint32_t)W);
        x2 = rotl(x2, 3) ^ s[b + W - 1] ^ rotl(s[b - 1], (uint32_t)(3 * W));
        h1[b] = x1;
        h2[b] = x2;
    }
}

// Same for a run-time W (any 1..64): bytes fetched one at a time.
__device__ __forceinline__ uint32_t byte_at(const uint8_t *base, uint64_t lead, uint64_t n, uint64_t q)
{
    return (q >= lead && q < lead + n) ? (uint32_t)base[q] : 0u;
}

__device__ __forceinline__ void hash_unit_generic(const uint8_t *base, uint64_t lead, uint64_t n, uint64_t u, uint32_t W,
                                                  const uint32_t *sb, uint32_t (&h1)[kUnit], uint32_t (&h2)[kUnit])
{
    const uint64_t q0 = u * 16u;
    uint32_t x1 = 0, x2 = 0;
    for (uint32_t j = 0; j < W; j++) {
        const uint32_t s = sb[byte_at(base, lead, n, q0 + j)];
        x1 = rotl(x1, 1) ^ s;
        x2 = rotl(x2, 3) ^ s;
    }
    h1[0] = x1;
    h2[0] = x2;
#pragma unroll
    for (int b = 1; b < kUnit; b++) {
        const uint32_t si = sb[byte_at(base, lead, n, q0 + (uint64_t)b + W - 1)];
        const uint32_t so = sb[byte_at(base, lead, n, q0 + (uint64_t)b - 1)];
        x1 = rotl(x1, 1) ^ si ^ rotl(so, W);
        x2 = rotl(x2, 3) ^ si ^ rotl(so, 3u * W);
        h1[b] = x1;
        h2[b] = x2;
    }
}

// Resolves up to 64 queued first-level survivors, one per lane: the remaining
// forward probes against F, then feed-forward and candidate append.
template <int MODE>
__device__ __forceinline__ void drain(const uint4 *q, uint32_t cnt, const ScanArgs &a)
{
    const int lane = lane_id();
    const bool act = (uint32_t)lane < cnt;
    uint4 e = make_uint4(0u, 0u, 0u, 0u);
    if (act) e = q[lane];
    const uint32_t h1 = e.z, h2 = e.w;
    bool hit = act;
    if (MODE != MODE_LDS) {
        // all remaining probes issued together: one round trip
        uint32_t words[5];
        uint32_t idx[5];
        constexpr int first = (MODE == MODE_GLOBAL) ? 1 : 0;
#pragma unroll
        for (int j = first; j < 5; j++) {
            idx[j] = fwd_idx(j, h1, h2) & a.mask;
            words[j] = (act && j < (int)a.k) ? a.F[idx[j] >> 5] : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int j = first; j < 5; j++) hit = hit && ((words[j] >> (idx[j] & 31u)) & 1u);
    }
    const uint64_t off = ((uint64_t)e.y << 32) | e.x;
    feed_forward(hit, h1, h2, a.k, a.mask, a.R);
    append_cand(hit, off, h1, h2, a.cand, a.cand_cap, a.counters + C_CAND);
}

// Text scan with feed-forward (GrepKernel, greptest.cu:125-251).  Persistent
// waves walk 1 KiB text slices (64 lanes x 16 offsets); every lane hashes its
// 16 windows, tests them first-level, and survivors go through a per-wave LDS
// queue that is resolved 64 at a time.
template <int WT, int MODE, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_scan(ScanArgs a)
{
    extern __shared__ uint32_t s_fold[];
    __shared__ uint32_t s_sbox[256];
    __shared__ uint4 s_q[BLOCK / 64][kQCap];

    for (uint32_t i = threadIdx.x; i < 256u; i += BLOCK) s_sbox[i] = a.sbox[i];
    if (MODE != MODE_GLOBAL) {
        const uint4 *src = (const uint4 *)a.fold;
        uint4 *dst = (uint4 *)s_fold;
        for (uint32_t i = threadIdx.x; i < a.fold_words / 4u; i += BLOCK) dst[i] = src[i];
    }
    __syncthreads();

    const uint32_t W = WT ? (uint32_t)WT : a.W;
    const int wv = (int)(threadIdx.x >> 6);
    const uint64_t lead = (uint64_t)((uintptr_t)a.text & 15u);
    const uint8_t *base = a.text - lead;
    const uint64_t npos = a.n >= W ? a.n - W + 1 : 0;
    const uint64_t units = (lead + npos + 15u) >> 4;
    const uint64_t iters = (units + 63u) >> 6;
    const uint64_t n_waves = (uint64_t)gridDim.x * (BLOCK / 64);
    uint4 *q = s_q[wv];
    uint32_t qn = 0;  // wave-uniform queue depth

    for (uint64_t it = (uint64_t)blockIdx.x * (BLOCK / 64) + wv; it < iters; it += n_waves) {
        const uint64_t u = (it << 6) + (uint64_t)lane_id();
        uint32_t h1[kUnit], h2[kUnit];
        if constexpr (WT > 0)
            hash_unit_fast<WT>(base, lead, a.n, u, s_sbox, h1, h2);
        else
            hash_unit_generic(base, lead, a.n, u, W, s_sbox, h1, h2);

        // window offset of slot b is 16u + b - lead; valid while in [0, npos)
        const int64_t p0 = (int64_t)(u * 16u) - (int64_t)lead;
        uint32_t pm = 0;
        if (MODE == MODE_GLOBAL) {
            uint32_t fw[kUnit];
#pragma unroll
            for (int b = 0; b < kUnit; b++) {
                const int64_t p = p0 + b;
                const bool v = p >= 0 && (uint64_t)p < npos;
                fw[b] = v ? a.F[(h1[b] & a.mask) >> 5] : 0u;
            }
#pragma unroll
            for (int b = 0; b < kUnit; b++) pm |= ((fw[b] >> (h1[b] & 31u)) & 1u) << b;
        } else {
#pragma unroll
            for (int b = 0; b < kUnit; b++) {
                const int64_t p = p0 + b;
                bool ok = p >= 0 && (uint64_t)p < npos;
#pragma unroll
                for (int j = 0; j < 5; j++) {
                    if (j < (int)a.k) ok = ok && test_bit(s_fold, fwd_idx(j, h1[b], h2[b]) & a.fold_mask);
                }
                pm |= (uint32_t)ok << b;
            }
        }

#pragma unroll
        for (int b = 0; b < kUnit; b++) {
            const bool pass = (pm >> b) & 1u;
            const uint64_t bal = __ballot(pass);
            if (bal) {
                if (pass) {
                    const uint64_t p = (uint64_t)(p0 + b);
                    q[qn + lanes_below(bal)] = make_uint4((uint32_t)p, (uint32_t)(p >> 32), h1[b], h2[b]);
                }
                qn += (uint32_t)__builtin_popcountll(bal);
                if (qn >= 64u) {
                    wave_sync();
                    drain<MODE>(q + (qn - 64u), 64u, a);
                    wave_sync();
                    qn -= 64u;
                }
            }
        }
    }
    if (qn) {
        wave_sync();
        drain<MODE>(q, qn, a);
    }
}

__global__ void k_fill32(uint32_t *p, uint64_t n, uint32_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n4 = n / 4u;
    uint4 *p4 = (uint4 *)p;
    const uint4 v4 = make_uint4(v, v, v, v);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) p4[i] = v4;
    for (uint64_t i = n4 * 4u + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Filter build (grepSetup, greptest.cu:96-123): one lane per pattern hashes
// its W-byte prefix and sets k forward bits; patterns shorter than W are
// skipped (ffbf.cpp:271-275).  The hash pair is kept for the pattern pass and
// verification; with F null only the hash pairs are computed.
__global__ __launch_bounds__(256) void k_build(const uint8_t *bytes, const uint64_t *off, uint64_t n,
                                               const uint32_t *sbox, uint32_t W, uint32_t k, uint32_t mask,
                                               uint32_t *F, uint2 *pat_hash, unsigned long long *counters)
{
    __shared__ uint32_t sb[256];
    for (uint32_t i = threadIdx.x; i < 256u; i += blockDim.x) sb[i] = sbox[i];
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t hashed = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t a0 = off[i], a1 = off[i + 1];
        uint32_t h1 = 0, h2 = 0;
        if (a1 - a0 >= W) {
            const uint8_t *p = bytes + a0;
            for (uint32_t j = 0; j < W; j++) {
                const uint32_t s = sb[p[j]];
                h1 = rotl(h1, 1) ^ s;
                h2 = rotl(h2, 3) ^ s;
            }
            if (F) {
                for (int j = 0; j < 5; j++) {
                    if (j >= (int)k) break;
                    const uint32_t idx = fwd_idx(j, h1, h2) & mask;
                    atomicOr(&F[idx >> 5], 1u << (idx & 31u));
                }
            }
            hashed++;
        }
        pat_hash[i] = make_uint2(h1, h2);
    }
    if (hashed) atomicAdd(counters + C_HASHED, (unsigned long long)hashed);
}

// OR-fold of F onto fold_words words: fold bit x is set iff some F bit
// congruent to x mod fold_words*32 is set, so a window whose k forward bits are
// set in F has its k folded bits set too (exact superset test).
__global__ void k_fold(const uint32_t *F, uint64_t f_words, uint32_t *fold, uint32_t fold_words)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < fold_words; j += stride) {
        uint32_t v = 0;
        for (uint64_t t = j; t < f_words; t += fold_words) v |= F[t];
        fold[j] = v;
    }
}

// Pattern pass (filterPatterns, greptest.cu:253-290): each pattern's prefix
// hash pair against R.  Each wave owns 64 consecutive ids, so its ballot is
// two whole bitmap words; survivors are compacted with a wave prefix count
// and indexed by prefix-hash key for verification.
__global__ __launch_bounds__(256) void k_pattern_pass(PassArgs a)
{
    const int lane = lane_id();
    const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u));
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n_words = (a.n + 31u) / 32u;
    for (uint64_t i0 = wave0; i0 < a.n; i0 += stride) {
        const uint64_t i = i0 + (uint64_t)lane;
        bool ok = false;
        uint2 h = make_uint2(0u, 0u);
        if (i < a.n) {
            ok = a.off[i + 1] - a.off[i] >= a.W;
            h = a.pat_hash[i];
            uint32_t words[5], idx[5];
#pragma unroll
            for (int j = 0; j < 5; j++) {
                idx[j] = rev_idx(j, h.x, h.y) & a.mask;
                words[j] = (ok && j < (int)a.k) ? a.R[idx[j] >> 5] : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int j = 0; j < 5; j++) ok = ok && ((words[j] >> (idx[j] & 31u)) & 1u);
        }
        const uint64_t bal = __ballot(ok);
        const uint64_t w = i0 >> 5;
        if (lane == 0 && w < n_words) a.bitmap[w] = (uint32_t)bal;
        if (lane == 32 && w + 1 < n_words) a.bitmap[w + 1] = (uint32_t)(bal >> 32);
        if (!bal) continue;
        const int leader = __builtin_ctzll(bal);
        unsigned long long first = 0;
        if (lane == leader) first = atomicAdd(a.counters + C_SURV, (unsigned long long)__builtin_popcountll(bal));
        first = readlane64(first, leader);
        if (!ok) continue;
        a.surv_ids[first + lanes_below(bal)] = (uint32_t)i;
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
    }
}

__device__ __forceinline__ bool bytes_equal(const uint8_t *x, const uint8_t *y, uint64_t len)
{
    for (uint64_t i = 0; i < len; i++)
        if (x[i] != y[i]) return false;
    return true;
}

// Exact verification: each candidate window looks its hash pair up among the
// survivors' prefix pairs, then compares the whole pattern byte for byte.
__global__ __launch_bounds__(256) void k_verify(VerifyArgs a)
{
    const unsigned long long total = a.counters[C_CAND];
    const uint64_t count = total < a.cand_cap ? total : a.cand_cap;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < count; c += stride) {
        const Cand e = a.cand[c];
        uint32_t slot = key_slot(e.h1, e.h2) & a.table_mask;
        for (;;) {
            const uint32_t t = a.table[slot];
            if (t == 0u) break;
            const uint2 ph = a.pat_hash[t - 1u];
            if (ph.x == e.h1 && ph.y == e.h2) {
                uint32_t id1 = t;
                uint32_t next = a.slot_chain[slot];
                while (id1) {
                    const uint32_t id = id1 - 1u;
                    const uint64_t p0 = a.off[id], len = a.off[id + 1] - p0;
                    if (e.offset + len <= a.n && bytes_equal(a.text + e.offset, a.bytes + p0, len)) {
                        const unsigned long long s = atomicAdd(a.counters + C_MATCH, 1ull);
                        if (s < a.out_cap) {
                            MatchRec m;
                            m.id = id;
                            m.reserved = 0u;
                            m.offset = e.offset;
                            a.out[s] = m;
                        }
                    }
                    id1 = next;
                    if (next) next = a.chain_next[next - 1u];
                }
                break;
            }
            slot = (slot + 1u) & a.table_mask;
        }
    }
}

template <int WT, int MODE, int BLOCK>
hipError_t launch_scan_t(const ScanArgs &a, int n_cu, hipStream_t s)
{
    const size_t dyn = (MODE == MODE_GLOBAL) ? 0 : (size_t)a.fold_words * 4u;
    if (dyn > 0) {
        hipError_t e = hipFuncSetAttribute((const void *)k_scan<WT, MODE, BLOCK>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
    }
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_scan<WT, MODE, BLOCK>, BLOCK, dyn);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    const uint64_t W = WT ? (uint64_t)WT : a.W;
    const uint64_t npos = a.n >= W ? a.n - W + 1 : 0;
    const uint64_t iters = (npos + 16u + 1023u) / 1024u;
    uint64_t grid = (uint64_t)n_cu * (uint64_t)per_cu;
    const uint64_t need = (iters + (BLOCK / 64) - 1) / (BLOCK / 64);
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL((k_scan<WT, MODE, BLOCK>), dim3((uint32_t)grid), dim3(BLOCK), dyn, s, a);
    return hipGetLastError();
}

template <int WT>
hipError_t launch_scan_w(const ScanArgs &a, uint32_t mode, int n_cu, hipStream_t s)
{
    switch (mode) {
    case MODE_GLOBAL: return launch_scan_t<WT, MODE_GLOBAL, 256>(a, n_cu, s);
    case MODE_FOLD: return launch_scan_t<WT, MODE_FOLD, 512>(a, n_cu, s);
    case MODE_LDS: return launch_scan_t<WT, MODE_LDS, 512>(a, n_cu, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool scan_has_fast_path(uint32_t W) { return W == 19 || W == 24 || W == 32; }

uint32_t max_fold_words(int device)
{
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || lds <= 0)
        lds = 64 * 1024;
    // static LDS of the 512-thread scan: sbox 1 KiB + 8 queues of 2 KiB
    const int static_lds = 1024 + 8 * kQCap * 16;
    uint32_t words = 1u << 15;  // 128 KiB
    while (words > 256u && (int)(words * 4u) + static_lds > lds) words >>= 1;
    return words;
}

hipError_t launch_fill32(uint32_t *p, uint64_t n_words, uint32_t value, hipStream_t s)
{
    if (n_words == 0) return hipSuccess;
    uint64_t blocks = (n_words / 4u + 255u) / 256u;
    if (blocks > 4096u) blocks = 4096u;
    if (blocks < 1u) blocks = 1u;
    hipLaunchKernelGGL(k_fill32, dim3((uint32_t)blocks), dim3(256), 0, s, p, n_words, value);
    return hipGetLastError();
}

hipError_t launch_build(const uint8_t *bytes, const uint64_t *off, uint64_t n, const uint32_t *sbox, uint32_t W,
                        uint32_t k, uint32_t mask, uint32_t *F, uint2 *pat_hash, unsigned long long *counters,
                        hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255u) / 256u;
    if (blocks > 8192u) blocks = 8192u;
    hipLaunchKernelGGL(k_build, dim3((uint32_t)blocks), dim3(256), 0, s, bytes, off, n, sbox, W, k, mask, F, pat_hash,
                       counters);
    return hipGetLastError();
}

hipError_t launch_fold(const uint32_t *F, uint64_t f_words, uint32_t *fold, uint32_t fold_words, hipStream_t s)
{
    uint64_t blocks = ((uint64_t)fold_words + 255u) / 256u;
    if (blocks < 1u) blocks = 1u;
    hipLaunchKernelGGL(k_fold, dim3((uint32_t)blocks), dim3(256), 0, s, F, f_words, fold, fold_words);
    return hipGetLastError();
}

hipError_t launch_scan(const ScanArgs &a, uint32_t mode, int n_cu, hipStream_t s)
{
    switch (a.W) {
    case 19: return launch_scan_w<19>(a, mode, n_cu, s);
    case 24: return launch_scan_w<24>(a, mode, n_cu, s);
    case 32: return launch_scan_w<32>(a, mode, n_cu, s);
    default: return launch_scan_w<0>(a, mode, n_cu, s);
    }
}

hipError_t launch_pattern_pass(const PassArgs &a, int n_cu, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    uint64_t blocks = (a.n + 255u) / 256u;
    const uint64_t cap = (uint64_t)n_cu * 8u;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_pattern_pass, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_verify(const VerifyArgs &a, int n_cu, hipStream_t s)
{
    hipLaunchKernelGGL(k_verify, dim3((uint32_t)(n_cu * 4)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace ffbf
