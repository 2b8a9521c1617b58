// ffbf_common.h -- device helpers shared by the FFBF kernels (internal).
//
// Hash pair of the W-byte window at offset o (greptest.cu:106-114):
//   h1 = XOR_j rotl(S[t[o+j]], W-1-j),  h2 = XOR_j rotl(S[t[o+j]], 3(W-1-j))
// Forward bit indices (greptest.cu:115-120, 199-202; the fifth is the
// commented-out (h1<<16)|(h2>>16) of :121-122):   h1, h2, h1+h2, h1+5h2
// Reverse bit indices (greptest.cu:231-234, 273-276; the fifth is the
// commented-out 11h1+h2 of :235):                 h1+7h2, h1+3h2, (h1<<1)+h2, (h1<<2)+h2
// All mod 2^32, then masked to the filter size.  Bit b of a filter is bit
// b&31 of word b>>5 (greptest.cu:80-89).
#pragma once

#include "ffbf_device.h"

namespace ffbf {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t rotl(uint32_t x, uint32_t r) { return __builtin_rotateleft32(x, r); }

// a ^ b ^ c in one VALU op (gfx950 three-input bitwise op, truth table 0x96);
// the compiler does not form it from XOR chains by itself.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

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

// Split method (src/rabin-karp/ffbf.cpp): forward functions of the 64-bit
// hash's halves (getHashValue, :135-159) and reverse ones (getHashValue_2,
// :161-185), before placement.
__device__ __forceinline__ uint32_t split_fwd_fn(int j, uint32_t h1, uint32_t h2)
{
    switch (j) {
    case 0: return h1;
    case 1: return h2;
    case 2: return rotl(h1, 16);
    case 3: return rotl(h2, 16);
    case 4: return h1 + h2;
    case 5: return h1 + (h2 << 1);
    case 6: return h1 + (h2 << 2);
    case 7: return h1 + 3u * h2;
    case 8: return h1 + 5u * h2;
    case 9: return h1 + 6u * h2;
    default: return h1 + 7u * h2;
    }
}

__device__ __forceinline__ uint32_t split_rev_fn(int j, uint32_t h1, uint32_t h2)
{
    switch (j) {
    case 0: return h1 + (h2 << 1);
    case 1: return h1 + (h2 << 2);
    case 2: return h1 + 3u * h2;
    case 3: return h1 + 5u * h2;
    case 4: return h1 + 6u * h2;
    case 5: return h1 + 7u * h2;
    case 6: return h1;
    case 7: return h2;
    case 8: return rotl(h1, 16);
    case 9: return rotl(h2, 16);
    default: return h1 + h2;
    }
}

__device__ __forceinline__ uint32_t split_place(int j, uint32_t x, const SplitMap &m)
{
    return (uint32_t)j < m.s ? (x & m.cmask) : (x & m.emask) + m.cache;
}

// 64-bit rotation left of (lo, hi) by r mod 64, as two funnel shifts
// (v_alignbit_b32 takes the high word first); r is a constant after
// unrolling wherever speed matters.
__device__ __forceinline__ void rol64(uint32_t lo, uint32_t hi, uint32_t r, uint32_t &olo, uint32_t &ohi)
{
    r &= 63u;
    if (r & 32u) {
        const uint32_t t = lo;
        lo = hi;
        hi = t;
    }
    r &= 31u;
    if (r == 0u) {
        olo = lo;
        ohi = hi;
        return;
    }
    olo = __builtin_amdgcn_alignbit(lo, hi, 32u - r);
    ohi = __builtin_amdgcn_alignbit(hi, lo, 32u - r);
}

__device__ __forceinline__ uint32_t bit_of(const uint32_t *v, uint32_t b) { return (v[b >> 5] >> (b & 31u)) & 1u; }

// Bit x (taken modulo 32*2^log2_words) of a filter held in LDS: word index by
// one bit-field extract, bit by another (v_bfe_u32 reads only the low 5 bits
// of its offset operand, so x needs no masking).
__device__ __forceinline__ uint32_t lds_bit(const uint32_t *s, uint32_t x, uint32_t log2_words)
{
    const uint32_t w = s[__builtin_amdgcn_ubfe(x, 5u, log2_words)];
    return __builtin_amdgcn_ubfe(w, x, 1u);
}

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

// OR of v over the whole wave (uniform result): inclusive OR scan inside each
// 16-lane row by DPP row shifts, then the four row totals.
__device__ __forceinline__ uint32_t wave_or(uint32_t v)
{
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111 /* row_shr:1 */, 0xF, 0xF, true);
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112 /* row_shr:2 */, 0xF, 0xF, true);
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114 /* row_shr:4 */, 0xF, 0xF, true);
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118 /* row_shr:8 */, 0xF, 0xF, true);
    return __builtin_amdgcn_readlane(v, 15) | __builtin_amdgcn_readlane(v, 31) | __builtin_amdgcn_readlane(v, 47) |
           __builtin_amdgcn_readlane(v, 63);
}

// Sum of v over lanes 0..lane_id() (inclusive): DPP row shifts build the sums
// inside each 16-lane row, then row broadcasts carry the row totals upward.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111 /* row_shr:1 */, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112 /* row_shr:2 */, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114 /* row_shr:4 */, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118 /* row_shr:8 */, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142 /* row_bcast:15 */, 0xA, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143 /* row_bcast:31 */, 0xC, 0xF, false);
    return v;
}

// Orders this wave's LDS traffic across its lanes: writes before the call are
// seen by reads after it, from any lane of the same wave.
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

// Whether some lane 1..6 below this one in its 16-lane row also hits and has
// the same value: drained candidates sit in queue order, so the repeats of a
// short tandem repeat land in nearby lanes with identical hashes.
__device__ __forceinline__ bool dup_below(bool hit, uint32_t v)
{
    const int h = hit ? 1 : 0;
    bool dup = false;
#define FFBF_DUP_AT(n)                                                                                        \
    {                                                                                                         \
        const uint32_t w = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x110 + (n), 0xF, 0xF, false);    \
        const int hh = __builtin_amdgcn_update_dpp(0, h, 0x110 + (n), 0xF, 0xF, false);                       \
        dup = dup || (hh != 0 && w == v);                                                                     \
    }
    FFBF_DUP_AT(1) FFBF_DUP_AT(2) FFBF_DUP_AT(3) FFBF_DUP_AT(4) FFBF_DUP_AT(5) FFBF_DUP_AT(6)
#undef FFBF_DUP_AT
    return hit && dup;
}

// Feed-forward: set the k reverse bits of every lane with `hit` in R (the
// reference issues one atomicOr per bit, greptest.cu:86-89, 236-239);
// bit_of(j) is lane's bit index of function j.  A few hits: lanes that
// target the same word are merged into one atomicOr, one word at a time.
// Many hits (dense candidates): one atomicOr per lane and bit, except where
// a lane close below sets the same bit.  Must be called by all lanes.
template <int KMAX, class BitOf>
__device__ __forceinline__ void feed_forward_bits(bool hit, uint32_t k, uint32_t *R, BitOf bit_of)
{
    const uint64_t hits = __ballot(hit);
    if (!hits) return;
    const bool dense = __builtin_popcountll(hits) > 4;
    const int lane = lane_id();
    // a rolled loop (bit_of(j) then selects by j at run time): it runs for
    // waves holding hits only, and its per-word merge loop keeps the
    // compiler from flattening it anyway
#pragma nounroll
    for (int j = 0; j < KMAX; j++) {
        if (j >= (int)k) break;
        const uint32_t idx = bit_of(j);
        if (dense) {
            const bool dup = dup_below(hit, idx);  // all lanes: DPP reads neighbours
            if (hit && !dup) atomicOr(&R[idx >> 5], 1u << (idx & 31u));
            continue;
        }
        const uint32_t w = idx >> 5, bit = 1u << (idx & 31u);
        uint64_t act = hits;
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

__device__ __forceinline__ void feed_forward(bool hit, uint32_t h1, uint32_t h2, uint32_t k, uint32_t mask, uint32_t *R)
{
    feed_forward_bits<5>(hit, k, R, [&](int j) { return rev_idx(j, h1, h2) & mask; });
}

// Feed-forward into a workgroup's LDS copy of R: LDS atomics, no merging (a
// small R in global memory concentrates every candidate's atomics on a few
// cache lines; in LDS they stay on the CU).
__device__ __forceinline__ void feed_forward_lds(bool hit, uint32_t h1, uint32_t h2, uint32_t k, uint32_t mask,
                                                 uint32_t *s_R)
{
    if (!hit) return;
#pragma unroll
    for (int j = 0; j < 5; j++) {
        if (j >= (int)k) break;
        const uint32_t idx = rev_idx(j, h1, h2) & mask;
        atomicOr(&s_R[idx >> 5], 1u << (idx & 31u));
    }
}

// ORs a workgroup's LDS copy of R (words words) into R; every thread of the
// block calls it after a barrier.
__device__ __forceinline__ void flush_lds_r(const uint32_t *s_R, uint32_t words, uint32_t *R)
{
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) {
        const uint32_t v = s_R[i];
        if (v) atomicOr(&R[i], v);
    }
}

__device__ __forceinline__ void zero_lds(uint32_t *s, uint32_t words)
{
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) s[i] = 0u;
}

// The same for the split method: k (<= 11) reverse bits placed in the cache
// or external part (setBitsInFilter_2, ffbf.cpp:205-212).
__device__ __forceinline__ void feed_forward_split(bool hit, uint32_t h1, uint32_t h2, uint32_t k, const SplitMap &m,
                                                   uint32_t *R)
{
    feed_forward_bits<kSplitMaxK>(hit, k, R, [&](int j) { return split_place(j, split_rev_fn(j, h1, h2), m); });
}

// Appends (offset, h1, h2) of every lane with `hit`: one atomicAdd per wave,
// slots by lane rank.  Must be called by all lanes of the wave.
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

// x[0..len) == y[0..len), eight bytes of each per step: the step's loads are
// independent, so one memory latency per eight bytes rather than per byte.
// The last step of a run of at least 8 bytes is the 8 ending at len
// (overlapping the one before it), so nothing past len is read.
__device__ __forceinline__ bool bytes_equal8(const uint8_t *x, const uint8_t *y)
{
    uint32_t d = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8u; j++) d |= (uint32_t)(x[j] ^ y[j]);
    return d == 0u;
}

__device__ __forceinline__ bool bytes_equal(const uint8_t *x, const uint8_t *y, uint64_t len)
{
    if (len < 8u) {
        for (uint64_t i = 0; i < len; i++)
            if (x[i] != y[i]) return false;
        return true;
    }
    for (uint64_t i = 0; i + 8u < len; i += 8u)
        if (!bytes_equal8(x + i, y + i)) return false;
    return bytes_equal8(x + len - 8u, y + len - 8u);
}

// Exact check of one candidate window (offset off of text[0..n), hash pair
// h1,h2): find the survivors whose prefix-window hash pair is (h1,h2) in the
// open-addressed index, compare each whole pattern byte for byte, and call
// emit(id) for every pattern that occurs at off.
template <class Emit>
__device__ __forceinline__ void verify_candidate(const VerifyArgs &f, const uint8_t *text, uint64_t n, uint64_t off,
                                                 uint32_t h1, uint32_t h2, Emit emit)
{
    uint32_t slot = key_slot(h1, h2) & f.table_mask;
    for (;;) {
        const uint32_t t = f.table[slot];
        if (t == 0u) return;
        const uint2 ph = f.pat_hash[t - 1u];
        if (ph.x == h1 && ph.y == h2) {
            uint32_t id1 = t;
            uint32_t next = f.slot_chain[slot];
            while (id1) {
                const uint32_t id = id1 - 1u;
                const uint64_t p0 = f.off[id], len = f.off[id + 1] - p0;
                if (off + len <= n && bytes_equal(text + off, f.bytes + p0, len)) emit(id);
                id1 = next;
                if (next) next = f.chain_next[next - 1u];
            }
            return;
        }
        slot = (slot + 1u) & f.table_mask;
    }
}

}  // namespace ffbf
