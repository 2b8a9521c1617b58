// ffbf_scan_common.h -- per-unit text loading and hashing shared by the plain
// and batched scan kernels (internal).
//
// Work layout: the text is cut into 16-byte units aligned to 16 bytes in
// memory; one lane owns the 16 windows that start in one unit.  A lane loads
// its unit with one 128-bit non-temporal load, looks its bytes up in an LDS
// copy of the sbox, receives the sbox values of the following bytes from the
// lanes above and computes its 16 hash pairs as one full window plus 15
// rolling slides (hash_rot_sbox_2.h:34-51).
#pragma once

#include "ffbf_common.h"

namespace ffbf {

constexpr int kUnit = 16;

// (m & x) | (~m & y) as one v_bfi_b32.  Written in asm so that the compiler
// keeps a select tree over register arrays as selects: it would otherwise
// see an indexed read of the array and place the array in scratch memory.
static __device__ __forceinline__ uint32_t bit_select(uint32_t m, uint32_t x, uint32_t y)
{
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(x), "v"(y));
    return r;
}

// Per-wave survivor queue: capacity and the depth at which it is drained
// (64 entries at a time, one per lane).  A slot-by-slot append adds at most 64
// entries, so cap >= drain_at + 63; a bulk append is taken only when it fits.
template <int BLOCK>
struct QueueGeom {
    static constexpr int cap = 128;
    static constexpr int drain_at = 64;
};
// 1024-thread blocks beside a 128 KiB prefilter: 16 queues, sized so that
// the split method's 2 KiB sbox and the batched scan's chunk word still fit
// in 160 KiB
template <>
struct QueueGeom<1024> {
    static constexpr int cap = 119;
    static constexpr int drain_at = 56;
};
// 512-thread blocks: what LDS is left beside a 128 KiB prefilter, for filters
// that pass many windows
template <>
struct QueueGeom<512> {
    static constexpr int cap = 240;
    static constexpr int drain_at = 64;
};
// the global-prefilter scan passes ~1 window in 8: room for two drains' worth
template <>
struct QueueGeom<256> {
    static constexpr int cap = 256;
    static constexpr int drain_at = 64;
};

// One aligned 16-byte text chunk c (index relative to the 16-byte aligned base
// below the text); bytes outside [lead, lead+n) read as zero.
static __device__ __forceinline__ uint4 load_chunk(const uint8_t *base, uint64_t lead, uint64_t n, uint64_t c)
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

// Lane L receives lane (L+1) mod 64's value (DPP wavefront rotate).
static __device__ __forceinline__ uint32_t lane_rol1(uint32_t v)
{
    // every lane has a source in a rotation, so no "old" value is needed
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x134 /* wave_rol:1 */, 0xF, 0xF, false);
}

template <int W>
struct Chunks {
    static constexpr int NB = kUnit + W - 1;  // bytes behind a lane's 16 windows
    static constexpr int NCH = (NB + 15) / 16;
    static_assert(NCH <= 64 - (int)kIterUnits + 1, "window longer than the lanes above can supply");
};

// sbox values of the 16 bytes of one unit, from the LDS table.
template <int N>
static __device__ __forceinline__ void sbox_unit(const uint4 own, const uint32_t *sb, uint32_t (&s)[N], int at)
{
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int d = j >> 2;
        const uint32_t w = d == 0 ? own.x : d == 1 ? own.y : d == 2 ? own.z : own.w;
        s[at + j] = sb[(w >> (8 * (j & 3))) & 0xFFu];
    }
}

// Hash pair of the window whose bytes' sbox values are s[0..W): a sum of
// independent rotated terms (pairs fuse into three-input XORs).
template <int W, int N>
static __device__ __forceinline__ void first_window(const uint32_t (&s)[N], uint32_t &x1, uint32_t &x2)
{
    x1 = s[W - 1];
    x2 = s[W - 1];
#pragma unroll
    for (int j = 0; j + 1 < W - 1; j += 2) {
        x1 = xor3(x1, rotl(s[j], (uint32_t)((W - 1 - j) & 31)), rotl(s[j + 1], (uint32_t)((W - 2 - j) & 31)));
        x2 = xor3(x2, rotl(s[j], (uint32_t)((3 * (W - 1 - j)) & 31)), rotl(s[j + 1], (uint32_t)((3 * (W - 2 - j)) & 31)));
    }
    if constexpr (((W - 1) & 1) != 0) {
        x1 ^= rotl(s[W - 2], 1u);
        x2 ^= rotl(s[W - 2], 3u);
    }
}

// Windows b0+1 .. b0+15 from window b0's pair by rolling (hash_rot_sbox_2.h:34-51):
// slot i of h1/h2 gets window b0+i (slot 0 is window b0 itself).
template <int W, int N>
static __device__ __forceinline__ void roll16(const uint32_t (&s)[N], int b0, uint32_t &x1, uint32_t &x2,
                                              uint32_t (&h1)[kUnit], uint32_t (&h2)[kUnit])
{
    h1[0] = x1;
    h2[0] = x2;
#pragma unroll
    for (int i = 1; i < kUnit; i++) {
        const int b = b0 + i;
        x1 = xor3(rotl(x1, 1), s[b + W - 1], rotl(s[b - 1], (uint32_t)W));
        x2 = xor3(rotl(x2, 3), s[b + W - 1], rotl(s[b - 1], (uint32_t)(3 * W)));
        h1[i] = x1;
        h2[i] = x2;
    }
}

// Hash pairs of the 16 windows of a lane's unit (W compile-time).  The lane
// looks up the sbox values of its own 16 bytes in LDS; those of the next
// NB-16 bytes are the lanes above's, rotated down the wave with DPP (lane L
// gets lane L+1's values, L+2's on the second rotation): one LDS lookup per
// text byte.  Lanes kIterUnits..63 get wrapped-around values and own no
// windows.
template <int W>
static __device__ __forceinline__ void hash_unit(const uint4 own, const uint32_t *sb, uint32_t (&h1)[kUnit],
                                                 uint32_t (&h2)[kUnit])
{
    constexpr int NB = Chunks<W>::NB;
    uint32_t s[Chunks<W>::NCH * 16];
    sbox_unit(own, sb, s, 0);
#pragma unroll
    for (int j = 16; j < NB; j++) s[j] = lane_rol1(s[j - 16]);
    uint32_t x1, x2;
    first_window<W>(s, x1, x2);
    roll16<W>(s, 0, x1, x2, h1, h2);
}

// Two units per lane: the sbox values of the lane's own 32 bytes and of the
// W-1 bytes after them (the lane above's, by one DPP rotation each), then the
// 32 hash pairs as one full window and 31 slides, handed out 16 at a time.
// The first window is paid once per 32 windows instead of once per 16.
template <int W>
struct HashPair32 {
    static_assert(W >= 1 && W <= 33, "the lane above supplies at most 32 bytes");
    static constexpr int NB = 32 + W - 1;
    uint32_t s[NB];
    uint32_t x1 = 0, x2 = 0;

    __device__ __forceinline__ void load(const uint4 u0, const uint4 u1, const uint32_t *sb)
    {
        sbox_unit(u0, sb, s, 0);
        sbox_unit(u1, sb, s, 16);
#pragma unroll
        for (int j = 32; j < NB; j++) s[j] = lane_rol1(s[j - 32]);
    }
    // windows 0..15
    __device__ __forceinline__ void first(uint32_t (&h1)[kUnit], uint32_t (&h2)[kUnit])
    {
        first_window<W>(s, x1, x2);
        roll16<W>(s, 0, x1, x2, h1, h2);
    }
    // windows 16..31
    __device__ __forceinline__ void second(uint32_t (&h1)[kUnit], uint32_t (&h2)[kUnit])
    {
        x1 = xor3(rotl(x1, 1), s[16 + W - 1], rotl(s[15], (uint32_t)W));
        x2 = xor3(rotl(x2, 3), s[16 + W - 1], rotl(s[15], (uint32_t)(3 * W)));
        roll16<W>(s, 16, x1, x2, h1, h2);
    }
};

// Run-time W (1..64): bytes fetched one at a time.
static __device__ __forceinline__ uint32_t byte_at(const uint8_t *base, uint64_t lead, uint64_t n, uint64_t q)
{
    return (q >= lead && q < lead + n) ? (uint32_t)base[q] : 0u;
}

static __device__ __forceinline__ void hash_unit_generic(const uint8_t *base, uint64_t lead, uint64_t n, uint64_t u, uint32_t W,
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

// Bit i (0..3) set where byte i of w is '\n'.  Exact per byte: v's bytes
// are 0 exactly at newlines, and ((v & 0x7F) + 0x7F) | v has bit 7 clear only
// for a zero byte, with no carry between bytes; the multiply gathers bits
// 7, 15, 23, 31 into bits 21..24 (the partial products do not overlap).
static __device__ __forceinline__ uint32_t nl_nibble(uint32_t w)
{
    const uint32_t v = w ^ 0x0A0A0A0Au;
    const uint32_t z = ~(((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u;
    return (((z >> 7) * 0x00204081u) >> 21) & 0xFu;
}

// Split method: 64-bit hashes of the 16 windows of a lane's unit (low halves
// in h1, high in h2) from a 256-entry (lo, hi) sbox whose entries 0x80..0xFF
// equal entry 0 (ffbf.cpp:124-133), one full window plus 15 rolling slides
// (hash_rot_sbox_pre_2.h:41-50: h = rol64(h,1) ^ S[in] ^ rol64(S[out], W)).
// Returns the mask of windows that hold no newline: only those are tested,
// as the reference restarts the hash at every '\n' (ffbf.cpp:342-358) and a
// hash is tested once W bytes followed the restart.
template <int W>
static __device__ __forceinline__ uint32_t hash_unit64(const uint4 own, const uint2 *sb, uint32_t (&h1)[kUnit],
                                                       uint32_t (&h2)[kUnit])
{
    constexpr int NB = Chunks<W>::NB;
    constexpr int NCH = Chunks<W>::NCH;
    static_assert(NCH <= 4, "the newline mask spans 64 bytes");
    uint32_t lo[NCH * 16], hi[NCH * 16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int d = j >> 2;
        const uint32_t w = d == 0 ? own.x : d == 1 ? own.y : d == 2 ? own.z : own.w;
        const uint2 v = sb[(w >> (8 * (j & 3))) & 0xFFu];
        lo[j] = v.x;
        hi[j] = v.y;
    }
#pragma unroll
    for (int j = 16; j < NB; j++) {
        lo[j] = lane_rol1(lo[j - 16]);
        hi[j] = lane_rol1(hi[j - 16]);
    }
    uint32_t xl = lo[W - 1], xh = hi[W - 1];
#pragma unroll
    for (int j = 0; j + 1 < W - 1; j += 2) {
        uint32_t al, ah, bl, bh;
        rol64(lo[j], hi[j], (uint32_t)(W - 1 - j), al, ah);
        rol64(lo[j + 1], hi[j + 1], (uint32_t)(W - 2 - j), bl, bh);
        xl = xor3(xl, al, bl);
        xh = xor3(xh, ah, bh);
    }
    if constexpr (((W - 1) & 1) != 0) {
        uint32_t al, ah;
        rol64(lo[W - 2], hi[W - 2], 1u, al, ah);
        xl ^= al;
        xh ^= ah;
    }
    h1[0] = xl;
    h2[0] = xh;
#pragma unroll
    for (int b = 1; b < kUnit; b++) {
        uint32_t rl, rh, ol, oh;
        rol64(xl, xh, 1u, rl, rh);
        rol64(lo[b - 1], hi[b - 1], (uint32_t)W, ol, oh);
        xl = xor3(rl, lo[b + W - 1], ol);
        xh = xor3(rh, hi[b + W - 1], oh);
        h1[b] = xl;
        h2[b] = xh;
    }
    // newlines in the bytes behind the 16 windows: own unit, then the units
    // of the lanes above; bit i of `bad` marks a newline in [i, i+W)
    uint32_t t = nl_nibble(own.x) | (nl_nibble(own.y) << 4) | (nl_nibble(own.z) << 8) | (nl_nibble(own.w) << 12);
    uint64_t m = t;
#pragma unroll
    for (int c = 1; c < NCH; c++) {
        t = lane_rol1(t);
        m |= (uint64_t)t << (16 * c);
    }
    uint64_t bad = m;
    int cov = 1;  // bad covers shifts [0, cov)
#pragma unroll
    for (; 2 * cov <= W; cov *= 2) bad |= bad >> cov;
    if (cov < W) bad |= bad >> (W - cov);
    return ~(uint32_t)bad & 0xFFFFu;
}

// Run-time W for the split method: bytes fetched one at a time.
static __device__ __forceinline__ uint32_t hash_unit_generic64(const uint8_t *base, uint64_t lead, uint64_t n,
                                                               uint64_t u, uint32_t W, const uint2 *sb,
                                                               uint32_t (&h1)[kUnit], uint32_t (&h2)[kUnit])
{
    const uint64_t q0 = u * 16u;
    uint32_t xl = 0, xh = 0;
    int last_nl = -1;  // position of the last newline seen, relative to q0
    for (uint32_t j = 0; j < W; j++) {
        const uint32_t c = byte_at(base, lead, n, q0 + j);
        if (c == 10u) last_nl = (int)j;
        const uint2 v = sb[c];
        uint32_t rl, rh;
        rol64(xl, xh, 1u, rl, rh);
        xl = rl ^ v.x;
        xh = rh ^ v.y;
    }
    uint32_t ok = last_nl < 0 ? 1u : 0u;
    h1[0] = xl;
    h2[0] = xh;
#pragma unroll
    for (int b = 1; b < kUnit; b++) {
        const uint32_t ci = byte_at(base, lead, n, q0 + (uint64_t)b + W - 1);
        if (ci == 10u) last_nl = b + (int)W - 1;
        const uint2 si = sb[ci], so = sb[byte_at(base, lead, n, q0 + (uint64_t)b - 1)];
        uint32_t rl, rh, ol, oh;
        rol64(xl, xh, 1u, rl, rh);
        rol64(so.x, so.y, W, ol, oh);
        xl = rl ^ si.x ^ ol;
        xh = rh ^ si.y ^ oh;
        h1[b] = xl;
        h2[b] = xh;
        ok |= (last_nl < b ? 1u : 0u) << b;
    }
    return ok;
}

// A lane's candidate not yet appended.  Appends go out in bulk: a wave
// flushes all its held candidates when a new hit lands on a lane that still
// holds one, and at the end of its work.  A flush writes into the wave's own
// region of the raw candidate buffer (no atomics); only a wave whose region
// is full takes slots in the shared spill area, with one counter atomic (a
// device-wide atomic on one address that every CU would otherwise hit).
struct PendingCand {
    uint64_t off = 0;
    uint32_t h1 = 0, h2 = 0;
    bool valid = false;
    uint32_t kept = 0;  // wave-uniform: entries written to the wave's region
};

static __device__ __forceinline__ uint32_t region_slot(const ScanArgs &a)
{
    return a.wave0 + blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
}

// Called by the whole wave.
static __device__ __forceinline__ void flush_pending(PendingCand &pc, const ScanArgs &a)
{
    const uint64_t hb = __ballot(pc.valid);
    if (!hb) return;
    const uint32_t n = (uint32_t)__builtin_popcountll(hb);
    const uint32_t slot_id = region_slot(a);
    uint64_t at;
    if (slot_id < a.region_end && pc.kept + n <= a.region) {
        at = (uint64_t)slot_id * a.region + pc.kept;
        pc.kept += n;
    } else {
        const int leader = __builtin_ctzll(hb);
        unsigned long long first = 0;
        if (lane_id() == leader) first = atomicAdd(a.counters + C_SPILL, (unsigned long long)n);
        at = a.spill_base + readlane64(first, leader);
    }
    if (pc.valid) {
        const uint64_t e = at + lanes_below(hb);
        if (e < a.cand_cap) {
            Cand c;
            c.offset = pc.off;
            c.h1 = pc.h1;
            c.h2 = pc.h2;
            a.cand[e] = c;
        }
    }
    pc.valid = false;
}

// End of a wave's work: flush, then record the region's fill for the
// compaction pass (counts start at zero, so an empty region writes nothing).
static __device__ __forceinline__ void finish_pending(PendingCand &pc, const ScanArgs &a)
{
    flush_pending(pc, a);
    if (pc.kept && lane_id() == 0) a.wave_counts[region_slot(a)] = pc.kept;  // kept > 0: slot < region_end
}

// Called by the whole wave.
static __device__ __forceinline__ void stage_cand(PendingCand &pc, bool hit, uint64_t off, uint32_t h1, uint32_t h2,
                                                  const ScanArgs &a)
{
    if (__ballot(hit && pc.valid)) flush_pending(pc, a);
    if (hit) {
        pc.off = off;
        pc.h1 = h1;
        pc.h2 = h2;
        pc.valid = true;
    }
}

// Finishes up to 64 queued windows, one per lane: the prefilter bits not yet
// tested, then (for a fold) the k forward bits of F, then feed-forward and
// candidate staging.  Called by the whole wave.
template <bool LDSPRE, int P, bool SPLIT = false>
static __device__ __forceinline__ void drain(const uint4 *q, uint32_t cnt, const ScanArgs &a, uint32_t *s_pre,
                                             PendingCand &pend)
{
    const bool act = (uint32_t)lane_id() < cnt;
    uint4 e = make_uint4(0u, 0u, 0u, 0u);
    if (act) e = q[lane_id()];
    const uint32_t h1 = e.z, h2 = e.w;
    uint32_t ok = act ? 1u : 0u;
    if constexpr (SPLIT) {
        // the fold covers the cache part: its bits P..s-1, then F (the
        // global prefilter already was F's bit 0), one bit for every window
        // still passing and the rest only where it is set
        if constexpr (LDSPRE) {
            for (int j = P; j < kSplitMaxK; j++) {
                if (j >= (int)a.split.s) break;
                ok &= lds_bit(s_pre, split_fwd_fn(j, h1, h2), a.pre_log2w);
            }
        }
        const int j0 = LDSPRE ? 0 : 1;
        if (ok && j0 < (int)a.k) {
            const uint32_t i0 = split_place(j0, split_fwd_fn(j0, h1, h2), a.split);
            ok &= (a.F[i0 >> 5] >> (i0 & 31u)) & 1u;
        }
        if (ok) {
            for (int j = j0 + 1; j < kSplitMaxK; j++) {
                if (j >= (int)a.k) break;
                const uint32_t ix = split_place(j, split_fwd_fn(j, h1, h2), a.split);
                ok &= (a.F[ix >> 5] >> (ix & 31u)) & 1u;
            }
        }
        const bool hit = ok != 0u;
        const uint64_t off = ((uint64_t)e.y << 32) | e.x;
        feed_forward_split(hit, h1, h2, a.k, a.split, a.R);
        stage_cand(pend, hit, off, h1, h2, a);
        return;
    }
    if constexpr (LDSPRE) {
        // bits P..k-1 of the LDS prefilter, loads issued back to back
        uint32_t bits[5];
#pragma unroll
        for (int j = P; j < 5; j++) bits[j] = lds_bit(s_pre, fwd_idx(j, h1, h2), a.pre_log2w);
#pragma unroll
        for (int j = P; j < 5; j++) ok &= ((uint32_t)j < a.k) ? bits[j] : 1u;
        if (a.pre2) {
            // the L2-resident fold: its first untested bit j0 for every window
            // still passing (bit 0 was tested before queueing when f_first is
            // set), the rest only where that one is set
            const uint32_t j0 = a.f_first ? 1u : 0u;
            if (ok && j0 < a.k) {
                const uint32_t i0 = fwd_idx((int)j0, h1, h2) & a.pre2_mask;
                ok &= (a.pre2[i0 >> 5] >> (i0 & 31u)) & 1u;
            }
            if (ok && a.k > j0 + 1u) {
                uint32_t words[5], idx[5];
#pragma unroll
                for (int j = 1; j < 5; j++) {
                    idx[j] = fwd_idx(j, h1, h2) & a.pre2_mask;
                    words[j] = ((uint32_t)j > j0 && (uint32_t)j < a.k) ? a.pre2[idx[j] >> 5] : 0xFFFFFFFFu;
                }
#pragma unroll
                for (int j = 1; j < 5; j++) ok &= (words[j] >> (idx[j] & 31u)) & 1u;
            }
        }
    } else {
        // the global prefilter tested bit 0 only; bit 1 next, then the rest
        // only for the few lanes still passing (each step is one round trip
        // to L2, and L2 requests are what bounds this mode)
        if (act && a.k > 1u) {
            const uint32_t i1 = fwd_idx(1, h1, h2) & a.pre_mask;
            ok &= (a.pre[i1 >> 5] >> (i1 & 31u)) & 1u;
        }
        if (ok && a.k > 2u) {
            uint32_t words[5], idx[5];
#pragma unroll
            for (int j = 2; j < 5; j++) {
                idx[j] = fwd_idx(j, h1, h2) & a.pre_mask;
                words[j] = ((uint32_t)j < a.k) ? a.pre[idx[j] >> 5] : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int j = 2; j < 5; j++) ok &= (words[j] >> (idx[j] & 31u)) & 1u;
        }
    }
    if (!a.pre_exact) {
        // F itself: its first bit for every window still passing, the rest
        // only where that bit is set.  F is too big for L2 whenever a fold is
        // in use, and its fill is low, so this takes ~1 random access per
        // window instead of k
        if (ok && !(LDSPRE && a.f_first && !a.pre2)) {  // F's own first bit may have been tested before queueing
            const uint32_t i0 = h1 & a.mask;
            ok &= (a.F[i0 >> 5] >> (i0 & 31u)) & 1u;
        }
        if (ok && a.k > 1u) {
            uint32_t words[5], idx[5];
#pragma unroll
            for (int j = 1; j < 5; j++) {
                idx[j] = fwd_idx(j, h1, h2) & a.mask;
                words[j] = ((uint32_t)j < a.k) ? a.F[idx[j] >> 5] : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int j = 1; j < 5; j++) ok &= (words[j] >> (idx[j] & 31u)) & 1u;
        }
    }
    const bool hit = ok != 0u;
    const uint64_t off = ((uint64_t)e.y << 32) | e.x;
    if (LDSPRE && a.r_lds)
        feed_forward_lds(hit, h1, h2, a.k, a.mask, s_pre + a.pre_words);
    else
        feed_forward(hit, h1, h2, a.k, a.mask, a.R);
    stage_cand(pend, hit, off, h1, h2, a);
}

// Prefilter test of one lane's 16 windows (window b starts at text offset
// p0+b; only offsets in [0, npos) are real) and wave-aggregated push of the
// survivors onto the wave's LDS queue, draining it 64 at a time.  `tag` is
// added to the recorded offsets (the offset of a text piece, or the batched
// scan's text id in the top bits).  Called by the whole wave.
template <bool LDSPRE, int BLOCK, int P, bool SPLIT = false, int OWN_LANES = (int)kIterUnits>
static __device__ __forceinline__ void test_and_queue(const ScanArgs &a, uint32_t *s_pre, const uint32_t (&h1)[kUnit],
                                                      const uint32_t (&h2)[kUnit], int64_t p0, uint64_t npos,
                                                      uint64_t tag, uint4 *q, uint32_t &qn, PendingCand &pend,
                                                      uint32_t line_ok = 0xFFFFu)
{
    using QG = QueueGeom<BLOCK>;
    // the real windows of this lane (0 <= p0 + b < npos) are one run [lo, hi)
    const int64_t lo64 = p0 < 0 ? -p0 : 0;
    const int64_t hi64 = (int64_t)npos - p0;
    const uint32_t lo = (uint32_t)(lo64 < kUnit ? lo64 : kUnit);
    const uint32_t hi = (uint32_t)(hi64 <= 0 ? 0 : (hi64 < kUnit ? hi64 : kUnit));
    uint32_t vm = hi > lo ? ((0xFFFFu >> (kUnit - hi)) & (0xFFFFu << lo)) : 0u;
    if (lane_id() >= OWN_LANES) vm = 0;  // these units belong to the next wave-iteration
    vm &= line_ok;
    uint32_t pm = 0;
    if constexpr (LDSPRE) {
        // P bits per window, all loads in flight together
        uint32_t bits[P][kUnit];
#pragma unroll
        for (int j = 0; j < P; j++)
#pragma unroll
            for (int b = 0; b < kUnit; b++)
                bits[j][b] = lds_bit(s_pre, SPLIT ? split_fwd_fn(j, h1[b], h2[b]) : fwd_idx(j, h1[b], h2[b]),
                                     a.pre_log2w);
#pragma unroll
        for (int b = 0; b < kUnit; b++) {
            uint32_t ok = bits[0][b];
#pragma unroll
            for (int j = 1; j < P; j++) ok &= bits[j][b];
            pm |= ok << b;
        }
    } else {
        uint32_t fw[kUnit];
#pragma unroll
        for (int b = 0; b < kUnit; b++) fw[b] = a.pre[(h1[b] & a.pre_mask) >> 5];
#pragma unroll
        for (int b = 0; b < kUnit; b++) pm |= ((fw[b] >> (h1[b] & 31u)) & 1u) << b;
    }
    pm &= vm;
    if constexpr (LDSPRE && !SPLIT) {
        // a fold that passes many windows: the first forward bit (h1) of F,
        // or of the L2-resident fold behind it, now, so that only windows
        // passing it are queued
        if (a.f_first && pm) {
            if (a.pre2 && (uint32_t)P < a.k) {
                // the LDS fold's next bit first: each window it rejects saves
                // an L2 request, and L2 requests bound this mode (the drain
                // tests the bit again, to the same result)
                uint32_t nb = 0;
#pragma unroll
                for (int b = 0; b < kUnit; b++) nb |= lds_bit(s_pre, fwd_idx(P, h1[b], h2[b]), a.pre_log2w) << b;
                pm &= nb;
            }
            const uint32_t *src = a.pre2 ? a.pre2 : a.F;
            const uint32_t smask = a.pre2 ? a.pre2_mask : a.mask;
            uint32_t fw[kUnit];
#pragma unroll
            for (int b = 0; b < kUnit; b++) fw[b] = ((pm >> b) & 1u) ? src[(h1[b] & smask) >> 5] : 0xFFFFFFFFu;
            uint32_t keep = 0;
#pragma unroll
            for (int b = 0; b < kUnit; b++) keep |= ((fw[b] >> (h1[b] & 31u)) & 1u) << b;
            pm &= keep;
        }
    }
    const uint32_t mine = (uint32_t)__builtin_popcount(pm);
    const uint32_t incl = wave_incl_sum(mine);
    const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
    if (total == 0u) return;
    if (qn + total <= (uint32_t)QG::cap) {
        // room for every survivor: each lane writes its own at its rank in
        // the wave (exclusive prefix of the per-lane counts), no ballots
        uint4 *dst = q + (qn + incl - mine);
        // p0 itself may be negative on a text's first unit; p0 + b is not
        // for the windows written
        const uint64_t pt = tag + (uint64_t)p0;
        // Most lanes pass at most one window: every lane's first survivor
        // takes its hash pair through a 4-level select on its slot number
        // (no per-slot branches), then the slots of second and later
        // survivors, where any lane has one.
        {
            const uint32_t b1 = (uint32_t)__builtin_ctz(pm | 0x10000u);  // 16 (slot 0 bits) when none
            uint32_t s1[8], s2[8];
            const uint32_t m3 = 0u - ((b1 >> 3) & 1u), m2 = 0u - ((b1 >> 2) & 1u), m1 = 0u - ((b1 >> 1) & 1u),
                           m0 = 0u - (b1 & 1u);
#pragma unroll
            for (int i = 0; i < 8; i++) {
                s1[i] = bit_select(m3, h1[i + 8], h1[i]);
                s2[i] = bit_select(m3, h2[i + 8], h2[i]);
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                s1[i] = bit_select(m2, s1[i + 4], s1[i]);
                s2[i] = bit_select(m2, s2[i + 4], s2[i]);
            }
#pragma unroll
            for (int i = 0; i < 2; i++) {
                s1[i] = bit_select(m1, s1[i + 2], s1[i]);
                s2[i] = bit_select(m1, s2[i + 2], s2[i]);
            }
            const uint32_t v1 = bit_select(m0, s1[1], s1[0]), v2 = bit_select(m0, s2[1], s2[0]);
            if (pm) {
                const uint64_t p = pt + (uint64_t)b1;
                *dst = make_uint4((uint32_t)p, (uint32_t)(p >> 32), v1, v2);
                ++dst;
            }
        }
        const uint32_t rest = pm & (pm - 1u);
        for (uint32_t any = wave_or(rest); any; any &= any - 1u) {
            const uint32_t b = (uint32_t)__builtin_ctz(any);  // wave-uniform: registers indexed directly
            if ((rest >> b) & 1u) {
                const uint64_t p = pt + (uint64_t)b;
                *dst = make_uint4((uint32_t)p, (uint32_t)(p >> 32), h1[b], h2[b]);
                ++dst;
            }
        }
        qn += total;
        while (qn >= (uint32_t)QG::drain_at) {
            const uint32_t cnt = qn < 64u ? qn : 64u;
            wave_sync();
            drain<LDSPRE, P, SPLIT>(q + (qn - cnt), cnt, a, s_pre, pend);
            wave_sync();
            qn -= cnt;
        }
        return;
    }
    // more survivors than queue room: slot by slot, draining as it fills.
    // Visit only the slots where some lane passed; b is wave-uniform, so
    // h1[b]/h2[b] index registers directly (no scratch)
    for (uint32_t any = wave_or(pm); any; any &= any - 1u) {
        const uint32_t b = (uint32_t)__builtin_ctz(any);
        const bool pass = (pm >> b) & 1u;
        const uint64_t bal = __ballot(pass);
        if (pass) {
            const uint64_t p = tag + (uint64_t)(p0 + (int64_t)b);
            q[qn + lanes_below(bal)] = make_uint4((uint32_t)p, (uint32_t)(p >> 32), h1[b], h2[b]);
        }
        qn += (uint32_t)__builtin_popcountll(bal);
        if (qn >= (uint32_t)QG::drain_at) {
            const uint32_t cnt = qn < 64u ? qn : 64u;
            wave_sync();
            drain<LDSPRE, P, SPLIT>(q + (qn - cnt), cnt, a, s_pre, pend);
            wave_sync();
            qn -= cnt;
        }
    }
}

}  // namespace ffbf
