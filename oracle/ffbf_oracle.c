/*
 * ffbf_oracle.c -- CPU restatement of the reference FFBF matcher.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP kernels are
 * compared against (tests/, the smoke check) and the CPU baseline timed by
 * bench.py.  The shipped library never links, loads or calls it.
 *
 * Each function cites the reference lines it restates.  Deliberate
 * differences from the reference (documented, not reproduced):
 *   - only offsets o in [0, n-W] are evaluated; the reference GPU kernel also
 *     hashes the last W-1 offsets over uninitialised bytes (greptest.cu:143-152);
 *   - offsets are 64-bit (the reference uses int, greptest.cu:134);
 *   - rotation by a multiple of 32 is the identity (the reference rol32 shifts
 *     by 32, which C leaves undefined, hash_common.h:8-11).
 */
#include "ffbf_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static inline uint32_t rotl32(uint32_t x, uint32_t r)
{
    r &= 31u;
    return r ? (x << r) | (x >> (32u - r)) : x;
}

static inline uint64_t filter_words(uint32_t log2_bits)
{
    return log2_bits >= 5 ? (1ull << (log2_bits - 5)) : 1ull;
}

static inline uint32_t filter_mask(uint32_t log2_bits)
{
    return log2_bits >= 32 ? 0xFFFFFFFFu : ((1u << log2_bits) - 1u);
}

static inline void set_bit(uint32_t *v, uint32_t b) { v[b >> 5] |= 1u << (b & 31u); }
static inline int test_bit(const uint32_t *v, uint32_t b) { return (v[b >> 5] >> (b & 31u)) & 1u; }

/* greptest.cu:106-114 (also 157-166, 264-272): h1 = rotl(h1,1)^S[c], h2 = rotl(h2,3)^S[c]. */
void orc_hash_window(const orc_params *p, const uint8_t *win, uint32_t *h1, uint32_t *h2)
{
    uint32_t a = 0, b = 0;
    for (uint32_t j = 0; j < p->window; j++) {
        uint32_t s = p->sbox[win[j]];
        a = rotl32(a, 1) ^ s;
        b = rotl32(b, 3) ^ s;
    }
    *h1 = a;
    *h2 = b;
}

/* hash_rot_sbox_2.h:34-51: after W bytes the byte leaving the window is
 * removed with rotl(S[out], W) from h1 and rotl(S[out], 3W) from h2
 * (rotate_old = 2*LEN % 32 applied on top of the LEN rotation, :18, :44-46). */
void orc_hash_rolling(const orc_params *p, const uint8_t *text, uint64_t n,
                      uint32_t *h1, uint32_t *h2)
{
    const uint32_t W = p->window;
    uint32_t a = 0, b = 0;
    for (uint64_t i = 0; i < n; i++) {
        uint32_t s = p->sbox[text[i]];
        a = rotl32(a, 1) ^ s;
        b = rotl32(b, 3) ^ s;
        if (i >= W) {
            uint32_t so = p->sbox[text[i - W]];
            a ^= rotl32(so, W);
            b ^= rotl32(so, 3u * W);
        }
        if (i + 1 >= W) {
            h1[i + 1 - W] = a;
            h2[i + 1 - W] = b;
        }
    }
}

/* greptest.cu:115-120 and 199-202; the fifth function is the commented-out
 * (h1<<16)|(h2>>16) of greptest.cu:121-122, 203. */
void orc_fwd_indices(const orc_params *p, uint32_t h1, uint32_t h2, uint32_t *out)
{
    const uint32_t m = filter_mask(p->log2_bits);
    uint32_t f[5];
    f[0] = h1;
    f[1] = h2;
    f[2] = h1 + h2;
    f[3] = h1 + 5u * h2;
    f[4] = (h1 << 16) | (h2 >> 16);
    for (uint32_t i = 0; i < p->k; i++) out[i] = f[i] & m;
}

/* greptest.cu:231-234 and 273-276; the fifth is the commented-out
 * h1*11+h2 of greptest.cu:235, 277. */
void orc_rev_indices(const orc_params *p, uint32_t h1, uint32_t h2, uint32_t *out)
{
    const uint32_t m = filter_mask(p->log2_bits);
    uint32_t r[5];
    r[0] = h1 + 7u * h2;
    r[1] = h1 + 3u * h2;
    r[2] = (h1 << 1) + h2;
    r[3] = (h1 << 2) + h2;
    r[4] = h1 * 11u + h2;
    for (uint32_t i = 0; i < p->k; i++) out[i] = r[i] & m;
}

/* greptest.cu:455-474 (clear F, then grepSetup :96-123 per pattern) with the
 * CPU flow's skip of patterns shorter than W (ffbf.cpp:268-283). */
uint64_t orc_build(const orc_params *p, const uint8_t *bytes, const uint64_t *off,
                   uint64_t n_patterns, uint32_t *F)
{
    memset(F, 0, filter_words(p->log2_bits) * sizeof(uint32_t));
    uint64_t hashed = 0;
    for (uint64_t i = 0; i < n_patterns; i++) {
        if (off[i + 1] - off[i] < p->window) continue;
        uint32_t h1, h2, idx[5];
        orc_hash_window(p, bytes + off[i], &h1, &h2);
        orc_fwd_indices(p, h1, h2, idx);
        for (uint32_t j = 0; j < p->k; j++) set_bit(F, idx[j]);
        hashed++;
    }
    return hashed;
}

/* Scan of text[lo .. hi) window starts, rolling hash as in the CPU flow
 * (ffbf.cpp:340-370, hash_rot_sbox_2.h:34-51), probe cascade and
 * feed-forward as in GrepKernel (greptest.cu:199-246).  R is not cleared. */
static uint64_t scan_range(const orc_params *p, const uint8_t *text, uint64_t n,
                           uint64_t lo, uint64_t hi, const uint32_t *F, uint32_t *R,
                           uint64_t *cand, uint64_t cap)
{
    const uint32_t W = p->window;
    uint64_t found = 0;
    if (n < W || lo >= hi) return 0;
    uint32_t a = 0, b = 0;
    /* prime with the first W-1 bytes of the window starting at lo */
    for (uint64_t i = lo; i < lo + W - 1; i++) {
        uint32_t s = p->sbox[text[i]];
        a = rotl32(a, 1) ^ s;
        b = rotl32(b, 3) ^ s;
    }
    for (uint64_t o = lo; o < hi; o++) {
        uint32_t s = p->sbox[text[o + W - 1]];
        a = rotl32(a, 1) ^ s;
        b = rotl32(b, 3) ^ s;
        if (o > lo) {
            uint32_t so = p->sbox[text[o - 1]];
            a ^= rotl32(so, W);
            b ^= rotl32(so, 3u * W);
        }
        uint32_t idx[5];
        orc_fwd_indices(p, a, b, idx);
        int hit = 1;
        for (uint32_t j = 0; j < p->k; j++) {
            if (!test_bit(F, idx[j])) { hit = 0; break; }
        }
        if (!hit) continue;
        orc_rev_indices(p, a, b, idx);
        for (uint32_t j = 0; j < p->k; j++) set_bit(R, idx[j]);
        if (found < cap) cand[found] = o;
        found++;
    }
    return found;
}

/* executeGrepOverlap (greptest.cu:505-576) clears R (:528-529) then runs
 * GrepKernel over the text; here only offsets [0, n-W] are evaluated. */
uint64_t orc_scan(const orc_params *p, const uint8_t *text, uint64_t n,
                  const uint32_t *F, uint32_t *R, uint64_t *cand, uint64_t cap)
{
    memset(R, 0, filter_words(p->log2_bits) * sizeof(uint32_t));
    if (n < p->window) return 0;
    return scan_range(p, text, n, 0, n - p->window + 1, F, R, cand, cap);
}

typedef struct scan_job {
    const orc_params *p;
    const uint8_t *text;
    uint64_t n, lo, hi;
    const uint32_t *F;
    uint32_t *R;
    uint64_t *cand;
    uint64_t cap, found;
} scan_job;

static void *scan_worker(void *arg)
{
    scan_job *j = (scan_job *)arg;
    j->found = scan_range(j->p, j->text, j->n, j->lo, j->hi, j->F, j->R, j->cand, j->cap);
    return NULL;
}

uint64_t orc_scan_mt(const orc_params *p, const uint8_t *text, uint64_t n,
                     const uint32_t *F, uint32_t *R, uint64_t *cand, uint64_t cap,
                     int nthreads)
{
    const uint64_t words = filter_words(p->log2_bits);
    memset(R, 0, words * sizeof(uint32_t));
    if (n < p->window) return 0;
    const uint64_t npos = n - p->window + 1;
    if (nthreads < 1) nthreads = 1;
    if ((uint64_t)nthreads > npos) nthreads = (int)npos;
    scan_job *jobs = (scan_job *)calloc((size_t)nthreads, sizeof(scan_job));
    pthread_t *tid = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    const uint64_t per = (npos + (uint64_t)nthreads - 1) / (uint64_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        scan_job *j = &jobs[t];
        j->p = p; j->text = text; j->n = n; j->F = F;
        j->lo = (uint64_t)t * per;
        j->hi = j->lo + per < npos ? j->lo + per : npos;
        j->R = t == 0 ? R : (uint32_t *)calloc(words, sizeof(uint32_t));
        /* each thread keeps its own candidates; capacity per thread = its range */
        j->cap = j->hi > j->lo ? j->hi - j->lo : 0;
        if (j->cap > cap) j->cap = cap;
        j->cand = (uint64_t *)malloc((j->cap ? j->cap : 1) * sizeof(uint64_t));
        pthread_create(&tid[t], NULL, scan_worker, j);
    }
    uint64_t total = 0;
    for (int t = 0; t < nthreads; t++) {
        scan_job *j = &jobs[t];
        pthread_join(tid[t], NULL);
        if (t > 0) {
            for (uint64_t w = 0; w < words; w++) R[w] |= j->R[w];
            free(j->R);
        }
        uint64_t kept = j->found < j->cap ? j->found : j->cap;
        for (uint64_t i = 0; i < kept && total + i < cap; i++) cand[total + i] = j->cand[i];
        total += j->found;
        free(j->cand);
    }
    free(jobs);
    free(tid);
    return total;
}

/* filterPatterns (greptest.cu:253-290): each pattern's prefix hash is tested
 * against R's reverse bits; skip-short and 0-based ids as ffbf.cpp:404-461. */
uint64_t orc_pattern_pass(const orc_params *p, const uint8_t *bytes, const uint64_t *off,
                          uint64_t n_patterns, const uint32_t *R, uint32_t *surv_bitmap)
{
    memset(surv_bitmap, 0, ((n_patterns + 31) / 32) * sizeof(uint32_t));
    uint64_t survivors = 0;
    for (uint64_t i = 0; i < n_patterns; i++) {
        if (off[i + 1] - off[i] < p->window) continue;
        uint32_t h1, h2, idx[5];
        orc_hash_window(p, bytes + off[i], &h1, &h2);
        orc_rev_indices(p, h1, h2, idx);
        int ok = 1;
        for (uint32_t j = 0; j < p->k; j++) {
            if (!test_bit(R, idx[j])) { ok = 0; break; }
        }
        if (ok) {
            surv_bitmap[i >> 5] |= 1u << (i & 31u);
            survivors++;
        }
    }
    return survivors;
}

typedef struct keyed { uint64_t key; uint32_t id; uint32_t pad; } keyed;

static int cmp_keyed(const void *x, const void *y)
{
    const keyed *a = (const keyed *)x, *b = (const keyed *)y;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return a->id < b->id ? -1 : (a->id > b->id);
}

typedef struct match_rec { uint32_t id; uint32_t pad; uint64_t off; } match_rec;

static int cmp_match(const void *x, const void *y)
{
    const match_rec *a = (const match_rec *)x, *b = (const match_rec *)y;
    if (a->id != b->id) return a->id < b->id ? -1 : 1;
    return a->off < b->off ? -1 : (a->off > b->off);
}

typedef void (*window_hash_fn)(const void *p, const uint8_t *win, uint32_t *h1, uint32_t *h2);

/* Exact verification.  The reference hands survivors and filtered lines to
 * `grep -Ff` (src/rabin-karp/README:49-52); this restates what that step
 * decides at (pattern id, offset) granularity: join candidates to survivors
 * on the prefix-window hash pair, then compare every byte. */
static uint64_t verify_core(const void *p, window_hash_fn hash, uint32_t W, const uint8_t *text, uint64_t n,
                            const uint8_t *bytes, const uint64_t *off, uint64_t n_patterns,
                            const uint64_t *cand, uint64_t n_cand, const uint32_t *surv_bitmap,
                            uint32_t *out_id, uint64_t *out_off, uint64_t cap)
{
    uint64_t ns = 0;
    for (uint64_t i = 0; i < n_patterns; i++)
        if ((surv_bitmap[i >> 5] >> (i & 31u)) & 1u) ns++;
    keyed *tab = (keyed *)malloc((ns ? ns : 1) * sizeof(keyed));
    uint64_t t = 0;
    for (uint64_t i = 0; i < n_patterns; i++) {
        if (!((surv_bitmap[i >> 5] >> (i & 31u)) & 1u)) continue;
        uint32_t h1, h2;
        hash(p, bytes + off[i], &h1, &h2);
        tab[t].key = ((uint64_t)h1 << 32) | h2;
        tab[t].id = (uint32_t)i;
        tab[t].pad = 0;
        t++;
    }
    qsort(tab, (size_t)ns, sizeof(keyed), cmp_keyed);

    uint64_t mcap = 1024, nm = 0;
    match_rec *m = (match_rec *)malloc(mcap * sizeof(match_rec));
    for (uint64_t c = 0; c < n_cand; c++) {
        uint64_t o = cand[c];
        if (o + W > n) continue;
        uint32_t h1, h2;
        hash(p, text + o, &h1, &h2);
        uint64_t key = ((uint64_t)h1 << 32) | h2;
        /* lower bound */
        uint64_t lo = 0, hi = ns;
        while (lo < hi) {
            uint64_t mid = (lo + hi) / 2;
            if (tab[mid].key < key) lo = mid + 1; else hi = mid;
        }
        for (uint64_t s = lo; s < ns && tab[s].key == key; s++) {
            uint32_t id = tab[s].id;
            uint64_t len = off[id + 1] - off[id];
            if (o + len > n) continue;
            if (memcmp(text + o, bytes + off[id], (size_t)len) != 0) continue;
            if (nm == mcap) {
                mcap *= 2;
                m = (match_rec *)realloc(m, mcap * sizeof(match_rec));
            }
            m[nm].id = id;
            m[nm].pad = 0;
            m[nm].off = o;
            nm++;
        }
    }
    qsort(m, (size_t)nm, sizeof(match_rec), cmp_match);
    for (uint64_t i = 0; i < nm && i < cap; i++) {
        out_id[i] = m[i].id;
        out_off[i] = m[i].off;
    }
    free(m);
    free(tab);
    return nm;
}

static void pair_hash(const void *p, const uint8_t *win, uint32_t *h1, uint32_t *h2)
{
    orc_hash_window((const orc_params *)p, win, h1, h2);
}

uint64_t orc_verify(const orc_params *p, const uint8_t *text, uint64_t n,
                    const uint8_t *bytes, const uint64_t *off, uint64_t n_patterns,
                    const uint64_t *cand, uint64_t n_cand, const uint32_t *surv_bitmap,
                    uint32_t *out_id, uint64_t *out_off, uint64_t cap)
{
    return verify_core(p, pair_hash, p->window, text, n, bytes, off, n_patterns, cand, n_cand, surv_bitmap,
                       out_id, out_off, cap);
}

uint64_t orc_brute_matches(const uint8_t *text, uint64_t n,
                           const uint8_t *bytes, const uint64_t *off, uint64_t n_patterns,
                           uint32_t min_len, uint32_t *out_id, uint64_t *out_off, uint64_t cap)
{
    uint64_t nm = 0;
    for (uint64_t i = 0; i < n_patterns; i++) {
        const uint64_t len = off[i + 1] - off[i];
        if (len < min_len || len == 0 || len > n) continue;
        const uint8_t *pat = bytes + off[i];
        for (uint64_t o = 0; o + len <= n; o++) {
            if (text[o] != pat[0]) continue;
            if (memcmp(text + o, pat, (size_t)len) == 0) {
                if (nm < cap) {
                    out_id[nm] = (uint32_t)i;
                    out_off[nm] = o;
                }
                nm++;
            }
        }
    }
    return nm;
}

/* saveFilterToFile (ffbf.cpp:508-527): one uint32 delta per set bit. */
uint64_t orc_filter_save_deltas(const uint32_t *F, uint32_t log2_bits, uint32_t *out, uint64_t cap)
{
    const uint64_t words = filter_words(log2_bits);
    const uint64_t bits = log2_bits >= 5 ? words * 32ull : (1ull << log2_bits);
    uint32_t prev = 0;
    uint64_t nout = 0;
    for (uint64_t i = 0; i < bits; i++) {
        if ((F[i >> 5] >> (i & 31u)) & 1u) {
            if (nout < cap) out[nout] = (uint32_t)i - prev;
            prev = (uint32_t)i;
            nout++;
        }
    }
    return nout;
}

/* loadFilterFromFile (ffbf.cpp:487-506): running sum of the deltas. */
void orc_filter_load_deltas(const uint32_t *deltas, uint64_t n, uint32_t log2_bits, uint32_t *F)
{
    memset(F, 0, filter_words(log2_bits) * sizeof(uint32_t));
    uint32_t prev = 0;
    for (uint64_t i = 0; i < n; i++) {
        prev += deltas[i];
        F[prev >> 5] |= 1u << (prev & 31u);
    }
}

uint64_t orc_filter_save_deltas_bits(const uint32_t *F, uint64_t n_bits, uint32_t *out, uint64_t cap)
{
    uint32_t prev = 0;
    uint64_t nout = 0;
    for (uint64_t i = 0; i < n_bits; i++) {
        if ((F[i >> 5] >> (i & 31u)) & 1u) {
            if (nout < cap) out[nout] = (uint32_t)i - prev;
            prev = (uint32_t)i;
            nout++;
        }
    }
    return nout;
}

/* ---------------- split method (src/rabin-karp/ffbf.cpp) ---------------- */

static inline uint64_t rotl64(uint64_t x, uint32_t r)
{
    r &= 63u;
    return r ? (x << r) | (x >> (64u - r)) : x;
}

uint64_t orc_split_bits(const orc_split_params *p)
{
    return (1ull << p->log2_cache_bits) + (1ull << p->log2_ext_bits);
}

/* The rolling hash state of hash_rot_sbox_pre_2 (hash_rot_sbox_pre_2.h:8-51):
 * h, a ring of the last W rotated sbox values and the count since reset. */
typedef struct split_hash {
    uint64_t h;
    uint64_t ring[64];
    uint32_t ptr;
    int full;
} split_hash;

/* reset (hash_rot_sbox_pre_2.h:29-36) */
static inline void split_reset(split_hash *st)
{
    st->h = 0;
    st->ptr = 0;
    st->full = 0;
    memset(st->ring, 0, sizeof st->ring);
}

/* updateHashes (ffbf.cpp:124-133): bytes >= 0x80 become 0, then
 * hash_rot_sbox_pre_2::update (:41-50): rotate by one, add the new byte's
 * sbox value, remove the value W bytes back (kept pre-rotated by W). */
static inline void split_update(const orc_split_params *p, split_hash *st, uint8_t c)
{
    if (c & 0x80u) c = 0;
    const uint64_t sc = p->sbox[c];
    st->h = rotl64(st->h, 1) ^ sc ^ st->ring[st->ptr];
    st->ring[st->ptr] = rotl64(sc, p->window);
    if (++st->ptr >= p->window) {
        st->ptr = 0;
        st->full = 1;
    }
}

void orc_split_hash_window(const orc_split_params *p, const uint8_t *win, uint32_t *h1, uint32_t *h2)
{
    split_hash st;
    split_reset(&st);
    for (uint32_t j = 0; j < p->window; j++) split_update(p, &st, win[j]);
    *h1 = (uint32_t)st.h;          /* hval1 (:26) */
    *h2 = (uint32_t)(st.h >> 32);  /* hval2 (:27) */
}

static inline uint32_t rot16(uint32_t x) { return (x << 16) | (x >> 16); }

static inline uint32_t split_place(const orc_split_params *p, uint32_t i, uint32_t x)
{
    if (i < p->s) return x & (uint32_t)((1ull << p->log2_cache_bits) - 1u);
    return (x & (uint32_t)((1ull << p->log2_ext_bits) - 1u)) + (uint32_t)(1ull << p->log2_cache_bits);
}

/* getHashValue (ffbf.cpp:135-159) */
void orc_split_fwd_indices(const orc_split_params *p, uint32_t h1, uint32_t h2, uint32_t *out)
{
    uint32_t f[11];
    f[0] = h1;
    f[1] = h2;
    f[2] = rot16(h1);
    f[3] = rot16(h2);
    f[4] = h1 + h2;
    f[5] = h1 + (h2 << 1);
    f[6] = h1 + (h2 << 2);
    f[7] = h1 + 3u * h2;
    f[8] = h1 + 5u * h2;
    f[9] = h1 + 6u * h2;
    f[10] = h1 + 7u * h2;
    for (uint32_t i = 0; i < p->k; i++) out[i] = split_place(p, i, f[i]);
}

/* getHashValue_2 (ffbf.cpp:161-185) */
void orc_split_rev_indices(const orc_split_params *p, uint32_t h1, uint32_t h2, uint32_t *out)
{
    uint32_t r[11];
    r[0] = h1 + (h2 << 1);
    r[1] = h1 + (h2 << 2);
    r[2] = h1 + 3u * h2;
    r[3] = h1 + 5u * h2;
    r[4] = h1 + 6u * h2;
    r[5] = h1 + 7u * h2;
    r[6] = h1;
    r[7] = h2;
    r[8] = rot16(h1);
    r[9] = rot16(h2);
    r[10] = h1 + h2;
    for (uint32_t i = 0; i < p->k; i++) out[i] = split_place(p, i, r[i]);
}

static inline uint64_t split_words(const orc_split_params *p) { return orc_split_bits(p) / 32u; }

/* Filter::operator<< (ffbf.cpp:268-283) per pattern, over a cleared F. */
uint64_t orc_split_build(const orc_split_params *p, const uint8_t *bytes, const uint64_t *off,
                         uint64_t n_patterns, uint32_t *F)
{
    memset(F, 0, split_words(p) * sizeof(uint32_t));
    uint64_t hashed = 0;
    for (uint64_t i = 0; i < n_patterns; i++) {
        if (off[i + 1] - off[i] < p->window) continue;
        uint32_t h1, h2, idx[11];
        orc_split_hash_window(p, bytes + off[i], &h1, &h2);
        orc_split_fwd_indices(p, h1, h2, idx);
        for (uint32_t j = 0; j < p->k; j++) set_bit(F, idx[j]);
        hashed++;
    }
    return hashed;
}

/* Filter::processFile's loop (ffbf.cpp:340-370) over the window starts
 * [lo, hi): restart at '\n', hash every other byte, test F once the hash is
 * full (checkInFilter, :214-239) and on a hit set the reverse bits in R
 * (setBitsInFilter_2, :205-212).  Starting from a reset at lo is exact: a
 * window's hash depends only on its own W bytes. */
static uint64_t split_scan_range(const orc_split_params *p, const uint8_t *text, uint64_t n,
                                 uint64_t lo, uint64_t hi, const uint32_t *F, uint32_t *R,
                                 uint64_t *cand, uint64_t cap)
{
    const uint32_t W = p->window;
    uint64_t found = 0;
    if (n < W || lo >= hi) return 0;
    const uint64_t end = hi + W - 1 < n ? hi + W - 1 : n;
    split_hash st;
    split_reset(&st);
    for (uint64_t i = lo; i < end; i++) {
        if (text[i] == '\n') {
            split_reset(&st);
            continue;
        }
        split_update(p, &st, text[i]);
        if (!st.full) continue;
        const uint32_t h1 = (uint32_t)st.h, h2 = (uint32_t)(st.h >> 32);
        uint32_t idx[11];
        orc_split_fwd_indices(p, h1, h2, idx);
        int hit = 1;
        for (uint32_t j = 0; j < p->k; j++) {
            if (!test_bit(F, idx[j])) { hit = 0; break; }
        }
        if (!hit) continue;
        orc_split_rev_indices(p, h1, h2, idx);
        for (uint32_t j = 0; j < p->k; j++) set_bit(R, idx[j]);
        if (found < cap) cand[found] = i + 1 - W;
        found++;
    }
    return found;
}

uint64_t orc_split_scan(const orc_split_params *p, const uint8_t *text, uint64_t n,
                        const uint32_t *F, uint32_t *R, uint64_t *cand, uint64_t cap)
{
    memset(R, 0, split_words(p) * sizeof(uint32_t));
    if (n < p->window) return 0;
    return split_scan_range(p, text, n, 0, n - p->window + 1, F, R, cand, cap);
}

typedef struct split_job {
    const orc_split_params *p;
    const uint8_t *text;
    uint64_t n, lo, hi;
    const uint32_t *F;
    uint32_t *R;
    uint64_t *cand;
    uint64_t cap, found;
} split_job;

static void *split_worker(void *arg)
{
    split_job *j = (split_job *)arg;
    j->found = split_scan_range(j->p, j->text, j->n, j->lo, j->hi, j->F, j->R, j->cand, j->cap);
    return NULL;
}

uint64_t orc_split_scan_mt(const orc_split_params *p, const uint8_t *text, uint64_t n,
                           const uint32_t *F, uint32_t *R, uint64_t *cand, uint64_t cap, int nthreads)
{
    const uint64_t words = split_words(p);
    memset(R, 0, words * sizeof(uint32_t));
    if (n < p->window) return 0;
    const uint64_t npos = n - p->window + 1;
    if (nthreads < 1) nthreads = 1;
    if ((uint64_t)nthreads > npos) nthreads = (int)npos;
    split_job *jobs = (split_job *)calloc((size_t)nthreads, sizeof(split_job));
    pthread_t *tid = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    const uint64_t per = (npos + (uint64_t)nthreads - 1) / (uint64_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        split_job *j = &jobs[t];
        j->p = p; j->text = text; j->n = n; j->F = F;
        j->lo = (uint64_t)t * per;
        j->hi = j->lo + per < npos ? j->lo + per : npos;
        j->R = t == 0 ? R : (uint32_t *)calloc(words, sizeof(uint32_t));
        j->cap = j->hi > j->lo ? j->hi - j->lo : 0;
        if (j->cap > cap) j->cap = cap;
        j->cand = (uint64_t *)malloc((j->cap ? j->cap : 1) * sizeof(uint64_t));
        pthread_create(&tid[t], NULL, split_worker, j);
    }
    uint64_t total = 0;
    for (int t = 0; t < nthreads; t++) {
        split_job *j = &jobs[t];
        pthread_join(tid[t], NULL);
        if (t > 0) {
            for (uint64_t w = 0; w < words; w++) R[w] |= j->R[w];
            free(j->R);
        }
        uint64_t kept = j->found < j->cap ? j->found : j->cap;
        for (uint64_t i = 0; i < kept && total + i < cap; i++) cand[total + i] = j->cand[i];
        total += j->found;
        free(j->cand);
    }
    free(jobs);
    free(tid);
    return total;
}

/* Filter::filterPhrases (ffbf.cpp:394-485): the first full window of each
 * pattern tested against R (checkInFilter_2, :241-266); ids are positions. */
uint64_t orc_split_pattern_pass(const orc_split_params *p, const uint8_t *bytes, const uint64_t *off,
                                uint64_t n_patterns, const uint32_t *R, uint32_t *surv_bitmap)
{
    memset(surv_bitmap, 0, ((n_patterns + 31) / 32) * sizeof(uint32_t));
    uint64_t survivors = 0;
    for (uint64_t i = 0; i < n_patterns; i++) {
        if (off[i + 1] - off[i] < p->window) continue;
        uint32_t h1, h2, idx[11];
        orc_split_hash_window(p, bytes + off[i], &h1, &h2);
        orc_split_rev_indices(p, h1, h2, idx);
        int ok = 1;
        for (uint32_t j = 0; j < p->k; j++) {
            if (!test_bit(R, idx[j])) { ok = 0; break; }
        }
        if (ok) {
            surv_bitmap[i >> 5] |= 1u << (i & 31u);
            survivors++;
        }
    }
    return survivors;
}

static void split_hash_cb(const void *p, const uint8_t *win, uint32_t *h1, uint32_t *h2)
{
    orc_split_hash_window((const orc_split_params *)p, win, h1, h2);
}

uint64_t orc_split_verify(const orc_split_params *p, const uint8_t *text, uint64_t n,
                          const uint8_t *bytes, const uint64_t *off, uint64_t n_patterns,
                          const uint64_t *cand, uint64_t n_cand, const uint32_t *surv_bitmap,
                          uint32_t *out_id, uint64_t *out_off, uint64_t cap)
{
    return verify_core(p, split_hash_cb, p->window, text, n, bytes, off, n_patterns, cand, n_cand, surv_bitmap,
                       out_id, out_off, cap);
}

/* Filter::operator[] (ffbf.cpp:285-308) over many lines: a line shorter
 * than W is false; otherwise the hash runs over the line from a reset, byte
 * by byte (the line is not restarted inside: operator[] takes one line), and
 * the line is true as soon as one window has its k forward bits set in F
 * (checkInFilter, :214-239).  Nothing is fed forward.  Line i is
 * bytes[off[i] .. off[i+1]); hits[i] gets 0 or 1.  Returns the number of
 * true lines.  The pair method: the GPU arithmetic's window hash and forward
 * family on a 2^log2_bits-bit F. */
uint64_t orc_query_lines(const orc_params *p, const uint8_t *bytes, const uint64_t *off, uint64_t n_lines,
                         const uint32_t *F, uint8_t *hits)
{
    const uint32_t W = p->window;
    uint64_t n_hit = 0;
    for (uint64_t i = 0; i < n_lines; i++) {
        const uint64_t len = off[i + 1] - off[i];
        const uint8_t *line = bytes + off[i];
        int hit = 0;
        for (uint64_t o = 0; len >= W && o + W <= len && !hit; o++) {
            uint32_t h1, h2, idx[5];
            orc_hash_window(p, line + o, &h1, &h2);
            orc_fwd_indices(p, h1, h2, idx);
            hit = 1;
            for (uint32_t j = 0; j < p->k && hit; j++) hit = test_bit(F, idx[j]);
        }
        hits[i] = (uint8_t)hit;
        n_hit += (uint64_t)hit;
    }
    return n_hit;
}

/* The same for the CPU binary's split filter: its ring-buffer update from a
 * reset (bytes >= 0x80 hashed as 0), tested once the hash is full. */
uint64_t orc_split_query_lines(const orc_split_params *p, const uint8_t *bytes, const uint64_t *off,
                               uint64_t n_lines, const uint32_t *F, uint8_t *hits)
{
    uint64_t n_hit = 0;
    for (uint64_t i = 0; i < n_lines; i++) {
        const uint64_t len = off[i + 1] - off[i];
        const uint8_t *line = bytes + off[i];
        int hit = 0;
        if (len >= p->window) {
            split_hash st;
            split_reset(&st);
            for (uint64_t c = 0; c < len && !hit; c++) {
                split_update(p, &st, line[c]);
                if (!st.full) continue;
                uint32_t idx[11];
                orc_split_fwd_indices(p, (uint32_t)st.h, (uint32_t)(st.h >> 32), idx);
                hit = 1;
                for (uint32_t j = 0; j < p->k && hit; j++) hit = test_bit(F, idx[j]);
            }
        }
        hits[i] = (uint8_t)hit;
        n_hit += (uint64_t)hit;
    }
    return n_hit;
}
