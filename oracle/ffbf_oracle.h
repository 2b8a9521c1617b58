/*
 * ffbf_oracle.h -- CPU restatement of the reference feed-forward Bloom filter
 * (FFBF) matcher, used ONLY as test infrastructure: the checker that the HIP
 * path is compared against, and the CPU baseline timed beside it.  Nothing in
 * the shipped library links, loads or calls this code.
 *
 * Reference anchors (efficient/ffbf):
 *   window hash pair        src/gpgpu/greptest.cu:106-114, 157-166, 264-272
 *   rolling form            src/hashes/hash_rot_sbox_2.h:34-51
 *   forward index family    src/gpgpu/greptest.cu:115-120, 199-202
 *   reverse index family    src/gpgpu/greptest.cu:231-234, 273-276
 *   bit layout              src/gpgpu/greptest.cu:80-89
 *   build / scan / pattern pass flow   src/rabin-karp/ffbf.cpp:268-283, 310-392, 394-485
 *
 * Filters are arrays of uint32 words; bit b lives in word b>>5 under mask
 * 1u<<(b&31).  Offsets are 64-bit.  All arithmetic is mod 2^32.
 */
#ifndef FFBF_ORACLE_H
#define FFBF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    uint32_t window;     /* W: bytes per hashed window (>= 1) */
    uint32_t k;          /* hash functions per filter, 1..5 */
    uint32_t log2_bits;  /* filter holds 2^log2_bits bits, 5..32 */
    uint32_t reserved;
    uint32_t sbox[256];
} orc_params;

/* Loop-form hash pair of one W-byte window (greptest.cu:106-114). */
void orc_hash_window(const orc_params *p, const uint8_t *win,
                     uint32_t *h1, uint32_t *h2);

/* Rolling-form hash pairs of every window of text[0..n) (hash_rot_sbox_2.h:34-51).
 * Writes n-W+1 pairs (nothing if n < W). */
void orc_hash_rolling(const orc_params *p, const uint8_t *text, uint64_t n,
                      uint32_t *h1, uint32_t *h2);

/* Forward / reverse bit indices (already masked to the filter size). */
void orc_fwd_indices(const orc_params *p, uint32_t h1, uint32_t h2, uint32_t *out);
void orc_rev_indices(const orc_params *p, uint32_t h1, uint32_t h2, uint32_t *out);

/* Filter build: clears F (2^log2_bits bits) then sets k forward bits for the
 * first W bytes of every pattern with length >= W.  Patterns are
 * bytes[off[i] .. off[i+1]).  Returns the number of patterns hashed. */
uint64_t orc_build(const orc_params *p, const uint8_t *bytes, const uint64_t *off,
                   uint64_t n_patterns, uint32_t *F);

/* Scan with feed-forward: clears R, then for every offset o in [0, n-W] whose
 * k forward bits are all set in F, sets its k reverse bits in R and records o.
 * Candidates are written ascending, at most cap of them; returns the total
 * number of candidates. */
uint64_t orc_scan(const orc_params *p, const uint8_t *text, uint64_t n,
                  const uint32_t *F, uint32_t *R, uint64_t *cand, uint64_t cap);

/* Same, split over nthreads POSIX threads (text cut with a W-1 byte overlap,
 * private R per thread OR-merged).  Candidates ascending. */
uint64_t orc_scan_mt(const orc_params *p, const uint8_t *text, uint64_t n,
                     const uint32_t *F, uint32_t *R, uint64_t *cand, uint64_t cap,
                     int nthreads);

/* Pattern pass: clears the survivor bitmap (ceil(n/32) words), then marks
 * each pattern with length >= W whose k reverse bits are set in R.  Returns
 * the number of survivors. */
uint64_t orc_pattern_pass(const orc_params *p, const uint8_t *bytes, const uint64_t *off,
                          uint64_t n_patterns, const uint32_t *R, uint32_t *surv_bitmap);

/* Exact verification: every (pattern id, offset) with the pattern a survivor,
 * the offset a candidate, the pattern's W-byte prefix equal to the window at
 * the offset and the whole pattern equal to text[offset .. offset+len).
 * Output sorted by (id, offset); at most cap written; returns the total. */
uint64_t orc_verify(const orc_params *p, const uint8_t *text, uint64_t n,
                    const uint8_t *bytes, const uint64_t *off, uint64_t n_patterns,
                    const uint64_t *cand, uint64_t n_cand, const uint32_t *surv_bitmap,
                    uint32_t *out_id, uint64_t *out_off, uint64_t cap);

/* Independent naive exact matcher (no hashing): every occurrence of every
 * pattern with length >= min_len.  Sorted by (id, offset). */
uint64_t orc_brute_matches(const uint8_t *text, uint64_t n,
                           const uint8_t *bytes, const uint64_t *off, uint64_t n_patterns,
                           uint32_t min_len, uint32_t *out_id, uint64_t *out_off, uint64_t cap);

/* Filter persistence in the reference's cache format (ffbf.cpp:487-527):
 * the ascending indices of set bits, each stored as the uint32 difference to
 * the previous set index (the first relative to 0).  Returns the number of
 * uint32 values written (or that would be written if cap is too small). */
uint64_t orc_filter_save_deltas(const uint32_t *F, uint32_t log2_bits, uint32_t *out, uint64_t cap);
void orc_filter_load_deltas(const uint32_t *deltas, uint64_t n, uint32_t log2_bits, uint32_t *F);

/* ---- Split method: the CPU binary's filter (src/rabin-karp/ffbf.cpp) ----
 *
 * A 64-bit rolling hash over the window (hash_rot_sbox_pre_2.h:34-51,
 * 64-bit sbox :55-120): h = rol64(h,1) ^ S[c] ^ rol64(S[c_out], W), whose low
 * and high halves are h1 and h2.  Bytes >= 0x80 hash as 0 (ffbf.cpp:124-133).
 * The filter has a cache part of 2^c bits followed by an external part of
 * 2^e bits; hash function i < s addresses the cache part (x & (2^c-1)), the
 * others the external part ((x & (2^e-1)) + 2^c) (ffbf.cpp:135-185).  A text
 * is hashed line by line: the hash restarts after every '\n', so a window
 * containing '\n' is never tested (ffbf.cpp:340-370).  Rotation amounts are
 * taken mod 64 (the reference's rol64 shifts by 64 for W = 64, which C leaves
 * undefined, hash_common.h:13-16). */
typedef struct orc_split_params {
    uint32_t window;          /* W (LEN; DEF_PATT_LEN 19), 1..64 */
    uint32_t k;               /* hash functions per filter, 1..11 (DEF_K 5) */
    uint32_t s;               /* functions in the cache part, 0..k (DEF_S 2) */
    uint32_t log2_cache_bits; /* cache part 2^c bits, 5..31 (DEF_CACHE_SIZE 2^24) */
    uint32_t log2_ext_bits;   /* external part 2^e bits, 5..31 (DEF_EXT_SIZE 2^28) */
    uint32_t reserved;
    uint64_t sbox[256];
} orc_split_params;

/* Bits of a split filter: 2^c + 2^e. */
uint64_t orc_split_bits(const orc_split_params *p);

/* Hash of one W-byte window by the reference's ring-buffer update from a
 * reset state (hash_rot_sbox_pre_2.h:32-51, after ffbf.cpp:124-133). */
void orc_split_hash_window(const orc_split_params *p, const uint8_t *win, uint32_t *h1, uint32_t *h2);

/* Forward (getHashValue, ffbf.cpp:135-159) and reverse (getHashValue_2,
 * ffbf.cpp:161-185) bit indices, placed in the cache or external part. */
void orc_split_fwd_indices(const orc_split_params *p, uint32_t h1, uint32_t h2, uint32_t *out);
void orc_split_rev_indices(const orc_split_params *p, uint32_t h1, uint32_t h2, uint32_t *out);

/* Filter build (Filter::operator<<, ffbf.cpp:268-283): clears F, then sets
 * k bits for the first W bytes of every pattern with length >= W. */
uint64_t orc_split_build(const orc_split_params *p, const uint8_t *bytes, const uint64_t *off,
                         uint64_t n_patterns, uint32_t *F);

/* Scan with feed-forward (Filter::processFile, ffbf.cpp:310-392): clears R;
 * every window o in [0, n-W] without a '\n' whose k forward bits are set in
 * F is a candidate, sets its k reverse bits in R and is recorded (ascending,
 * at most cap).  Returns the number of candidates. */
uint64_t orc_split_scan(const orc_split_params *p, const uint8_t *text, uint64_t n,
                        const uint32_t *F, uint32_t *R, uint64_t *cand, uint64_t cap);
/* The same over nthreads POSIX threads (ranges of window starts). */
uint64_t orc_split_scan_mt(const orc_split_params *p, const uint8_t *text, uint64_t n,
                           const uint32_t *F, uint32_t *R, uint64_t *cand, uint64_t cap, int nthreads);

/* Pattern pass (Filter::filterPhrases, ffbf.cpp:394-485): each pattern with
 * length >= W whose k reverse bits of its prefix window are set in R. */
uint64_t orc_split_pattern_pass(const orc_split_params *p, const uint8_t *bytes, const uint64_t *off,
                                uint64_t n_patterns, const uint32_t *R, uint32_t *surv_bitmap);

/* Exact verification as orc_verify, with the split method's window hash. */
uint64_t orc_split_verify(const orc_split_params *p, const uint8_t *text, uint64_t n,
                          const uint8_t *bytes, const uint64_t *off, uint64_t n_patterns,
                          const uint64_t *cand, uint64_t n_cand, const uint32_t *surv_bitmap,
                          uint32_t *out_id, uint64_t *out_off, uint64_t cap);

/* Filter persistence over a filter of n_bits bits (the split filter's
 * 2^c + 2^e bits; ffbf.cpp:487-527). */
uint64_t orc_filter_save_deltas_bits(const uint32_t *F, uint64_t n_bits, uint32_t *out, uint64_t cap);

/* Filter::operator[] (ffbf.cpp:285-308) for many lines: hits[i] = 1 iff
 * line i (bytes[off[i] .. off[i+1])) is at least W bytes long and one of its
 * windows, hashed from a reset over the line without restarting inside it,
 * has all k forward bits set in F.  Returns the number of such lines. */
uint64_t orc_query_lines(const orc_params *p, const uint8_t *bytes, const uint64_t *off, uint64_t n_lines,
                         const uint32_t *F, uint8_t *hits);
uint64_t orc_split_query_lines(const orc_split_params *p, const uint8_t *bytes, const uint64_t *off,
                               uint64_t n_lines, const uint32_t *F, uint8_t *hits);

#ifdef __cplusplus
}
#endif
#endif
