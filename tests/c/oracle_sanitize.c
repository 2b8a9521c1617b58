/* Exercises every function of the C oracle (oracle/ffbf_oracle.c) on small
 * seeded cases, so that a build under AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_sanitize.py) sees each code path:
 * W from 1 to 64, k 1..5 (split: 1..11, s 0..k), empty and short texts and
 * patterns, output caps below the totals, threaded scans, filter persistence.
 * Along the way it checks the properties that do not need a GPU: the rolling
 * hash equals the loop hash, the threaded scan equals the single-threaded
 * one, verification equals the naive matcher, a saved filter loads back.
 * Exits 0 when every check passes. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ffbf_oracle.h"

static int failures = 0;
#define CHECK(c)                                                            \
    do {                                                                    \
        if (!(c)) {                                                         \
            fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            failures++;                                                     \
        }                                                                   \
    } while (0)

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t rnd(void)
{
    /* splitmix64 */
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* patterns over `alpha`, some shorter than W; some copied into the text */
static void make_case(uint32_t W, const char *alpha, int high_and_nl, uint64_t n_text, uint64_t n_pat, uint8_t **text,
                      uint8_t **bytes, uint64_t **off)
{
    const size_t na = strlen(alpha);
    *text = malloc(n_text ? n_text : 1);
    for (uint64_t i = 0; i < n_text; i++) {
        uint8_t c = (uint8_t)alpha[rnd() % na];
        if (high_and_nl && rnd() % 37 == 0) c = '\n';
        if (high_and_nl && rnd() % 41 == 0) c = (uint8_t)(0x80 + rnd() % 128);
        (*text)[i] = c;
    }
    *off = malloc((n_pat + 1) * sizeof(uint64_t));
    uint64_t total = 0;
    (*off)[0] = 0;
    uint64_t *len = malloc((n_pat ? n_pat : 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < n_pat; i++) {
        len[i] = rnd() % 4 == 0 ? rnd() % (W + 1) : W + rnd() % 9; /* a quarter at most W bytes */
        total += len[i];
        (*off)[i + 1] = total;
    }
    *bytes = malloc(total ? total : 1);
    for (uint64_t i = 0; i < n_pat; i++) {
        uint8_t *p = *bytes + (*off)[i];
        if (n_text >= len[i] && rnd() % 2) {
            memcpy(p, *text + rnd() % (n_text - len[i] + 1), len[i]);
            for (uint64_t j = 0; j < len[i]; j++)
                if (p[j] == '\n') p[j] = (uint8_t)alpha[0]; /* patterns are lines */
        } else {
            for (uint64_t j = 0; j < len[i]; j++) p[j] = (uint8_t)alpha[rnd() % na];
        }
    }
    free(len);
}

static int same_matches(const uint32_t *a_id, const uint64_t *a_off, uint64_t na, const uint32_t *b_id,
                        const uint64_t *b_off, uint64_t nb)
{
    if (na != nb) return 0;
    for (uint64_t i = 0; i < na; i++)
        if (a_id[i] != b_id[i] || a_off[i] != b_off[i]) return 0;
    return 1;
}

static void pair_case(uint32_t W, uint32_t k, uint32_t log2, uint64_t n_text, uint64_t n_pat, const char *alpha)
{
    orc_params p;
    memset(&p, 0, sizeof p);
    p.window = W;
    p.k = k;
    p.log2_bits = log2;
    for (int i = 0; i < 256; i++) p.sbox[i] = (uint32_t)rnd();
    uint8_t *text, *bytes;
    uint64_t *off;
    make_case(W, alpha, 0, n_text, n_pat, &text, &bytes, &off);
    const uint64_t words = (1ull << log2) / 32 ? (1ull << log2) / 32 : 1;
    uint32_t *F = calloc(words, 4), *R = calloc(words, 4), *R2 = calloc(words, 4), *F2 = calloc(words, 4);
    /* hashes: rolling equals loop on every window */
    const uint64_t nw = n_text >= W ? n_text - W + 1 : 0;
    uint32_t *r1 = calloc(nw ? nw : 1, 4), *r2 = calloc(nw ? nw : 1, 4);
    orc_hash_rolling(&p, text, n_text, r1, r2);
    for (uint64_t o = 0; o < nw; o++) {
        uint32_t a, b;
        orc_hash_window(&p, text + o, &a, &b);
        CHECK(a == r1[o] && b == r2[o]);
    }
    uint32_t idx[5];
    orc_fwd_indices(&p, r1[0], r2[0], idx);
    orc_rev_indices(&p, r1[0], r2[0], idx);
    orc_build(&p, bytes, off, n_pat, F);
    uint64_t *cand = malloc((nw + 1) * 8), *cand2 = malloc((nw + 1) * 8);
    const uint64_t nc = orc_scan(&p, text, n_text, F, R, cand, nw + 1);
    const uint64_t nc_short = orc_scan(&p, text, n_text, F, R2, cand2, nc / 2); /* cap below the total */
    CHECK(nc_short == nc);
    const uint64_t nc2 = orc_scan_mt(&p, text, n_text, F, R2, cand2, nw + 1, 3);
    CHECK(nc2 == nc && memcmp(cand, cand2, nc * 8) == 0 && memcmp(R, R2, words * 4) == 0);
    uint32_t *surv = calloc((n_pat + 31) / 32 + 1, 4);
    orc_pattern_pass(&p, bytes, off, n_pat, R, surv);
    /* totals first (nothing written at cap 0), then the lists */
    const uint64_t nm = orc_verify(&p, text, n_text, bytes, off, n_pat, cand, nc, surv, NULL, NULL, 0);
    const uint64_t nb = orc_brute_matches(text, n_text, bytes, off, n_pat, W, NULL, NULL, 0);
    const uint64_t cap = (nm > nb ? nm : nb) + 1;
    uint32_t *id = malloc(cap * 4), *bid = malloc(cap * 4);
    uint64_t *mo = malloc(cap * 8), *bo = malloc(cap * 8);
    CHECK(orc_verify(&p, text, n_text, bytes, off, n_pat, cand, nc, surv, id, mo, cap) == nm);
    CHECK(orc_brute_matches(text, n_text, bytes, off, n_pat, W, bid, bo, cap) == nb);
    CHECK(same_matches(id, mo, nm, bid, bo, nb));
    CHECK(orc_verify(&p, text, n_text, bytes, off, n_pat, cand, nc, surv, id, mo, nm / 2) == nm);
    /* persistence */
    const uint64_t nd = orc_filter_save_deltas(F, log2, NULL, 0);
    uint32_t *d = malloc((nd ? nd : 1) * 4);
    CHECK(orc_filter_save_deltas(F, log2, d, nd) == nd);
    orc_filter_load_deltas(d, nd, log2, F2);
    CHECK(memcmp(F, F2, words * 4) == 0);
    free(d);
    free(text), free(bytes), free(off), free(F), free(R), free(R2), free(F2), free(r1), free(r2);
    free(cand), free(cand2), free(surv), free(id), free(bid), free(mo), free(bo);
}

static void split_case(uint32_t W, uint32_t k, uint32_t s, uint32_t c, uint32_t e, uint64_t n_text, uint64_t n_pat)
{
    orc_split_params p;
    memset(&p, 0, sizeof p);
    p.window = W;
    p.k = k;
    p.s = s;
    p.log2_cache_bits = c;
    p.log2_ext_bits = e;
    for (int i = 0; i < 256; i++) p.sbox[i] = rnd();
    uint8_t *text, *bytes;
    uint64_t *off;
    make_case(W, "ACGTacgt ", 1, n_text, n_pat, &text, &bytes, &off);
    const uint64_t bits = orc_split_bits(&p), words = bits / 32;
    uint32_t *F = calloc(words, 4), *R = calloc(words, 4), *R2 = calloc(words, 4);
    uint32_t h1 = 0, h2 = 0, idx[11];
    if (n_text >= W) orc_split_hash_window(&p, text, &h1, &h2);
    orc_split_fwd_indices(&p, h1, h2, idx);
    for (uint32_t j = 0; j < k; j++) CHECK(idx[j] < bits);
    orc_split_rev_indices(&p, h1, h2, idx);
    for (uint32_t j = 0; j < k; j++) CHECK(idx[j] < bits);
    orc_split_build(&p, bytes, off, n_pat, F);
    const uint64_t nw = n_text >= W ? n_text - W + 1 : 0;
    uint64_t *cand = malloc((nw + 1) * 8), *cand2 = malloc((nw + 1) * 8);
    const uint64_t nc = orc_split_scan(&p, text, n_text, F, R, cand, nw + 1);
    const uint64_t nc2 = orc_split_scan_mt(&p, text, n_text, F, R2, cand2, nw + 1, 4);
    CHECK(nc2 == nc && memcmp(cand, cand2, nc * 8) == 0 && memcmp(R, R2, words * 4) == 0);
    for (uint64_t i = 0; i < nc; i++) CHECK(memchr(text + cand[i], '\n', W) == NULL);
    uint32_t *surv = calloc((n_pat + 31) / 32 + 1, 4);
    orc_split_pattern_pass(&p, bytes, off, n_pat, R, surv);
    const uint64_t nm = orc_split_verify(&p, text, n_text, bytes, off, n_pat, cand, nc, surv, NULL, NULL, 0);
    const uint64_t nb = orc_brute_matches(text, n_text, bytes, off, n_pat, W, NULL, NULL, 0);
    const uint64_t cap = (nm > nb ? nm : nb) + 1;
    uint32_t *id = malloc(cap * 4), *bid = malloc(cap * 4);
    uint64_t *mo = malloc(cap * 8), *bo = malloc(cap * 8);
    CHECK(orc_split_verify(&p, text, n_text, bytes, off, n_pat, cand, nc, surv, id, mo, cap) == nm);
    /* patterns hold no newline, so every occurrence's window is tested;
     * high bytes hash as 0 but verification compares raw bytes */
    CHECK(orc_brute_matches(text, n_text, bytes, off, n_pat, W, bid, bo, cap) == nb);
    CHECK(same_matches(id, mo, nm, bid, bo, nb));
    const uint64_t nd = orc_filter_save_deltas_bits(F, bits, NULL, 0);
    uint32_t *d = malloc((nd ? nd : 1) * 4);
    CHECK(orc_filter_save_deltas_bits(F, bits, d, nd) == nd);
    free(d);
    free(text), free(bytes), free(off), free(F), free(R), free(R2), free(cand), free(cand2), free(surv);
    free(id), free(bid), free(mo), free(bo);
}

int main(void)
{
    for (uint32_t W = 1; W <= 64; W++) {
        const uint32_t k = 1 + W % 5;
        pair_case(W, k, 5 + W % 14, W * 40 + (W % 7), 20 + W % 13, W % 2 ? "ACGT" : "abcdefghij");
        split_case(W, 1 + W % 11, (W % 11) % (1 + W % 11 + 1), 5 + W % 9, 6 + W % 12, W * 30 + 5, 15 + W % 7);
    }
    /* empty and degenerate inputs */
    pair_case(19, 4, 16, 0, 0, "ab");
    pair_case(19, 4, 16, 18, 5, "ab");
    pair_case(19, 4, 16, 19, 5, "ab");
    pair_case(8, 1, 5, 3000, 200, "a"); /* one letter: every window equal, saturated tiny filter */
    split_case(19, 5, 2, 10, 12, 0, 0);
    split_case(19, 5, 0, 8, 10, 18, 4);
    split_case(19, 11, 11, 6, 6, 2000, 60);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("oracle_sanitize: all checks passed\n");
    return 0;
}
