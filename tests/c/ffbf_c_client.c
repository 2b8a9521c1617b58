/* A plain C99 caller of the C ABI, the way a reference-side driver binds it:
 * create a context, load newline-separated patterns, build the filter, match
 * one text, print "match <pattern id> <offset>" per match in list order.
 * Without a usable device it prints the error status and message instead;
 * it then exits 0 unless "gpu" was given as the first argument. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ffbf.h"

static int check(int st, const char *what)
{
    if (st != FFBF_OK) printf("%s status %d: %s\n", what, st, ffbf_last_error());
    return st;
}

int main(int argc, char **argv)
{
    const int need_gpu = argc > 1 && strcmp(argv[1], "gpu") == 0;
    printf("version %s\n", ffbf_version());

    ffbf_params p;
    memset(&p, 0, sizeof p);
    p.window = 8;
    p.k = 4;
    p.log2_bits = 16;
    for (uint32_t i = 0; i < 256; i++) p.sbox[i] = (i * 2654435761u) ^ 0x9E3779B9u;

    ffbf_ctx *ctx = NULL;
    if (check(ffbf_create(&p, 0, &ctx), "create") != FFBF_OK) return need_gpu ? 1 : 0;

    /* "abc" is shorter than the window: it keeps id 2 and never matches */
    static const char patterns[] = "needle in\nhaystack\nabc\nrolling hash\n";
    static const char text[] = "a haystack with a needle in it, and a rolling hash over a haystack";
    ffbf_match *m = NULL;
    uint64_t n = 0;
    int rc = 1;
    if (check(ffbf_set_patterns_lines(ctx, patterns, sizeof patterns - 1), "set_patterns_lines") == FFBF_OK &&
        check(ffbf_build(ctx), "build") == FFBF_OK &&
        check(ffbf_match_text(ctx, (const uint8_t *)text, sizeof text - 1, &m, &n), "match_text") == FFBF_OK) {
        for (uint64_t i = 0; i < n; i++)
            printf("match %u %llu\n", m[i].pattern_id, (unsigned long long)m[i].offset);
        rc = 0;
    }
    ffbf_free_matches(m);
    ffbf_destroy(ctx);
    return rc;
}
