/*
 * ffbf_workload.h -- seeded synthetic inputs (texts, pattern sets, plants,
 * low-complexity repeats, sbox tables) for the FFBF scan path.
 *
 * These entry points have no counterpart in the reference, which reads its
 * patterns and corpus from files (src/gpgpu/greptest.cu:598-665,
 * src/rabin-karp/ffbf.cpp:529-601).  They exist so that tests, benchmarks and
 * the CPU baseline see byte-identical inputs on any machine: every value is a
 * pure function of (seed, index) built on counter-based SplitMix64, with no
 * libm and no sequential generator state.  All functions are host-only.
 */
#ifndef FFBF_WORKLOAD_H
#define FFBF_WORKLOAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFBF_ALPHA_ASCII 0u     /* 95 printable bytes 0x20..0x7E, uniform */
#define FFBF_ALPHA_DNA_GC40 1u  /* A,T at 30% each, C,G at 20% each */

/* Bytes [start, start+n) of the text stream of `seed` over `alphabet`.
 * nthreads <= 0 uses every hardware thread.  Returns 0 or -1. */
int ffbf_gen_bytes(uint32_t alphabet, uint64_t seed, uint64_t start, uint64_t n, uint8_t *out, int nthreads);

/* Offsets of n patterns with lengths uniform in [lmin, lmax]: off[0]=0,
 * off[i+1]-off[i] = length of pattern i.  off has n+1 entries. */
int ffbf_gen_pattern_offsets(uint64_t seed, uint64_t n, uint32_t lmin, uint32_t lmax, uint64_t *off);

/* The `total` concatenated pattern bytes of `seed` over `alphabet`. */
int ffbf_gen_pattern_bytes(uint32_t alphabet, uint64_t seed, uint64_t total, uint8_t *bytes, int nthreads);

/* Copies n_plants seeded patterns into the text, in order (a later plant may
 * overwrite an earlier one): plant j takes pattern U[0, n_patterns) at offset
 * U[0, n - len].  plant_id / plant_off (may be NULL) receive the choices;
 * a pattern longer than the text is not planted (offset UINT64_MAX). */
int ffbf_gen_plants(uint64_t seed, uint64_t n_plants, uint8_t *text, uint64_t n,
                    const uint8_t *bytes, const uint64_t *off, uint64_t n_patterns,
                    uint32_t *plant_id, uint64_t *plant_off);

/* Overwrites one tandem-repeat run per `block` bytes of DNA text: run length
 * U[min_len, max_len], period U[min_period, max_period] (<= 64), unit drawn as
 * GC-40 DNA, run start uniform inside the block. */
int ffbf_gen_tandem_repeats(uint64_t seed, uint8_t *text, uint64_t n, uint32_t block,
                            uint32_t min_len, uint32_t max_len, uint32_t min_period, uint32_t max_period,
                            int nthreads);

/* 256 seeded random uint32 substitution values. */
int ffbf_gen_sbox(uint64_t seed, uint32_t *sbox);

/* n log-normal sizes: median * exp(sigma * Z), Z the Irwin-Hall
 * approximation of a standard normal, rounded and clipped to [lo, hi].
 * Uses no libm, so every machine produces the same sizes. */
int ffbf_gen_lognormal_sizes(uint64_t seed, uint64_t n, double median, double sigma, uint64_t lo, uint64_t hi,
                             uint64_t *out);

/* Raw draw i of the stream (seed, tag): used for sizes and routing. */
uint64_t ffbf_gen_u64(uint64_t seed, uint64_t tag, uint64_t i);

#ifdef __cplusplus
}
#endif
#endif
