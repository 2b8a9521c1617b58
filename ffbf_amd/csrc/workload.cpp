// workload.cpp -- seeded synthetic inputs for the FFBF scan path.
//
// Every byte, length, plant and table entry is a pure function of
// (seed, stream tag, index): a counter-based SplitMix64 draw (Steele, Lea and
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014).
// No libm and no sequential state, so the host, any number of threads and
// any machine produce identical inputs.
#include "ffbf_workload.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Key of one named stream of a seed.
inline uint64_t stream_key(uint64_t seed, uint64_t tag) { return mix64(seed ^ mix64(tag * kGolden + 1)); }

// i-th draw of a stream: SplitMix64 output for state key + (i+1)*golden.
inline uint64_t draw(uint64_t key, uint64_t i) { return mix64(key + (i + 1) * kGolden); }

// Unbiased-enough bounded integer in [0, bound) from 64 random bits
// (multiply-high; bias below 2^-32 for bound < 2^32).
inline uint64_t bounded(uint64_t r, uint64_t bound)
{
    return (uint64_t)(((unsigned __int128)r * bound) >> 64);
}

enum : uint64_t {
    TAG_TEXT = 1, TAG_PAT_LEN = 2, TAG_PAT_BYTES = 3, TAG_PLANT_ID = 4, TAG_PLANT_OFF = 5,
    TAG_SBOX = 6, TAG_REPEAT = 7, TAG_SIZES = 8, TAG_ROUTE = 9,
};

// DNA at GC 40%: u = (r16 * 10) >> 16 in [0,10): 0-2 A, 3-5 T, 6-7 C, 8-9 G.
inline uint8_t dna_of(uint32_t r16)
{
    static const uint8_t lut[10] = {'A', 'A', 'A', 'T', 'T', 'T', 'C', 'C', 'G', 'G'};
    return lut[(r16 * 10u) >> 16];
}

// Printable ASCII 0x20..0x7E uniform: 0x20 + (r32 * 95) >> 32.
inline uint8_t ascii_of(uint32_t r32) { return (uint8_t)(0x20u + (uint32_t)(((uint64_t)r32 * 95u) >> 32)); }

void fill_range(uint32_t alphabet, uint64_t key, uint64_t start, uint64_t n, uint8_t *out)
{
    uint64_t i = 0;
    while (i < n) {
        const uint64_t g = start + i;
        if (alphabet == FFBF_ALPHA_DNA_GC40) {
            // four bases per draw
            const uint64_t r = draw(key, g >> 2);
            for (uint64_t j = g & 3; j < 4 && i < n; j++, i++) out[i] = dna_of((uint32_t)(r >> (16 * j)) & 0xFFFFu);
        } else {
            // two printable bytes per draw
            const uint64_t r = draw(key, g >> 1);
            for (uint64_t j = g & 1; j < 2 && i < n; j++, i++) out[i] = ascii_of((uint32_t)(r >> (32 * j)));
        }
    }
}

int clamp_threads(int nthreads, uint64_t n)
{
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    if (nthreads > 64) nthreads = 64;
    const uint64_t min_per = 1u << 20;
    if ((uint64_t)nthreads > n / min_per + 1) nthreads = (int)(n / min_per + 1);
    return nthreads;
}

template <class F>
void parallel_for(uint64_t n, int nthreads, F fn)
{
    nthreads = clamp_threads(nthreads, n);
    if (nthreads <= 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> ts;
    const uint64_t per = (n + (uint64_t)nthreads - 1) / (uint64_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        const uint64_t lo = std::min(n, (uint64_t)t * per), hi = std::min(n, lo + per);
        ts.emplace_back([=] { fn(lo, hi); });
    }
    for (auto &t : ts) t.join();
}

// exp(x) from IEEE-exact pieces only (floor, ldexp, +, *, /): range reduction
// by ln 2, then the Taylor series of exp(r) for |r| <= ln2/2 to degree 16.
// Deterministic on every machine, unlike a libm or SIMD exp.
double det_exp(double x)
{
    const double ln2 = 0.69314718055994530942;
    const double kf = std::floor(x / ln2 + 0.5);
    const double r = x - kf * ln2;
    double term = 1.0, sum = 1.0;
    for (int i = 1; i <= 16; i++) {
        term = term * r / (double)i;
        sum += term;
    }
    return std::ldexp(sum, (int)kf);
}

// Standard normal approximated by the Irwin-Hall sum of 12 uniforms minus 6
// (uniforms from 53-bit draws; exact arithmetic).
double approx_normal(uint64_t key, uint64_t i)
{
    double s = 0.0;
    for (uint64_t j = 0; j < 12; j++) s += (double)(draw(key, 12 * i + j) >> 11) * (1.0 / 9007199254740992.0);
    return s - 6.0;
}

}  // namespace

extern "C" {

int ffbf_gen_lognormal_sizes(uint64_t seed, uint64_t n, double median, double sigma, uint64_t lo, uint64_t hi,
                             uint64_t *out)
{
    if (!out || lo > hi || median <= 0.0 || sigma < 0.0) return -1;
    const uint64_t key = stream_key(seed, TAG_SIZES);
    for (uint64_t i = 0; i < n; i++) {
        const double v = median * det_exp(sigma * approx_normal(key, i));
        uint64_t x = v >= 1.8e19 ? UINT64_MAX : (uint64_t)(v + 0.5);
        out[i] = std::min(hi, std::max(lo, x));
    }
    return 0;
}


int ffbf_gen_bytes(uint32_t alphabet, uint64_t seed, uint64_t start, uint64_t n, uint8_t *out, int nthreads)
{
    if (alphabet > FFBF_ALPHA_DNA_GC40 || (!out && n)) return -1;
    const uint64_t key = stream_key(seed, TAG_TEXT);
    parallel_for(n, nthreads, [&](uint64_t lo, uint64_t hi) { fill_range(alphabet, key, start + lo, hi - lo, out + lo); });
    return 0;
}

int ffbf_gen_pattern_offsets(uint64_t seed, uint64_t n, uint32_t lmin, uint32_t lmax, uint64_t *off)
{
    if (!off || lmin > lmax) return -1;
    const uint64_t key = stream_key(seed, TAG_PAT_LEN);
    const uint64_t span = (uint64_t)lmax - lmin + 1;
    off[0] = 0;
    for (uint64_t i = 0; i < n; i++) off[i + 1] = off[i] + lmin + bounded(draw(key, i), span);
    return 0;
}

int ffbf_gen_pattern_bytes(uint32_t alphabet, uint64_t seed, uint64_t total, uint8_t *bytes, int nthreads)
{
    if (alphabet > FFBF_ALPHA_DNA_GC40 || (!bytes && total)) return -1;
    const uint64_t key = stream_key(seed, TAG_PAT_BYTES);
    parallel_for(total, nthreads, [&](uint64_t lo, uint64_t hi) { fill_range(alphabet, key, lo, hi - lo, bytes + lo); });
    return 0;
}

int ffbf_gen_plants(uint64_t seed, uint64_t n_plants, uint8_t *text, uint64_t n,
                    const uint8_t *bytes, const uint64_t *off, uint64_t n_patterns,
                    uint32_t *plant_id, uint64_t *plant_off)
{
    if (!text || !bytes || !off || n_patterns == 0) return n_plants ? -1 : 0;
    const uint64_t kid = stream_key(seed, TAG_PLANT_ID), koff = stream_key(seed, TAG_PLANT_OFF);
    for (uint64_t j = 0; j < n_plants; j++) {
        const uint64_t id = bounded(draw(kid, j), n_patterns);
        const uint64_t len = off[id + 1] - off[id];
        uint64_t o = UINT64_MAX;
        if (len <= n) {
            o = bounded(draw(koff, j), n - len + 1);
            std::memcpy(text + o, bytes + off[id], (size_t)len);
        }
        if (plant_id) plant_id[j] = (uint32_t)id;
        if (plant_off) plant_off[j] = o;
    }
    return 0;
}

int ffbf_gen_tandem_repeats(uint64_t seed, uint8_t *text, uint64_t n, uint32_t block,
                            uint32_t min_len, uint32_t max_len, uint32_t min_period, uint32_t max_period,
                            int nthreads)
{
    if (!text || block == 0 || min_len > max_len || max_len > block || min_period == 0 || min_period > max_period ||
        max_period > 64)
        return -1;
    const uint64_t key = stream_key(seed, TAG_REPEAT);
    const uint64_t nblocks = n / block;
    parallel_for(nblocks, nthreads, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t b = lo; b < hi; b++) {
            const uint64_t r0 = draw(key, 4 * b), r1 = draw(key, 4 * b + 1), r2 = draw(key, 4 * b + 2);
            const uint32_t len = min_len + (uint32_t)bounded(r0, max_len - min_len + 1);
            const uint32_t period = min_period + (uint32_t)bounded(r1, max_period - min_period + 1);
            const uint64_t start = b * (uint64_t)block + bounded(r2, block - len + 1);
            uint8_t unit[64];
            const uint64_t r3 = draw(key, 4 * b + 3);
            for (uint32_t i = 0; i < period; i++) unit[i] = dna_of((uint32_t)(mix64(r3 + i) & 0xFFFFu));
            for (uint32_t i = 0; i < len; i++) text[start + i] = unit[i % period];
        }
    });
    return 0;
}

int ffbf_gen_sbox(uint64_t seed, uint32_t *sbox)
{
    if (!sbox) return -1;
    const uint64_t key = stream_key(seed, TAG_SBOX);
    for (uint32_t i = 0; i < 256; i++) sbox[i] = (uint32_t)(draw(key, i) >> 32);
    return 0;
}

uint64_t ffbf_gen_u64(uint64_t seed, uint64_t tag, uint64_t i) { return draw(stream_key(seed, tag), i); }

}  // extern "C"
