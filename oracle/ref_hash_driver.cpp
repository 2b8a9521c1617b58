// A C ABI over the reference's own rolling hashes, compiled from the
// reference's unmodified headers (src/hashes/hash_rot_sbox_2.h and
// src/hashes/hash_rot_sbox_pre_2.h with hash_common.h and hash_buf.h) by
// oracle/ref.mk into oracle/_ref/ -- test infrastructure that pins the
// oracle's window hashes (orc_hash_window, orc_split_hash_window) against
// the reference's code rather than a restatement of it.  Never linked into
// the library.
//
// Each function feeds text[0..n) byte by byte through the reference's
// update() from a reset state and records the hash after every update that
// completes a window: entry o is the hash of window text[o .. o+W).
// ref_pair_hashes: hash_rot_sbox_2<W> (hval1, hval2; the sbox table of the
// header).  ref_split_hashes: hash_rot_sbox_pre_2<W> (low and high halves of
// its 64-bit hash; the 64-bit table of the header); the CPU binary maps bytes
// >= 0x80 to 0 before hashing (ffbf.cpp:124-133), which is left to the
// caller here.  W is 1..64; returns 0, or -1 for another W.
#include <stdint.h>

#include <utility>

#include "hash_rot_sbox_2.h"
#include "hash_rot_sbox_pre_2.h"

namespace {

template <int LEN>
void pair_roll(const uint8_t *text, uint64_t n, uint32_t *h1, uint32_t *h2)
{
    hash_rot_sbox_2<LEN> h;
    h.init();
    for (uint64_t i = 0; i < n; i++) {
        h.update(text[i]);
        if (i + 1 >= (uint64_t)LEN) {
            h1[i + 1 - LEN] = h.hval1();
            h2[i + 1 - LEN] = h.hval2();
        }
    }
}

template <int LEN>
void split_roll(const uint8_t *text, uint64_t n, uint32_t *h1, uint32_t *h2)
{
    hash_rot_sbox_pre_2<LEN> h;
    h.init();
    for (uint64_t i = 0; i < n; i++) {
        h.update(text[i]);
        if (i + 1 >= (uint64_t)LEN) {
            h1[i + 1 - LEN] = h.hval1();
            h2[i + 1 - LEN] = h.hval2();
        }
    }
}

typedef void (*roll_fn)(const uint8_t *, uint64_t, uint32_t *, uint32_t *);

template <size_t... I>
const roll_fn *pair_table(std::index_sequence<I...>)
{
    static const roll_fn t[] = {&pair_roll<(int)I + 1>...};
    return t;
}

template <size_t... I>
const roll_fn *split_table(std::index_sequence<I...>)
{
    static const roll_fn t[] = {&split_roll<(int)I + 1>...};
    return t;
}

}  // namespace

extern "C" int ref_pair_hashes(uint32_t W, const uint8_t *text, uint64_t n, uint32_t *h1, uint32_t *h2)
{
    if (W < 1 || W > 64) return -1;
    pair_table(std::make_index_sequence<64>())[W - 1](text, n, h1, h2);
    return 0;
}

extern "C" int ref_split_hashes(uint32_t W, const uint8_t *text, uint64_t n, uint32_t *h1, uint32_t *h2)
{
    if (W < 1 || W > 64) return -1;
    split_table(std::make_index_sequence<64>())[W - 1](text, n, h1, h2);
    return 0;
}
