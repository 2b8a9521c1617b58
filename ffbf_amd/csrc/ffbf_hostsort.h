// ffbf_hostsort.h -- host-side ordering of match records (host only, no HIP).
//
// Match lists come back from the device in the order verification appended
// them; the API returns them sorted by (pattern id, offset), or (text,
// pattern id, offset) for a batch.  A stable LSD radix sort does that in a
// few linear passes: one pass per 11-bit digit of each field that is not the
// same in every record, least significant field first.  Short lists of a
// short text take few passes; there is no packing limit on the fields.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

namespace ffbf {

// Stable sort of v[0..n) by key(v[i]) (an unsigned 64-bit value), using tmp
// as the second buffer.
template <class Rec, class Key>
void radix_pass_sort(Rec *v, size_t n, Key key, std::vector<Rec> &tmp)
{
    if (n < 2) return;
    uint64_t any = 0, all = ~0ull;
    for (size_t i = 0; i < n; i++) {
        const uint64_t k = key(v[i]);
        any |= k;
        all &= k;
    }
    const uint64_t varying = any ^ all;  // bits that differ between some records
    if (!varying) return;
    constexpr int kDigit = 11;
    constexpr uint64_t kMask = (1ull << kDigit) - 1u;
    tmp.resize(n);
    Rec *a = v, *b = tmp.data();
    for (int sh = 0; sh < 64; sh += kDigit) {
        if (((varying >> sh) & kMask) == 0) continue;
        uint32_t at[(1u << kDigit) + 1u] = {};
        for (size_t i = 0; i < n; i++) at[((key(a[i]) >> sh) & kMask) + 1u]++;
        for (uint32_t d = 0; d < (1u << kDigit); d++) at[d + 1u] += at[d];
        for (size_t i = 0; i < n; i++) b[at[(key(a[i]) >> sh) & kMask]++] = a[i];
        Rec *t = a;
        a = b;
        b = t;
    }
    if (a != v) std::memcpy(v, a, n * sizeof(Rec));
}

// Records with fields `major` (32-bit) and `offset` (64-bit): sorted by
// (major, offset).
template <class Rec, class Major>
void sort_by_major_offset(Rec *v, size_t n, Major major, std::vector<Rec> &tmp)
{
    radix_pass_sort(v, n, [](const Rec &r) { return (uint64_t)r.offset; }, tmp);
    radix_pass_sort(v, n, [&](const Rec &r) { return (uint64_t)major(r); }, tmp);
}

}  // namespace ffbf
