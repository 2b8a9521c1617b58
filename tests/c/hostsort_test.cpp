// The host's match-record radix sort (ffbf_amd/csrc/ffbf_hostsort.h) against
// std::stable_sort with the API's ordering: records {a, b, offset} sorted by
// (b, a, offset), as a batch's (text, pattern id, offset).  Empty and
// one-record lists, all-equal fields, full-width 32- and 64-bit values, and
// the ranges real match lists have.  Prints "all checks passed" and exits 0.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "ffbf_hostsort.h"

struct Rec {
    uint32_t a, b;
    uint64_t offset;
};

int main()
{
    std::mt19937_64 g(7);
    int bad = 0;
    for (int t = 0; t < 2000; t++) {
        const size_t n = (t % 7 == 0) ? (size_t)(g() % 3) : (size_t)(g() % 4000);
        const int mode = t % 5;
        std::vector<Rec> v(n);
        for (Rec &r : v) {
            r.a = mode == 0 ? 5u : (mode == 1 ? (uint32_t)g() : (uint32_t)(g() % 50));
            r.b = (uint32_t)(g() % 4);
            r.offset = mode == 2 ? 123u : (mode == 3 ? g() : g() % 100000);
        }
        std::vector<Rec> want = v, tmp;
        ffbf::sort_by_major_offset(v.data(), v.size(), [](const Rec &r) { return r.a; }, tmp);
        ffbf::radix_pass_sort(v.data(), v.size(), [](const Rec &r) { return r.b; }, tmp);
        std::stable_sort(want.begin(), want.end(), [](const Rec &x, const Rec &y) {
            if (x.b != y.b) return x.b < y.b;
            return x.a != y.a ? x.a < y.a : x.offset < y.offset;
        });
        for (size_t i = 0; i < n; i++)
            if (v[i].a != want[i].a || v[i].b != want[i].b || v[i].offset != want[i].offset) {
                std::printf("case %d (n %zu, mode %d): record %zu differs\n", t, n, mode, i);
                bad++;
                break;
            }
    }
    if (bad) return 1;
    std::printf("all checks passed\n");
    return 0;
}
