// ffbf_noalloc.hip -- ffbf_run_device without any device allocation.
//
// The same stages as ffbf_run_device (scan with feed-forward, candidate
// compaction, pattern pass, verification, device sort) and the same
// kernels, but every buffer the call itself needs -- scan counters,
// per-wave candidate staging and its counts and offsets, the candidate list,
// the match list, the sort keys and the radix sort's temporary storage --
// lives in one file-scope device array of fixed size.  Each GPU holds its
// own copy of the array (allocated when the code object loads, kept until
// the process ends); its address is looked up once per device with
// hipGetSymbolAddress.  The context's resident state (F, its prefilter, R,
// the pattern set and its survivor bitmap and index) was allocated when the
// context was created, given patterns and built; the caller passes the text
// and the output buffer.  Nothing here calls hipMalloc, hipFree or a library
// routine that allocates.
#include <mutex>

#include "ffbf_ctx.h"

namespace ffbf {

namespace {

constexpr uint64_t kScratchBytes = 32ull << 20;     // per GPU
constexpr uint64_t kNoallocCandidates = 1ull << 18;  // candidate windows per call
constexpr uint64_t kNoallocMatches = 1ull << 18;     // verified matches per call
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Where each buffer sits in one device's copy of the array.
struct Layout {
    bool ready = false;
    unsigned char *base = nullptr;  // the device address of `scratch` on this device
    int n_cu = 0;
    unsigned long long *counters = nullptr;
    uint32_t *counts = nullptr;  // per wave region
    uint64_t *offsets = nullptr;
    Cand *raw = nullptr;  // regions, then spill
    uint64_t raw_cap = 0;
    uint32_t n_regions = 0, region = 0;
    Cand *cand = nullptr;
    MatchRec *matches = nullptr;
    uint64_t *keys = nullptr;  // 2 x kNoallocMatches
    void *temp = nullptr;
    size_t temp_bytes = 0;
};

Layout g_layout[kMaxDevices];
std::mutex g_mu[kMaxDevices];  // one call at a time per device: calls share the device's array

// Looks the array up on `device` (current) and lays the buffers out in it.
int layout_for(int device, Layout &L)
{
    if (L.ready) return FFBF_OK;
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    int n_cu = 0;
    HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device));
    Layout l;
    l.base = (unsigned char *)p;
    l.n_cu = n_cu;
    l.n_regions = (uint32_t)n_cu * kRegionWavesPerCU;
    l.region = 16;
    size_t tb = 0;
    HIP_TRY(sort_temp_bytes(kNoallocMatches, 64, &tb));
    uint64_t at = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = at;
        at = (at + bytes + 255u) & ~255ull;  // every buffer 256-byte aligned
        return l.base + o;
    };
    for (;; l.region /= 2) {
        at = 0;
        l.counters = (unsigned long long *)take(C_COUNT * sizeof(unsigned long long));
        l.counts = (uint32_t *)take((uint64_t)l.n_regions * 4u);
        l.offsets = (uint64_t *)take(((uint64_t)l.n_regions + 2u) * 8u);
        l.raw_cap = (uint64_t)l.n_regions * l.region + kNoallocCandidates / 2u;
        l.raw = (Cand *)take(l.raw_cap * sizeof(Cand));
        l.cand = (Cand *)take(kNoallocCandidates * sizeof(Cand));
        l.matches = (MatchRec *)take(kNoallocMatches * sizeof(MatchRec));
        l.keys = (uint64_t *)take(2u * kNoallocMatches * 8u);
        l.temp = take(tb);
        l.temp_bytes = tb;
        if (at <= kScratchBytes) break;
        if (l.region <= 1)
            return fail(FFBF_ENOMEM, "allocation-free scratch: " + std::to_string(at) + " bytes needed on a device of " +
                                         std::to_string(n_cu) + " CUs, " + std::to_string(kScratchBytes) + " built in");
    }
    l.ready = true;
    L = l;
    return FFBF_OK;
}

int run_noalloc(ffbf_ctx *c, const Layout &L, const uint8_t *dev_text, uint64_t n, ffbf_match *dev_out, uint64_t cap,
                uint64_t *n_matches)
{
    hipStream_t s = c->stream;
    c->scan_state_clean = false;
    // the array keeps the previous call's contents: clear the counters and
    // the region counts; R and the survivor index are the context's own
    HIP_TRY(launch_fill32((uint32_t *)L.counters, C_COUNT * 2u, 0u, s));
    HIP_TRY(launch_fill32(L.counts, L.n_regions, 0u, s));
    HIP_TRY(launch_fill32(c->R.p, c->f_words, 0u, s));
    HIP_TRY(launch_fill32(c->table.p, 2ull * c->table_size, 0u, s));

    // scan with feed-forward, staging into the array's regions
    ScanArgs a{};
    filter_scan_args(c, a);
    a.text = dev_text;
    a.n = n;
    a.off_base = 0;
    a.cand = L.raw;
    a.cand_cap = L.raw_cap;
    a.wave_counts = L.counts;
    a.region = L.region;
    a.wave0 = 0;
    a.region_end = L.n_regions;
    a.spill_base = (uint64_t)L.n_regions * L.region;
    a.counters = L.counters;
    RET_IF(record(c, EV_S0));
    RET_IF(record(c, EV_S1));
    HIP_TRY(launch_scan(a, c->mode, scan_block(c), c->n_cu, s));
    RET_IF(record(c, EV_S2));
    HIP_TRY(launch_cand_compact(L.raw, L.raw_cap, L.counts, L.offsets, L.n_regions, L.region, L.cand,
                                kNoallocCandidates, L.counters, c->n_cu, s));
    RET_IF(record(c, EV_S3));

    // pattern pass into the context's survivor bitmap and index
    PassArgs pa{};
    pa.pat_hash = c->pat_hash.p;
    pa.off = c->off.p;
    pa.n = c->n_pat;
    pa.W = c->params.window;
    pa.k = c->params.k;
    pa.mask = c->mask;
    pa.R = c->R.p;
    pa.bitmap = c->bitmap.p;
    pa.counters = L.counters;
    pa.table = c->table.p;
    pa.slot_chain = c->table.p + c->table_size;
    pa.chain_next = c->chain_next.p;
    pa.table_mask = c->table_size - 1u;
    pa.method = c->split ? METHOD_SPLIT : METHOD_PAIR;
    pa.split = c->smap;
    RET_IF(record(c, EV_P0));
    HIP_TRY(launch_pattern_pass(pa, c->n_cu, s));
    RET_IF(record(c, EV_P1));

    // verification into the array's match list
    VerifyArgs va{};
    va.text = dev_text;
    va.n = n;
    va.cand = L.cand;
    va.cand_cap = kNoallocCandidates;
    va.bytes = c->bytes.p;
    va.off = c->off.p;
    va.pat_hash = c->pat_hash.p;
    va.table = c->table.p;
    va.slot_chain = c->table.p + c->table_size;
    va.chain_next = c->chain_next.p;
    va.table_mask = c->table_size - 1u;
    va.out = L.matches;
    va.out_cap = kNoallocMatches;
    va.counters = L.counters;
    RET_IF(record(c, EV_V0));
    HIP_TRY(launch_verify(va, c->n_cu, s));
    RET_IF(record(c, EV_V1));

    HIP_TRY(hipMemcpyAsync(c->h_counters, L.counters, C_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const uint64_t n_cand = c->h_counters[C_CAND], n_match = c->h_counters[C_MATCH];
    if (n_cand > kNoallocCandidates)
        return fail(FFBF_ECAPACITY, "allocation-free run: " + std::to_string(n_cand) +
                                        " candidate windows, more than the " + std::to_string(kNoallocCandidates) +
                                        " its scratch holds (use ffbf_run_device)");
    if (n_match > kNoallocMatches)
        return fail(FFBF_ECAPACITY, "allocation-free run: " + std::to_string(n_match) + " matches, more than the " +
                                        std::to_string(kNoallocMatches) + " its scratch holds (use ffbf_run_device)");
    c->n_cand = n_cand;
    c->n_surv = c->h_counters[C_SURV];
    c->n_match = n_match;
    if (n_matches) *n_matches = n_match;
    if (n_match > cap) return fail(FFBF_ECAPACITY, "match buffer too small");
    if (n_match == 0) return FFBF_OK;
    const int end_bit = sort_end_bit(c->n_pat);
    size_t need = 0;
    HIP_TRY(sort_temp_bytes(n_match, end_bit, &need));
    if (need > L.temp_bytes)
        return fail(FFBF_ECAPACITY, "allocation-free run: the sort needs " + std::to_string(need) +
                                        " bytes of temporary storage, " + std::to_string(L.temp_bytes) + " laid out");
    HIP_TRY(sort_matches(L.matches, n_match, L.keys, L.keys + n_match, L.temp, L.temp_bytes, end_bit, s));
    HIP_TRY(hipMemcpyAsync(dev_out, L.matches, n_match * sizeof(MatchRec), hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return FFBF_OK;
}

}  // namespace

}  // namespace ffbf

using namespace ffbf;

extern "C" int ffbf_run_device_noalloc(ffbf_ctx *c, const uint8_t *dev_text, uint64_t n, ffbf_match *dev_matches,
                                       uint64_t cap, uint64_t *n_matches)
{
    RET_IF(check_ctx(c));
    if (n_matches) *n_matches = 0;
    if (!c->built) return fail(FFBF_ESTATE, "scan before build (or filter load)");
    if (!dev_text && n) return fail(FFBF_EINVAL, "null text");
    if (!dev_matches && cap) return fail(FFBF_EINVAL, "null match buffer");
    if (c->device < 0 || c->device >= kMaxDevices) return fail(FFBF_EINVAL, "device index beyond the scratch table");
    if (!sort_packable(c->n_pat, n))
        return fail(FFBF_EINVAL, "allocation-free run: at most 2^24 patterns and 2^40 text bytes");
    std::lock_guard<std::mutex> lk(g_mu[c->device]);
    Layout &L = g_layout[c->device];
    RET_IF(layout_for(c->device, L));
    // the context's own candidate and match lists do not hold this call's
    // results: its getters of those need a scan through the other entry points
    c->scanned = c->passed = c->verified = false;
    c->sorted_valid = false;
    c->text = nullptr;
    c->text_n = 0;
    return run_noalloc(c, L, dev_text, n, dev_matches, cap, n_matches);
}
