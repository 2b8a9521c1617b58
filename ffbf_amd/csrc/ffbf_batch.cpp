// ffbf_batch.cpp -- batched scans: many texts, each routed to one of several
// resident filters (contexts), scanned in one persistent launch.
//
// One batch object binds a fixed set of built contexts (same device, window,
// k and sbox).  A run clears every filter's R and survivor index, runs the
// batched scan over all chunks, the pattern pass of every filter and the
// batched verification, and copies the counters back.  That sequence does not
// depend on the texts (chunk count and descriptors live in device memory), so
// it is captured once in a hipGraph and replayed for every batch until a
// buffer has to grow.  R and the survivors are per filter and per batch: a
// filter's R is the union over the texts routed to it, which only widens the
// survivor set; verification is per text, so every text's match list is the
// same as when it is scanned alone.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include <cmath>

#include "ffbf_ctx.h"
#include "ffbf_hostsort.h"

using namespace ffbf;

struct ffbf_batch {
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    std::vector<ffbf_ctx *> filters;
    uint32_t W = 0, k = 0;
    uint32_t pre_words_max = 0;  // largest LDS prefilter among the filters (dynamic LDS of the scan)
    bool has_lds = false, has_global = false;  // filters of each chunk class

    DevBuf<ScanArgs> d_scan;
    DevBuf<VerifyArgs> d_ver;
    DevBuf<PassArgs> d_pass;          // per filter, for the one-launch pattern pass
    DevBuf<uint64_t> d_wave_start;    // first 64-id wave of each filter's pass (nf+1)
    uint64_t pass_waves = 0;
    DevBuf<uint32_t> d_active;        // per filter: some text of the batch is routed to it
    DevBuf<ZeroSpan> d_zero;          // everything a run clears first, in one launch
    uint32_t n_zero = 0;
    DevBuf<BatchText> d_texts;
    DevBuf<BatchChunk> d_chunks;
    // per chunk class (0: LDS prefilter, 1: prefilter in global memory), four
    // words each: next chunk, number of chunks, index of the first; word 3
    // holds the number of texts (read by verification), word 7 is unused
    DevBuf<unsigned long long> ctl;
    DevBuf<unsigned long long> counters;  // C_CAND, C_MATCH
    unsigned long long *h_counters = nullptr;
    // pinned copy of the match list's first kMatchPrefetch records, made by
    // every run with its counters: a short list reaches the host with the
    // run's own synchronisation
    BatchMatchRec *h_match = nullptr;
    uint64_t prefetched = 0;  // records of the current list in h_match
    DevBuf<Cand> cand;  // compacted candidate list
    CandStage cstage;   // staging of both scan launches before compaction
    DevBuf<BatchMatchRec> matches;
    DevBuf<uint64_t> sort_keys;
    DevBuf<uint8_t> sort_temp;

    const uint8_t *texts = nullptr;
    uint32_t n_texts = 0, n_chunks = 0;
    uint64_t total_bytes = 0, max_text = 0;
    std::vector<BatchText> h_texts;
    bool have_texts = false;

    std::vector<uint64_t> gens;  // the filters' generations the descriptors were built from
    ffbf_batch_tuning tuning{32u, 0u, 1024u, 1u, 0u};
    uint32_t probes_eff = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e_run0 = nullptr, e_run1 = nullptr, e_scan0 = nullptr, e_scan1 = nullptr;
    float run_ms = 0.0f, scan_ms = -1.0f;
    bool timing = true;  // record e_run0/e_run1 around each run (ffbf_batch_set_timing)

    uint64_t n_cand = 0, n_match = 0;
    bool ran = false, sorted_on_device = false, sorted_valid = false;
    std::vector<BatchMatchRec> sorted;
};

namespace ffbf {
hipStream_t batch_stream(ffbf_batch *b) { return b->stream; }
}  // namespace ffbf

namespace {

void drop_graph(ffbf_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Whether the LDS-class scan has a 32-bytes-per-lane instance: the pair
// method at a compile-time W.
bool takes_32_bytes(const ffbf_batch *b)
{
    return !b->filters[0]->split && (b->W == 19 || b->W == 24 || b->W == 32);
}

// 16-byte units per lane of the LDS-class scan: two (32 window starts) where
// there is an instance for it, unless 16 bytes per lane were asked for.
uint32_t units_per_lane(const ffbf_batch *b)
{
    return (takes_32_bytes(b) && b->tuning.bytes_per_lane != 16u) ? 2u : 1u;
}

// Match records every run copies to the pinned h_match (a captured graph
// holds this count with the list's address; both change only together with
// the list, which drops the graph).
uint64_t prefetch_count(const ffbf_batch *b) { return std::min<uint64_t>(b->matches.cap, kMatchPrefetch); }

uint32_t bits_for(uint64_t v)
{
    uint32_t b = 1;
    while (b < 64 && (1ull << b) <= v) b++;
    return b;
}

// Rebuilds the per-filter descriptors from the contexts' current state.
int refresh_filters(ffbf_batch *b)
{
    const uint32_t nf = (uint32_t)b->filters.size();
    std::vector<ScanArgs> sa(nf);
    std::vector<VerifyArgs> va(nf);
    b->pre_words_max = 0;
    const bool had_lds = b->has_lds, had_global = b->has_global;
    b->has_lds = b->has_global = false;
    for (uint32_t f = 0; f < nf; f++) {
        const ffbf_ctx *c = b->filters[f];
        const bool lds = mode_in_lds(c->mode);
        // an R gathered in LDS sits right after the filter
        ScanArgs probe{};
        filter_scan_args(c, probe);
        if (lds) b->pre_words_max = std::max(b->pre_words_max, c->pre_words * (probe.r_lds ? 2u : 1u));
        (lds ? b->has_lds : b->has_global) = true;
    }
    if (b->has_lds != had_lds || b->has_global != had_global) drop_graph(b);
    // probes: requested, else the smallest auto choice among LDS filters
    uint32_t p = b->tuning.probes;
    if (!p) {
        p = 3;
        for (uint32_t f = 0; f < nf; f++) {
            const ffbf_ctx *c = b->filters[f];
            if (mode_in_lds(c->mode)) p = std::min(p, std::max(1u, c->probes));
        }
    }
    b->probes_eff = std::max(1u, std::min({p, 3u, b->k}));
    // a split filter's fold covers only its s cache-part functions: the
    // launch-wide up-front probes may not exceed any such s
    for (uint32_t f = 0; f < nf; f++) {
        const ffbf_ctx *c = b->filters[f];
        if (c->split && c->mode == MODE_FOLD) b->probes_eff = std::max(1u, std::min(b->probes_eff, c->smap.s));
    }
    for (uint32_t f = 0; f < nf; f++) {
        const ffbf_ctx *c = b->filters[f];
        ScanArgs &a = sa[f];
        a = ScanArgs{};
        const bool lds = mode_in_lds(c->mode);
        filter_scan_args(c, a);
        // region slots: LDS-class launch [0, half), global-class launch [half, all)
        const uint32_t half = b->cstage.n_regions / 2u;
        if (lds)
            b->cstage.fill(a, 0u, half);
        else
            b->cstage.fill(a, half, b->cstage.n_regions);
        a.counters = b->counters.p;
        a.P = lds ? b->probes_eff : 1u;
        a.f_first =
            ((c->mode == MODE_FOLD || c->mode == MODE_HYBRID) && !c->split && fold_passes_many(c->pre_fill, a.P)) ? 1u
                                                                                                                : 0u;
        VerifyArgs &v = va[f];
        v = VerifyArgs{};
        v.bytes = c->bytes.p;
        v.off = c->off.p;
        v.pat_hash = c->pat_hash.p;
        v.table = c->table.p;
        v.slot_chain = c->table.p + c->table_size;
        v.chain_next = c->chain_next.p;
        v.table_mask = c->table_size - 1u;
    }
    RET_IF(b->d_scan.ensure(nf, "batch filter descriptors"));
    RET_IF(b->d_ver.ensure(nf, "batch verify descriptors"));
    // pattern passes: descriptors and each filter's first wave
    std::vector<PassArgs> pa(nf);
    std::vector<uint64_t> ws(nf + 1, 0);
    for (uint32_t f = 0; f < nf; f++) {
        const ffbf_ctx *c = b->filters[f];
        PassArgs &p = pa[f];
        p = PassArgs{};
        p.pat_hash = c->pat_hash.p;
        p.off = c->off.p;
        p.n = c->n_pat;
        p.W = c->params.window;
        p.k = c->params.k;
        p.mask = c->mask;
        p.R = c->R.p;
        p.bitmap = c->bitmap.p;
        p.counters = c->counters.p;
        p.table = c->table.p;
        p.slot_chain = c->table.p + c->table_size;
        p.chain_next = c->chain_next.p;
        p.table_mask = c->table_size - 1u;
        p.method = c->split ? METHOD_SPLIT : METHOD_PAIR;
        p.split = c->smap;
        ws[f + 1] = ws[f] + (c->n_pat + 63u) / 64u;
    }
    b->pass_waves = ws[nf];
    // what a run clears before its scans, as spans of at most 8192 words
    std::vector<ZeroSpan> zs;
    auto zero = [&](uint32_t *p, uint64_t words) {
        for (uint64_t o = 0; o < words; o += 8192u)
            zs.push_back(ZeroSpan{p + o, (uint32_t)std::min<uint64_t>(8192u, words - o)});
    };
    zero((uint32_t *)b->ctl.p, 2u);        // next chunk, class 0
    zero((uint32_t *)(b->ctl.p + 4), 2u);  // next chunk, class 1
    zero((uint32_t *)b->counters.p, 2u * C_COUNT);
    zero(b->cstage.counts.p, b->cstage.n_regions);
    for (ffbf_ctx *c : b->filters) {
        zero(c->R.p, c->f_words);
        zero(c->table.p, 2ull * c->table_size);
    }
    b->n_zero = (uint32_t)zs.size();
    RET_IF(b->d_pass.ensure(nf, "batch pass descriptors"));
    RET_IF(b->d_wave_start.ensure(nf + 1, "batch pass waves"));
    RET_IF(b->d_zero.ensure(zs.size(), "batch clear list"));
    RET_IF(b->d_active.ensure(nf, "batch filter flags"));
    HIP_TRY(hipMemcpyAsync(b->d_scan.p, sa.data(), nf * sizeof(ScanArgs), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_ver.p, va.data(), nf * sizeof(VerifyArgs), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_pass.p, pa.data(), nf * sizeof(PassArgs), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_wave_start.p, ws.data(), (nf + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_zero.p, zs.data(), zs.size() * sizeof(ZeroSpan), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return FFBF_OK;
}

// The whole batch run on b->stream (capturable: kernels and one D2H copy).
int enqueue_run(ffbf_batch *b, bool timed)
{
    // chunk counters, batch counters, candidate region counts, and every
    // filter's R, survivor index and survivor count
    HIP_TRY(launch_zero_spans(b->d_zero.p, b->n_zero, b->n_cu, b->stream));
    if (timed) HIP_TRY(hipEventRecord(b->e_scan0, b->stream));
    BatchArgs ba{};
    ba.texts = b->texts;
    ba.tdesc = b->d_texts.p;
    ba.chunks = b->d_chunks.p;
    ba.filters = b->d_scan.p;
    ba.sbox = b->filters[0]->sbox.p;
    ba.W = b->W;
    ba.method = b->filters[0]->split ? METHOD_SPLIT : METHOD_PAIR;
    if (b->has_lds) {
        ba.ctl = b->ctl.p;
        HIP_TRY(launch_scan_batch(ba, b->tuning.block, b->probes_eff, units_per_lane(b), b->pre_words_max, b->n_cu,
                                  b->stream));
    }
    if (b->has_global) {
        ba.ctl = b->ctl.p + 4;
        HIP_TRY(launch_scan_batch_global(ba, b->n_cu, b->stream));
    }
    if (timed) HIP_TRY(hipEventRecord(b->e_scan1, b->stream));  // the scan kernels alone end here
    HIP_TRY(launch_cand_compact(b->cstage.raw.p, b->cstage.raw.cap, b->cstage.counts.p, b->cstage.offsets.p,
                                b->cstage.n_regions, b->cstage.region, b->cand.p, b->cand.cap, b->counters.p, b->n_cu,
                                b->stream));
    HIP_TRY(launch_pattern_pass_multi(b->d_pass.p, b->d_wave_start.p, (uint32_t)b->filters.size(), b->d_active.p,
                                      b->pass_waves, b->n_cu, b->stream, b->filters[0]->split));
    BatchVerifyArgs va{};
    va.texts = b->texts;
    va.tdesc = b->d_texts.p;
    va.cand = b->cand.p;
    va.cand_cap = b->cand.cap;
    va.filters = b->d_ver.p;
    va.out = b->matches.p;
    va.out_cap = b->matches.cap;
    va.counters = b->counters.p;
    va.ctl = b->ctl.p;
    HIP_TRY(launch_verify_batch(va, b->n_cu, b->stream));
    HIP_TRY(hipMemcpyAsync(b->h_counters, b->counters.p, C_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           b->stream));
    HIP_TRY(hipMemcpyAsync(b->h_match, b->matches.p, prefetch_count(b) * sizeof(BatchMatchRec), hipMemcpyDeviceToHost,
                           b->stream));
    return FFBF_OK;
}

int run_once(ffbf_batch *b)
{
    // the run writes its filters' R and survivor indexes
    for (ffbf_ctx *c : b->filters) c->scan_state_clean = false;
    const bool graph = b->tuning.use_graph != 0u;
    if (graph && !b->exec) {
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int st = enqueue_run(b, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (st != FFBF_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        HIP_TRY(e);
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    if (b->timing) HIP_TRY(hipEventRecord(b->e_run0, b->stream));
    if (graph) {
        HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    } else {
        RET_IF(enqueue_run(b, b->timing));
    }
    if (b->timing) HIP_TRY(hipEventRecord(b->e_run1, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->run_ms = 0.0f;
    b->scan_ms = -1.0f;
    if (b->timing) {
        (void)hipEventElapsedTime(&b->run_ms, b->e_run0, b->e_run1);
        if (!graph) (void)hipEventElapsedTime(&b->scan_ms, b->e_scan0, b->e_scan1);
    }
    return FFBF_OK;
}

int sort_on_device(ffbf_batch *b)
{
    b->sorted_on_device = false;
    // a few thousand records sort faster on the host at copy-out than the
    // device radix sort's launches take
    if (b->n_match < kDeviceSortMin) return FFBF_OK;
    uint64_t max_pat = 1;
    for (const ffbf_ctx *c : b->filters) max_pat = std::max<uint64_t>(max_pat, c->n_pat);
    const uint32_t bo = bits_for(b->max_text), bi = bits_for(max_pat), bt = bits_for(b->n_texts);
    if (bo + bi + bt > 64) return FFBF_OK;
    const int end_bit = (int)(bo + bi + bt);
    size_t tb = 0;
    HIP_TRY(sort_temp_bytes(b->n_match, end_bit, &tb));
    // geometric growth: a reallocation (hipFree) stalls the whole device
    if (2 * b->n_match > b->sort_keys.cap)
        RET_IF(b->sort_keys.ensure(std::max<uint64_t>(4 * b->n_match, 1 << 17), "batch sort keys"));
    if (tb > b->sort_temp.cap) RET_IF(b->sort_temp.ensure(std::max<uint64_t>(2 * tb, 1 << 20), "batch sort scratch"));
    tb = b->sort_temp.cap;
    HIP_TRY(sort_batch_matches(b->matches.p, b->n_match, b->sort_keys.p, b->sort_keys.p + b->n_match, b->sort_temp.p, tb,
                               (int)bo, (int)(bo + bi), end_bit, b->stream));
    b->sorted_on_device = true;
    b->prefetched = 0;  // the records were copied out before the sort reordered them
    return FFBF_OK;
}

int check_batch(ffbf_batch *b)
{
    if (!b) return fail(FFBF_EINVAL, "null batch");
    HIP_TRY(hipSetDevice(b->device));
    return FFBF_OK;
}

// Cuts the current texts into chunks again (new chunk size, or filters whose
// prefilter class changed).  Texts are contiguous ranges of one buffer, so
// their offsets follow from their starts and sizes.
int rechunk_texts(ffbf_batch *b)
{
    if (!b->have_texts) return FFBF_OK;
    std::vector<uint64_t> off(b->n_texts + 1, 0);
    std::vector<uint32_t> fil(b->n_texts);
    for (uint32_t i = 0; i < b->n_texts; i++) {
        off[i] = b->h_texts[i].start;
        off[i + 1] = b->h_texts[i].start + b->h_texts[i].n;
        fil[i] = b->h_texts[i].filter;
    }
    return ffbf_batch_set_texts(b, b->texts, off.data(), fil.data(), b->n_texts);
}

// A borrowed context was rebuilt, reloaded or given another scan mode since
// the descriptors were made: rebuild them, and cut the texts again (the
// chunk order depends on each filter's prefilter class).
// A run or new texts that follow need no re-cut (rechunk false).
int sync_filters(ffbf_batch *b, bool rechunk)
{
    bool changed = b->gens.size() != b->filters.size();
    for (size_t f = 0; !changed && f < b->filters.size(); f++) changed = b->gens[f] != b->filters[f]->generation;
    if (!changed) return FFBF_OK;
    for (ffbf_ctx *c : b->filters) {
        if (!c->built || !c->patterns) return fail(FFBF_ESTATE, "a filter of the batch has no built filter");
        HIP_TRY(hipStreamSynchronize(c->stream));  // a rebuild may still be running
    }
    drop_graph(b);
    RET_IF(refresh_filters(b));
    b->gens.resize(b->filters.size());
    for (size_t f = 0; f < b->filters.size(); f++) b->gens[f] = b->filters[f]->generation;
    return rechunk ? rechunk_texts(b) : FFBF_OK;
}

}  // namespace

namespace ffbf {
int batch_reserve(ffbf_batch *b, uint32_t max_texts, uint64_t max_bytes)
{
    RET_IF(check_batch(b));
    const uint32_t ci = b->tuning.chunk_iters ? b->tuning.chunk_iters : 32u;
    // chunks: every text adds at most two partial wave-iterations and one
    // partial chunk beyond its share of the bytes
    const uint64_t chunks = max_bytes / (16ull * kIterUnits * ci) + 3ull * max_texts + 16u;
    // candidates: twice the false-positive windows the densest filter is
    // expected to pass (fill^k per window), at least 1 per KiB and at most 1
    // per 16 bytes; a batch that overflows its list regrows it, which stalls
    // the device and re-captures the graph (filters of 16 bits per pattern
    // with k=3 pass ~0.5% of windows)
    double fp = 0.0;
    for (ffbf_ctx *c : b->filters) fp = std::max(fp, std::pow(f_fill(c), (double)c->params.k));
    const uint64_t fp_want = std::min<uint64_t>((uint64_t)(2.0 * fp * (double)max_bytes), max_bytes >> 4);
    const uint64_t cand = std::max<uint64_t>(std::max<uint64_t>(1 << 16, max_bytes >> 10), fp_want);
    bool moved = false;
    if (b->d_texts.cap < max_texts) {
        RET_IF(b->d_texts.ensure(std::max<uint64_t>(max_texts, 1024), "batch text descriptors"));
        moved = true;
    }
    if (b->d_chunks.cap < chunks) {
        RET_IF(b->d_chunks.ensure(chunks, "batch chunks"));
        moved = true;
    }
    if (b->cand.cap < cand) {
        RET_IF(b->cand.ensure(cand, "batch candidates"));
        RET_IF(b->cstage.ensure(b->n_cu, b->cand.cap));
        RET_IF(refresh_filters(b));
        moved = true;
    }
    if (moved) drop_graph(b);
    return FFBF_OK;
}
}  // namespace ffbf

extern "C" {

int ffbf_batch_create(ffbf_ctx *const *filters, uint32_t n_filters, ffbf_batch **out)
{
    if (!filters || !out || n_filters == 0) return fail(FFBF_EINVAL, "need at least one filter");
    *out = nullptr;
    const ffbf_ctx *f0 = filters[0];
    if (!f0) return fail(FFBF_EINVAL, "null filter");
    for (uint32_t i = 0; i < n_filters; i++) {
        const ffbf_ctx *c = filters[i];
        if (!c) return fail(FFBF_EINVAL, "null filter");
        if (!c->built || !c->patterns) return fail(FFBF_ESTATE, "every filter must have patterns and a built filter");
        if (c->device != f0->device || c->params.window != f0->params.window || c->params.k != f0->params.k ||
            std::memcmp(c->params.sbox, f0->params.sbox, sizeof f0->params.sbox) != 0)
            return fail(FFBF_EINVAL, "filters of a batch need the same device, window, k and sbox");
        // one hash method per batch; split filters also share their 64-bit
        // table (their s and part sizes may differ)
        if (c->split != f0->split ||
            (c->split && std::memcmp(c->split_params.sbox, f0->split_params.sbox, sizeof f0->split_params.sbox) != 0))
            return fail(FFBF_EINVAL, "filters of a batch need the same hash method (and split filters the same sbox)");
    }
    ffbf_batch *b = new (std::nothrow) ffbf_batch();
    if (!b) return fail(FFBF_ENOMEM, "batch allocation");
    b->filters.assign(filters, filters + n_filters);
    b->device = f0->device;
    b->W = f0->params.window;
    b->k = f0->params.k;
    auto init = [&]() -> int {
        HIP_TRY(hipSetDevice(b->device));
        HIP_TRY(hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, b->device));
        HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&b->e_run0));
        HIP_TRY(hipEventCreate(&b->e_run1));
        HIP_TRY(hipEventCreate(&b->e_scan0));
        HIP_TRY(hipEventCreate(&b->e_scan1));
        HIP_TRY(hipHostMalloc((void **)&b->h_counters, C_COUNT * sizeof(unsigned long long), hipHostMallocDefault));
        HIP_TRY(hipHostMalloc((void **)&b->h_match, kMatchPrefetch * sizeof(BatchMatchRec), hipHostMallocDefault));
        RET_IF(b->ctl.ensure(8, "batch control"));
        RET_IF(b->counters.ensure(C_COUNT, "batch counters"));
        RET_IF(b->cand.ensure(1 << 16, "batch candidates"));
        RET_IF(b->cstage.ensure(b->n_cu, b->cand.cap));
        RET_IF(b->matches.ensure(1 << 16, "batch matches"));
        return sync_filters(b, false);  // also waits for the filters' own streams
    };
    const int st = init();
    if (st != FFBF_OK) {
        const std::string msg = ffbf_last_error();
        ffbf_batch_destroy(b);
        fail(st, msg);
        return st;
    }
    *out = b;
    return FFBF_OK;
}

void ffbf_batch_destroy(ffbf_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    b->d_scan.release();
    b->d_ver.release();
    b->d_texts.release();
    b->d_chunks.release();
    b->ctl.release();
    b->counters.release();
    b->cand.release();
    b->cstage.release();
    b->d_pass.release();
    b->d_wave_start.release();
    b->d_active.release();
    b->d_zero.release();
    b->matches.release();
    b->sort_keys.release();
    b->sort_temp.release();
    if (b->h_counters) (void)hipHostFree(b->h_counters);
    if (b->h_match) (void)hipHostFree(b->h_match);
    for (hipEvent_t e : {b->e_run0, b->e_run1, b->e_scan0, b->e_scan1})
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    (void)hipGetLastError();
    delete b;
}

int ffbf_batch_set_texts(ffbf_batch *b, const uint8_t *dev_texts, const uint64_t *offsets, const uint32_t *filter_of_text,
                         uint32_t n_texts)
{
    RET_IF(check_batch(b));
    if (!offsets || (n_texts && (!filter_of_text || !dev_texts))) return fail(FFBF_EINVAL, "null argument");
    if (n_texts >= (1u << (64 - kBatchOffsetBits))) return fail(FFBF_EINVAL, "too many texts in one batch");
    RET_IF(sync_filters(b, false));
    const uint32_t nf = (uint32_t)b->filters.size();
    std::vector<BatchText> tx(n_texts);
    uint64_t total = 0, maxn = 0;
    for (uint32_t i = 0; i < n_texts; i++) {
        if (offsets[i + 1] < offsets[i]) return fail(FFBF_EINVAL, "offsets must be non-decreasing");
        if (filter_of_text[i] >= nf) return fail(FFBF_EINVAL, "filter id out of range");
        tx[i].start = offsets[i];
        tx[i].n = offsets[i + 1] - offsets[i];
        tx[i].filter = filter_of_text[i];
        tx[i].reserved = 0;
        if (tx[i].n >= (1ull << kBatchOffsetBits)) return fail(FFBF_EINVAL, "text longer than 2^40 bytes");
        total += tx[i].n;
        maxn = std::max(maxn, tx[i].n);
    }
    // chunks of chunk_iters wave-iterations (kIterUnits units of 16 bytes, or
    // kIterUnits2 in the LDS class at two units per lane), ordered by chunk
    // class, filter, then text
    std::vector<uint32_t> order(n_texts);
    for (uint32_t i = 0; i < n_texts; i++) order[i] = i;
    // chunk class of a filter: 0 when its prefilter is held in LDS
    auto cls = [&](uint32_t f) -> uint32_t {
        const ffbf_ctx *c = b->filters[f];
        return mode_in_lds(c->mode) ? 0u : 1u;
    };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        const uint32_t cx = cls(tx[x].filter), cy = cls(tx[y].filter);
        return cx != cy ? cx < cy : tx[x].filter < tx[y].filter;
    });
    const uint32_t ci = b->tuning.chunk_iters ? b->tuning.chunk_iters : 32u;
    std::vector<BatchChunk> chunks;
    chunks.reserve(total / (16ull * kIterUnits * ci) + n_texts + 1);
    for (uint32_t i : order) {
        const uint64_t lead = (uint64_t)((uintptr_t)(dev_texts + tx[i].start) & 15u);
        const uint64_t npos = tx[i].n >= b->W ? tx[i].n - b->W + 1 : 0;
        if (!npos) continue;
        const uint64_t units = (lead + npos + 15u) >> 4;
        const uint64_t iu = (cls(tx[i].filter) == 0u && units_per_lane(b) == 2u) ? kIterUnits2 : kIterUnits;
        const uint64_t iters = (units + iu - 1u) / iu;
        if (iters > 0xFFFFFFFFull) return fail(FFBF_EINVAL, "text too long for one batch");
        for (uint64_t it = 0; it < iters; it += ci) {
            BatchChunk c;
            c.text = i;
            c.filter = tx[i].filter;
            c.it0 = (uint32_t)it;
            c.n_it = (uint32_t)std::min<uint64_t>(ci, iters - it);
            chunks.push_back(c);
        }
    }
    // the captured graph holds these pointers: drop it whenever one changes
    // (a replay must never read a freed descriptor buffer)
    const BatchText *old_texts = b->d_texts.p;
    const BatchChunk *old_chunks = b->d_chunks.p;
    // grow geometrically so that a stream of batches settles quickly (every
    // move costs a re-capture)
    if (tx.size() > b->d_texts.cap)
        RET_IF(b->d_texts.ensure(std::max<uint64_t>(2 * tx.size(), 1024), "batch text descriptors"));
    if (chunks.size() > b->d_chunks.cap)
        RET_IF(b->d_chunks.ensure(std::max<uint64_t>(2 * chunks.size(), 4096), "batch chunks"));
    if (!b->d_texts.p) RET_IF(b->d_texts.ensure(1024, "batch text descriptors"));
    if (!b->d_chunks.p) RET_IF(b->d_chunks.ensure(4096, "batch chunks"));
    if (b->d_texts.p != old_texts || b->d_chunks.p != old_chunks || b->texts != dev_texts) drop_graph(b);
    b->texts = dev_texts;
    b->n_texts = n_texts;
    b->n_chunks = (uint32_t)chunks.size();
    b->total_bytes = total;
    b->max_text = maxn;
    b->h_texts = tx;
    const uint64_t want = std::max<uint64_t>(1 << 16, total >> 10);
    if (b->cand.cap < want) {
        drop_graph(b);
        RET_IF(b->cand.ensure(std::max(want, 2 * b->cand.cap), "batch candidates"));
        RET_IF(b->cstage.ensure(b->n_cu, b->cand.cap));
        RET_IF(refresh_filters(b));
    }
    if (n_texts) HIP_TRY(hipMemcpyAsync(b->d_texts.p, tx.data(), tx.size() * sizeof(BatchText), hipMemcpyHostToDevice, b->stream));
    if (!chunks.empty())
        HIP_TRY(hipMemcpyAsync(b->d_chunks.p, chunks.data(), chunks.size() * sizeof(BatchChunk), hipMemcpyHostToDevice,
                               b->stream));
    std::vector<uint32_t> active(b->filters.size(), 0u);
    for (const BatchChunk &ch : chunks) active[ch.filter] = 1u;  // texts with windows only
    HIP_TRY(hipMemcpyAsync(b->d_active.p, active.data(), active.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           b->stream));
    uint64_t n0 = 0;
    while (n0 < chunks.size() && cls(chunks[n0].filter) == 0u) n0++;
    const unsigned long long ctl[8] = {0ull, n0, 0ull, n_texts, 0ull, chunks.size() - n0, n0, 0ull};
    HIP_TRY(hipMemcpyAsync(b->ctl.p, ctl, sizeof ctl, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->have_texts = true;
    b->ran = false;
    return FFBF_OK;
}

int ffbf_batch_run(ffbf_batch *b, uint64_t *n_matches)
{
    RET_IF(check_batch(b));
    RET_IF(sync_filters(b, true));
    if (!b->have_texts) return fail(FFBF_ESTATE, "no texts set");
    for (int attempt = 0; attempt < 3; attempt++) {
        RET_IF(run_once(b));
        const uint64_t nc = b->h_counters[C_CAND], nm = b->h_counters[C_MATCH];
        if (nc <= b->cand.cap && nm <= b->matches.cap) break;
        // grow and run again: every stage is a pure function of its inputs
        drop_graph(b);
        if (nc > b->cand.cap) {
            RET_IF(b->cand.ensure(std::max(2 * nc, 2 * b->cand.cap), "batch candidates"));
            RET_IF(b->cstage.ensure(b->n_cu, b->cand.cap));
        }
        if (nm > b->matches.cap) RET_IF(b->matches.ensure(std::max(2 * nm, 2 * b->matches.cap), "batch matches"));
        RET_IF(refresh_filters(b));
        if (attempt == 2) return fail(FFBF_EDEVICE, "batch outputs kept overflowing");
    }
    b->n_cand = b->h_counters[C_CAND];
    b->n_match = b->h_counters[C_MATCH];
    b->ran = true;
    b->sorted_valid = false;
    b->prefetched = prefetch_count(b);
    RET_IF(sort_on_device(b));
    if (n_matches) *n_matches = b->n_match;
    return FFBF_OK;
}

int ffbf_batch_get_matches(ffbf_batch *b, ffbf_batch_match *out, uint64_t cap, uint64_t *n)
{
    RET_IF(check_batch(b));
    if (!b->ran) return fail(FFBF_ESTATE, "batch has not run");
    if (n) *n = b->n_match;
    if (cap < b->n_match) return fail(FFBF_ECAPACITY, "match buffer too small");
    if (b->n_match && !out) return fail(FFBF_EINVAL, "null output");
    if (!b->sorted_valid) {
        b->sorted.resize(b->n_match);
        if (b->n_match <= b->prefetched) {
            std::memcpy(b->sorted.data(), b->h_match, b->n_match * sizeof(BatchMatchRec));
        } else if (b->n_match) {
            HIP_TRY(hipMemcpyAsync(b->sorted.data(), b->matches.p, b->n_match * sizeof(BatchMatchRec),
                                   hipMemcpyDeviceToHost, b->stream));
            HIP_TRY(hipStreamSynchronize(b->stream));
        }
        if (!b->sorted_on_device) {
            // by (text, pattern id, offset): stable passes, least significant field first
            std::vector<BatchMatchRec> tmp;
            ffbf::sort_by_major_offset(b->sorted.data(), b->sorted.size(), [](const BatchMatchRec &r) { return r.id; },
                                       tmp);
            ffbf::radix_pass_sort(b->sorted.data(), b->sorted.size(), [](const BatchMatchRec &r) { return r.text; }, tmp);
        }
        b->sorted_valid = true;
    }
    for (uint64_t i = 0; i < b->n_match; i++) {
        out[i].text_id = b->sorted[i].text;
        out[i].pattern_id = b->sorted[i].id;
        out[i].offset = b->sorted[i].offset;
    }
    return FFBF_OK;
}

int ffbf_batch_run_per_text(ffbf_batch *b, uint64_t *n_matches)
{
    RET_IF(check_batch(b));
    RET_IF(sync_filters(b, true));
    if (!b->have_texts) return fail(FFBF_ESTATE, "no texts set");
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<BatchMatchRec> all;
    std::vector<ffbf_match> one;
    for (uint32_t i = 0; i < b->n_texts; i++) {
        ffbf_ctx *c = b->filters[b->h_texts[i].filter];
        uint64_t nm = 0;
        RET_IF(ffbf_run_device(c, b->texts + b->h_texts[i].start, b->h_texts[i].n, &nm));
        one.resize(nm);
        RET_IF(ffbf_get_matches(c, one.data(), nm, &nm));
        for (const ffbf_match &m : one) all.push_back(BatchMatchRec{i, m.pattern_id, m.offset});
    }
    // each text ran to completion (with its own synchronisation) inside the loop
    b->run_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    b->scan_ms = -1.0f;
    b->sorted = std::move(all);
    b->n_match = b->sorted.size();
    b->sorted_valid = true;
    b->prefetched = 0;
    b->ran = true;
    if (n_matches) *n_matches = b->n_match;
    return FFBF_OK;
}

int ffbf_batch_set_tuning(ffbf_batch *b, const ffbf_batch_tuning *t)
{
    RET_IF(check_batch(b));
    if (!t) return fail(FFBF_EINVAL, "null tuning");
    if (t->block != 0 && t->block != 512 && t->block != 1024) return fail(FFBF_EINVAL, "block must be 0, 512 or 1024");
    if (t->probes > b->k) return fail(FFBF_EINVAL, "probes must be 0..k");
    if (t->chunk_iters > (1u << 20)) return fail(FFBF_EINVAL, "chunk_iters too large");
    if (t->bytes_per_lane != 0 && t->bytes_per_lane != 16 && t->bytes_per_lane != 32)
        return fail(FFBF_EINVAL, "bytes_per_lane must be 0, 16 or 32");
    RET_IF(sync_filters(b, true));
    const uint32_t upl_before = units_per_lane(b);
    const bool new_chunk =
        (t->chunk_iters ? t->chunk_iters : 32u) != (b->tuning.chunk_iters ? b->tuning.chunk_iters : 32u);
    b->tuning = *t;
    const bool rechunk = new_chunk || units_per_lane(b) != upl_before;
    if (!b->tuning.block) b->tuning.block = 1024;
    if (!b->tuning.chunk_iters) b->tuning.chunk_iters = 32;
    drop_graph(b);
    RET_IF(refresh_filters(b));
    if (rechunk) RET_IF(rechunk_texts(b));  // re-cut the current texts with the new chunk geometry
    return FFBF_OK;
}

int ffbf_batch_get_tuning(ffbf_batch *b, ffbf_batch_tuning *t)
{
    RET_IF(check_batch(b));
    if (!t) return fail(FFBF_EINVAL, "null tuning");
    *t = b->tuning;
    t->probes = b->probes_eff;
    t->bytes_per_lane = 16u * units_per_lane(b);
    return FFBF_OK;
}

int ffbf_batch_autotune(ffbf_batch *b, ffbf_batch_tuning *best, float *best_ms)
{
    RET_IF(check_batch(b));
    if (!b->have_texts) return fail(FFBF_ESTATE, "autotune needs a batch of texts");
    RET_IF(sync_filters(b, true));
    const ffbf_batch_tuning start = b->tuning;
    ffbf_batch_tuning top = start;
    float top_ms = 1e30f;
    // the candidates are ranked by their timed runs
    struct TimingOn {
        ffbf_batch *b;
        bool was;
        ~TimingOn() { b->timing = was; }
    } timing_on{b, b->timing};
    b->timing = true;
    // 256 was 2% ahead of 128 on the 64-set batch, 512 and up fell behind
    // (fewer chunks than workgroups can balance); 16 never won
    const uint32_t chunk_opts[] = {32u, 64u, 128u, 256u};
    const uint32_t block_opts[] = {1024u, 512u};
    const bool any_lds = b->pre_words_max > 0;
    // 32 bytes per lane only where the LDS-class scan can take it
    const bool wide = any_lds && takes_32_bytes(b);
    const uint32_t bpl_opts[] = {32u, 16u};
    for (uint32_t bpl : bpl_opts) {
        if (bpl == 32u && !wide) continue;
        for (uint32_t blk : block_opts) {
            for (uint32_t p = 1; p <= std::min(3u, b->k); p++) {
                if (!any_lds && p > 1) break;
                for (uint32_t ci : chunk_opts) {
                    ffbf_batch_tuning t{ci, p, blk, start.use_graph, bpl};
                    RET_IF(ffbf_batch_set_tuning(b, &t));
                    uint64_t nm = 0;
                    RET_IF(ffbf_batch_run(b, &nm));  // warm (and capture)
                    RET_IF(ffbf_batch_run(b, &nm));
                    if (b->run_ms < top_ms) {
                        top_ms = b->run_ms;
                        top = t;
                    }
                }
            }
        }
    }
    RET_IF(ffbf_batch_set_tuning(b, &top));
    if (best) *best = top;
    if (best_ms) *best_ms = top_ms;
    return FFBF_OK;
}

int ffbf_batch_set_timing(ffbf_batch *b, int on)
{
    RET_IF(check_batch(b));
    b->timing = on != 0;
    return FFBF_OK;
}

int ffbf_batch_timings(ffbf_batch *b, float *run_ms, float *scan_ms)
{
    RET_IF(check_batch(b));
    if (run_ms) *run_ms = b->run_ms;
    if (scan_ms) *scan_ms = b->scan_ms;
    return FFBF_OK;
}

}  // extern "C"
