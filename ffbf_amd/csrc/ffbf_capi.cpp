// ffbf_capi.cpp -- host context behind the C ABI (include/ffbf.h).
//
// The context owns every device buffer of the path: F and R (resident between
// scans), the pattern set with its prefix hash pairs, the candidate list, the
// survivor bitmap/list/index and the match list.  Stages are queued on one
// HIP stream; counters come back through pinned memory.  Buffers that depend
// on data sizes (candidates, matches, owned text) grow on demand: a stage
// whose output overflowed is simply queued again with room for all of it,
// which gives the same result because every stage is a pure function of its
// inputs (filter bits are OR-idempotent).
#include "ffbf.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "ffbf_ctx.h"
#include "ffbf_hostsort.h"

using ffbf::Cand;
using ffbf::MatchRec;

namespace ffbf {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

uint32_t pow2_at_least(uint64_t n)
{
    uint32_t l = 0;
    while ((1ull << l) < n) l++;
    return l;
}

int use_device(ffbf_ctx *c)
{
    HIP_TRY(hipSetDevice(c->device));
    return FFBF_OK;
}

int record(ffbf_ctx *c, int e)
{
    c->ev_recorded[e] = false;
    if (!c->timing) return FFBF_OK;
    HIP_TRY(hipEventRecord(c->ev[e], c->stream));
    c->ev_recorded[e] = true;
    return FFBF_OK;
}

int sync_counters(ffbf_ctx *c)
{
    HIP_TRY(hipMemcpyAsync(c->h_counters, c->counters.p, ffbf::C_COUNT * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FFBF_OK;
}

int clear_counters(ffbf_ctx *c, int first, int count)
{
    HIP_TRY(ffbf::launch_fill32((uint32_t *)(c->counters.p + first), (uint64_t)count * 2u, 0u, c->stream));
    return FFBF_OK;
}

// Expected fraction of F's bits that are set: exact after a filter load,
// else k random bits per hashed pattern, 1 - (1 - 1/M)^(k n).
double f_fill(ffbf_ctx *c)
{
    const double M = (double)c->f_words * 32.0;
    if (c->f_set_bits >= 0) return std::min(1.0, (double)c->f_set_bits / M);
    return 1.0 - std::exp((double)c->params.k * (double)c->n_long * std::log1p(-1.0 / M));
}

// Expected fill of an OR-fold onto fold_bits bits.
double fold_fill(ffbf_ctx *c, uint64_t fold_bits)
{
    const double ratio = (double)c->f_words * 32.0 / (double)fold_bits;
    return 1.0 - std::pow(1.0 - f_fill(c), ratio);
}

// log2 bits of the largest fold that fits one XCD's L2 (4 MiB on MI355X;
// the attribute is capped there in case it reports the sum over XCDs).
uint32_t l2_fold_log2(ffbf_ctx *c)
{
    int l2 = 0;
    if (hipDeviceGetAttribute(&l2, hipDeviceAttributeL2CacheSize, c->device) != hipSuccess || l2 <= 0) {
        (void)hipGetLastError();
        l2 = 4 << 20;
    }
    const uint64_t bytes = std::min<uint64_t>((uint64_t)l2, 4ull << 20);
    uint32_t g = 10;
    while ((1ull << (g + 1 - 3)) <= bytes && g + 1 < c->params.log2_bits) g++;
    return g;
}

void add_fold_job(ffbf_ctx *c, const uint32_t *src, uint64_t src_words, uint32_t *dst, uint32_t dst_words)
{
    c->fold_jobs[c->n_fold_jobs++] = ffbf_ctx::FoldJob{src, src_words, dst, dst_words};
}

// Launches the planned folds, in order, from the finished F.
int run_fold_jobs(ffbf_ctx *c)
{
    for (uint32_t i = 0; i < c->n_fold_jobs; i++) {
        const ffbf_ctx::FoldJob &j = c->fold_jobs[i];
        HIP_TRY(ffbf::launch_fold(j.src, j.src_words, j.dst, j.dst_words, c->stream));
    }
    return FFBF_OK;
}

// Picks how the scan pre-tests windows, after F changed or a mode was
// requested.  Automatic choice: the whole filter in LDS when it fits; else
// the cheapest of the LDS fold, F probed directly, and (F beyond one XCD's
// L2) an L2-sized fold, by an expected-cost model over the analytic fills.
// Plans the folds the mode needs (run_fold_jobs launches them once F is
// complete); nothing is read back.
// Split filter: an LDS fold of the cache part (tested by the functions
// below s) or F probed directly, its cache part first.  The same cost model
// as below, over the cache part's expected fill.
int choose_mode_split(ffbf_ctx *c)
{
    const ffbf::SplitMap &m = c->smap;
    const uint32_t S = m.s;
    if (c->mode_req == ffbf::MODE_LDS || c->mode_req == ffbf::MODE_GFOLD || c->mode_req == ffbf::MODE_HYBRID)
        return fail(FFBF_EINVAL, "split filter: scan mode must be -1, 0 or 1");
    if (c->mode_req == ffbf::MODE_FOLD && S == 0)
        return fail(FFBF_EINVAL, "split filter with s = 0 has no cache part to fold");
    const uint64_t cache_words = (uint64_t)m.cache / 32u;
    // the scan copies a prefilter into LDS in 16-byte units: at least 4 words
    const bool foldable = S > 0 && cache_words >= 4u;
    if (c->mode_req == ffbf::MODE_FOLD && !foldable)
        return fail(FFBF_EINVAL, "split filter: a cache part below 2^7 bits cannot be folded");
    const uint32_t max_words = ffbf::max_fold_words(c->device, true);
    uint32_t words = (uint32_t)std::min<uint64_t>(max_words, cache_words);
    if (c->fold_log2_req) {
        if (c->fold_log2_req < 7 || (1ull << (c->fold_log2_req - 5)) > max_words ||
            (1ull << (c->fold_log2_req - 5)) > cache_words)
            return fail(FFBF_EINVAL, "fold size out of range");
        words = 1u << (c->fold_log2_req - 5);
    }
    // s bits per hashed pattern land in the cache part (after a filter load
    // the whole filter's exact fill stands in for it)
    const double fc = c->f_set_bits >= 0
                          ? std::min(1.0, (double)c->f_set_bits / ((double)c->f_words * 32.0))
                          : 1.0 - std::exp((double)S * (double)c->n_long * std::log1p(-1.0 / (double)m.cache));
    const double ff = 1.0 - std::pow(1.0 - fc, (double)cache_words / (double)words);
    uint32_t P = c->probes_req ? c->probes_req : (ff < 0.01 ? 1u : (ff < 0.2 ? 2u : 3u));
    P = std::max(1u, std::min(std::min(P, S), 3u));
    bool fold = false;
    if (foldable && c->mode_req != ffbf::MODE_GLOBAL) {
        if (c->mode_req == ffbf::MODE_FOLD) {
            fold = true;
        } else {
            constexpr double kLds = 1.25, kL2 = 3.8, kFar = 18.0, kQueue = 0.8;
            const double first = (uint64_t)m.cache / 8u <= (4ull << 20) ? kL2 : kFar;
            const double cost_global = first * (1.0 + fc);
            const double cost_fold = kLds + std::pow(ff, (double)P) * kQueue + std::pow(ff, (double)S) * first;
            fold = cost_fold <= cost_global;
        }
    }
    if (fold) {
        RET_IF(c->fold.ensure(words, "fold"));
        add_fold_job(c, c->F.p, cache_words, c->fold.p, words);
        c->mode = ffbf::MODE_FOLD;
        c->pre_words = words;
        c->pre_fill = ff;
        c->probes = P;
    } else {
        c->mode = ffbf::MODE_GLOBAL;
        c->pre_words = (uint32_t)(S > 0 ? cache_words : ((uint64_t)m.emask + 1u) / 32u);
        c->pre_fill = fc;
        c->probes = 1;
    }
    return FFBF_OK;
}

// LDS prefilter bits tested up front for a prefilter of fill `fill`: enough
// to leave few windows queued.  Measured on 20 MiB-1.5 GiB texts: with k=4,
// fill 0.04 (10k patterns on a 1 Mbit fold) is fastest at 2 and fill 0.32
// (100k patterns) at 3; with k=3 (the 64-set batch's filters, DNA with
// tandem repeats), 3 bits win from fill ~0.09 on, whole filters in LDS and
// folds alike (2.9% over 34 filters held whole in LDS).
uint32_t auto_probes(const ffbf_ctx *c, double fill)
{
    const uint32_t k = c->params.k;
    if (c->probes_req) return std::min(c->probes_req, k);
    const double three_from = k <= 3 ? 0.09 : 0.2;
    return std::min(fill < 0.01 ? 1u : (fill < three_from ? 2u : 3u), k);
}

int plan_mode(ffbf_ctx *c)
{
    c->generation++;
    c->gfold_words = 0;
    c->n_fold_jobs = 0;
    if (c->split) return choose_mode_split(c);
    const uint32_t max_words = ffbf::max_fold_words(c->device);
    const uint32_t k = c->params.k;
    int st = FFBF_OK;
    uint32_t mode = 0;
    if (c->mode_req == ffbf::MODE_LDS || (c->mode_req < 0 && c->f_words <= max_words)) {
        if (c->f_words > max_words) return fail(FFBF_EINVAL, "scan mode 2: filter does not fit in LDS");
        mode = ffbf::MODE_LDS;
        c->pre_words = (uint32_t)c->f_words;
        c->pre_fill = f_fill(c);
    } else if (c->mode_req == ffbf::MODE_GLOBAL) {
        mode = ffbf::MODE_GLOBAL;
        c->pre_words = (uint32_t)std::min<uint64_t>(c->f_words, 0xFFFFFFFFu);
    } else if (c->mode_req == ffbf::MODE_GFOLD) {
        const uint32_t l2 = c->fold_log2_req ? c->fold_log2_req : l2_fold_log2(c);
        if (l2 < 10 || l2 > c->params.log2_bits || l2 > 31) return fail(FFBF_EINVAL, "global fold size out of range");
        const uint32_t words = 1u << (l2 - 5);
        RET_IF(c->gfold.ensure(words, "global fold"));
        add_fold_job(c, c->F.p, c->f_words, c->gfold.p, words);
        mode = ffbf::MODE_GFOLD;
        c->pre_words = words;
        c->gfold_words = words;
    } else {
        uint32_t words = max_words;
        if (c->fold_log2_req) {
            if (c->fold_log2_req < 7 || (1ull << (c->fold_log2_req - 5)) > max_words ||
                (1ull << (c->fold_log2_req - 5)) > c->f_words)
                return fail(FFBF_EINVAL, "fold size out of range");
            words = 1u << (c->fold_log2_req - 5);
        }
        words = (uint32_t)std::min<uint64_t>(words, c->f_words);
        const double ff = fold_fill(c, (uint64_t)words * 32u);
        uint32_t choice = ffbf::MODE_FOLD;
        uint32_t gl2 = 0;
        if (c->mode_req != ffbf::MODE_FOLD && c->mode_req != ffbf::MODE_HYBRID) {
            // Expected scan cost per window (picoseconds of the whole chip),
            // fitted to MI355X measurements: the LDS stage (hash + up to 3
            // fold bits) ~1.25; one random probe ~3.8 while its array fits
            // one XCD's L2 (265 G probes/s), ~18 beyond (55 G/s); ~0.8 per
            // window queued past the up-front bits.  A fold passing many
            // windows tests F's first bit before queueing: every window
            // passing the up-front bits probes F once, and only those that
            // pass it are queued.
            constexpr double kLds = 1.25, kL2 = 3.8, kFar = 18.0, kQueue = 0.8;
            const bool f_in_l2 = c->f_words * 4ull <= (4ull << 20);
            const double far = f_in_l2 ? kL2 : kFar;
            const double f = f_fill(c);
            const uint32_t P = std::min(auto_probes(c, ff), 3u);
            const double pass = std::pow(ff, (double)P);
            const bool early = fold_passes_many(ff, P);
            const double probed = early ? pass : std::pow(ff, (double)k);
            const double queued = early ? pass * f : pass;
            const double cost_fold = kLds + probed * far + queued * kQueue;
            const double cost_global = (1.0 + f + f * f) * far;
            double cost_gfold = 1e30, cost_hybrid = 1e30;
            if (!f_in_l2) {
                gl2 = l2_fold_log2(c);
                const double g = fold_fill(c, 1ull << gl2);
                cost_gfold = (1.0 + g + g * g) * kL2 + std::pow(g, (double)k) * kFar;
                // an LDS fold passing many windows: the LDS stage overlaps the
                // L2 probes of the windows it lets through, whichever of the
                // two takes longer; else the windows passing all its bits
                // probe the L2 fold in the drain instead of F
                cost_hybrid = early ? std::max(kLds, pass * cost_gfold)
                                    : kLds + std::pow(ff, (double)k) * cost_gfold + pass * kQueue;
            }
            if (cost_global < cost_fold && cost_global <= cost_gfold && cost_global <= cost_hybrid)
                choice = ffbf::MODE_GLOBAL;
            else if (cost_hybrid < cost_fold && cost_hybrid < cost_gfold)
                choice = ffbf::MODE_HYBRID;
            else if (cost_gfold < cost_fold)
                choice = ffbf::MODE_GFOLD;
        } else if (c->mode_req == ffbf::MODE_HYBRID) {
            if (c->f_words * 4ull <= (4ull << 20)) return fail(FFBF_EINVAL, "scan mode 4: filter already fits L2");
            gl2 = l2_fold_log2(c);
            if ((1ull << (gl2 - 5)) <= words) return fail(FFBF_EINVAL, "scan mode 4: L2 fold not larger than the LDS fold");
            choice = ffbf::MODE_HYBRID;
        }
        if (choice == ffbf::MODE_FOLD || choice == ffbf::MODE_HYBRID) {
            RET_IF(c->fold.ensure(words, "fold"));
            if (choice == ffbf::MODE_HYBRID) {
                // the L2 fold first; the LDS fold is then the fold of the L2
                // fold (both sizes are powers of two, so folding onto 2^a
                // bits after 2^b > 2^a bits gives the same bits), which
                // reads 4 MiB instead of all of F
                const uint32_t gw = 1u << (gl2 - 5);
                RET_IF(c->gfold.ensure(gw, "global fold"));
                add_fold_job(c, c->F.p, c->f_words, c->gfold.p, gw);
                c->gfold_words = gw;
                if (gw > words)
                    add_fold_job(c, c->gfold.p, gw, c->fold.p, words);
                else
                    add_fold_job(c, c->F.p, c->f_words, c->fold.p, words);
                mode = ffbf::MODE_HYBRID;
            } else {
                add_fold_job(c, c->F.p, c->f_words, c->fold.p, words);
                mode = ffbf::MODE_FOLD;
            }
            c->pre_words = words;
            c->pre_fill = ff;
        } else if (choice == ffbf::MODE_GLOBAL) {
            mode = ffbf::MODE_GLOBAL;
            c->pre_words = (uint32_t)c->f_words;
        } else {
            const uint32_t gw = 1u << (gl2 - 5);
            RET_IF(c->gfold.ensure(gw, "global fold"));
            add_fold_job(c, c->F.p, c->f_words, c->gfold.p, gw);
            mode = ffbf::MODE_GFOLD;
            c->pre_words = gw;
            c->gfold_words = gw;
        }
    }
    RET_IF(st);
    c->mode = mode;
    // LDS prefilter bits tested up front: enough to leave few windows queued
    if (c->probes_req)
        c->probes = std::min(c->probes_req, k);
    else
        c->probes = ffbf::mode_in_lds(mode) ? auto_probes(c, c->pre_fill) : 1u;
    return FFBF_OK;
}

// The mode and its folds, from the F now on the stream.
int choose_mode(ffbf_ctx *c)
{
    RET_IF(plan_mode(c));
    return run_fold_jobs(c);
}

uint32_t scan_block(ffbf_ctx *c)
{
    if (c->split) return c->mode == ffbf::MODE_FOLD ? 1024u : 256u;
    if (ffbf::mode_in_lds(c->mode)) return c->block_req == 512u ? 512u : 1024u;
    return 256u;
}

int ensure_cand(ffbf_ctx *c, uint64_t n)
{
    if (n > c->cand.cap || !c->cand.p) RET_IF(c->cand.ensure(std::max(n, 2 * c->cand.cap), "candidates"));
    RET_IF(c->cstage.ensure(c->n_cu, c->cand.cap));
    return FFBF_OK;
}

// Extra atomics (patterns x k x folds) up to which the atomic build sets the
// folds' bits itself: 2^17 of them take ~5 us at ~27 G/s, less than a fold
// launch and the clear it saves.
constexpr uint64_t kFusedFoldAtomics = 1ull << 17;

// Appends the 8192-word spans of p[0..words) to a clear list.
void add_spans(std::vector<ffbf::ZeroSpan> &zs, uint32_t *p, uint64_t words)
{
    for (uint64_t o = 0; o < words; o += 8192u)
        zs.push_back(ffbf::ZeroSpan{p + o, (uint32_t)std::min<uint64_t>(8192u, words - o)});
}

// Puts a clear list on the device unless the copy there (`have`) is the same.
int upload_spans(ffbf_ctx *c, const std::vector<ffbf::ZeroSpan> &zs, ffbf::DevBuf<ffbf::ZeroSpan> &dst,
                 std::vector<ffbf::ZeroSpan> &have)
{
    if (dst.p && zs.size() == have.size() &&
        std::equal(zs.begin(), zs.end(), have.begin(),
                   [](const ffbf::ZeroSpan &a, const ffbf::ZeroSpan &b) { return a.p == b.p && a.words == b.words; }))
        return FFBF_OK;
    have.clear();
    RET_IF(dst.ensure(zs.size(), "clear list"));
    HIP_TRY(hipMemcpyAsync(dst.p, zs.data(), zs.size() * sizeof(ffbf::ZeroSpan), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    have = zs;
    return FFBF_OK;
}

// Appends the scan state a scan starts from zero: R, the candidate region
// counts, the scan's counters and the survivor index.
void add_scan_spans(ffbf_ctx *c, std::vector<ffbf::ZeroSpan> &zs)
{
    add_spans(zs, c->R.p, c->f_words);
    if (c->cstage.counts.p) add_spans(zs, c->cstage.counts.p, c->cstage.n_regions);
    add_spans(zs, (uint32_t *)(c->counters.p + ffbf::C_CAND), 6u);   // C_CAND, C_SURV, C_MATCH
    add_spans(zs, (uint32_t *)(c->counters.p + ffbf::C_SPILL), 4u);  // C_SPILL, C_CVALID
    if (c->table.p) add_spans(zs, c->table.p, 2ull * c->table_size);
}

// The atomic build's clear covered the scan state of the buffers in use now,
// and nothing has written it since.
bool scan_state_clean(const ffbf_ctx *c)
{
    return c->scan_state_clean && c->clean_key[0] == c->cstage.counts.p && c->clean_size[0] == c->cstage.n_regions &&
           c->clean_key[1] == c->table.p && c->clean_size[1] == c->table_size;
}

// The spans a scan clears: R, the candidate region counts, the scan's
// counters (C_CAND, C_SURV, C_MATCH, C_SPILL, C_CVALID) and the survivor
// index, which the pattern pass fills.
int build_zero_spans(ffbf_ctx *c)
{
    if (c->d_zero.p && c->zero_key[0] == c->cstage.counts.p && c->zero_key[1] == c->table.p &&
        c->zero_table_size == c->table_size)
        return FFBF_OK;
    std::vector<ffbf::ZeroSpan> zs;
    add_scan_spans(c, zs);
    RET_IF(c->d_zero.ensure(zs.size(), "clear list"));
    HIP_TRY(hipMemcpyAsync(c->d_zero.p, zs.data(), zs.size() * sizeof(ffbf::ZeroSpan), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->n_zero = (uint32_t)zs.size();
    c->zero_key[0] = c->cstage.counts.p;
    c->zero_key[1] = c->table.p;
    c->zero_table_size = c->table_size;
    return FFBF_OK;
}

// Clears R, the candidate list and its staging, the scan's counters and the
// survivor index before the first piece of a scan (one launch).
int enqueue_scan_begin(ffbf_ctx *c)
{
    RET_IF(build_zero_spans(c));
    RET_IF(record(c, EV_S0));
    if (!scan_state_clean(c)) HIP_TRY(ffbf::launch_zero_spans(c->d_zero.p, c->n_zero, c->n_cu, c->stream));
    c->scan_state_clean = false;
    RET_IF(record(c, EV_S1));
    return FFBF_OK;
}

constexpr uint32_t kLdsRMaxWords = 2048;  // 2^16 bits

void filter_scan_args(const ffbf_ctx *c, ffbf::ScanArgs &a)
{
    a.F = c->F.p;
    a.R = c->R.p;
    a.sbox = c->sbox.p;
    a.W = c->params.window;
    a.k = c->params.k;
    a.mask = c->mask;
    switch (c->mode) {
    case ffbf::MODE_FOLD:
    case ffbf::MODE_HYBRID: a.pre = c->fold.p; break;
    case ffbf::MODE_GFOLD: a.pre = c->gfold.p; break;
    default: a.pre = c->F.p; break;
    }
    a.pre_words = c->pre_words;
    a.pre_mask = c->mode == ffbf::MODE_GLOBAL ? c->mask : (uint32_t)((uint64_t)c->pre_words * 32u - 1u);
    a.pre_log2w = pow2_at_least(c->pre_words);
    a.P = ffbf::mode_in_lds(c->mode) ? std::min(c->probes, 3u) : 1u;
    a.pre_exact = (c->mode == ffbf::MODE_GLOBAL || c->mode == ffbf::MODE_LDS) ? 1u : 0u;
    a.lds = ffbf::mode_in_lds(c->mode) ? 1u : 0u;
    a.f_first = (c->mode == ffbf::MODE_FOLD && !c->split && fold_passes_many(c->pre_fill, a.P)) ? 1u : 0u;
    if (c->mode == ffbf::MODE_HYBRID) {
        // the L2-resident fold, then F, behind the LDS fold: its first bit
        // before queueing when the LDS fold passes many windows, the rest in
        // the drain
        a.pre2 = c->gfold.p;
        a.pre2_mask = (uint32_t)((uint64_t)c->gfold_words * 32u - 1u);
        a.f_first = fold_passes_many(c->pre_fill, a.P) ? 1u : 0u;
    }
    // whole F in LDS and small: R is gathered beside it in LDS.  Below 2^17
    // bits the candidates' atomics crowd onto a few cache lines of a global
    // R (a 2^15-bit filter with dense candidates scanned 2.9x slower for
    // it); at 2^17 bits and above they do not, and a larger copy would only
    // add to each workgroup's final OR
    a.r_lds = (c->mode == ffbf::MODE_LDS && !c->split && c->pre_words <= kLdsRMaxWords &&
               2ull * c->pre_words <= (uint64_t)ffbf::max_fold_words(c->device))
                  ? 1u
                  : 0u;
    a.method = ffbf::METHOD_PAIR;
    if (c->split) {
        a.method = ffbf::METHOD_SPLIT;
        a.split = c->smap;
        a.pre_exact = 0;
        if (c->mode == ffbf::MODE_GLOBAL) {
            // F's bit of function 0: the cache part, or the external part
            // when no function addresses the cache part
            a.pre = c->smap.s > 0 ? c->F.p : c->F.p + c->smap.cache / 32u;
            a.pre_mask = c->smap.s > 0 ? c->smap.cmask : c->smap.emask;
        }
    }
}

// Scans the windows of text[0..n) (device memory), recording their offsets
// plus base: a text can be scanned piece by piece, each piece starting at the
// first window the previous one did not cover and carrying its W-1 byte tail.
// defer_gather: only the compaction's offsets here; the gather goes into the
// pattern pass's launch (enqueue_pass), which must follow.
int enqueue_scan_range(ffbf_ctx *c, const uint8_t *text, uint64_t n, uint64_t base, bool mark_end, bool defer_gather)
{
    c->scan_state_clean = false;
    ffbf::ScanArgs a{};
    filter_scan_args(c, a);
    a.text = text;
    a.n = n;
    a.off_base = base;
    c->cstage.fill(a, 0u, c->cstage.n_regions);
    a.counters = c->counters.p;
    uint32_t waves = 0;
    HIP_TRY(ffbf::launch_scan(a, c->mode, scan_block(c), c->n_cu, c->stream, &waves));
    if (mark_end) RET_IF(record(c, EV_S2));  // the scan kernel alone ends here
    c->gather_pending = false;
    if (defer_gather) {
        HIP_TRY(ffbf::launch_cand_offsets(c->cstage.raw.p, c->cstage.raw.cap, c->cstage.counts.p, c->cstage.offsets.p,
                                          c->cstage.n_regions, c->cstage.region, c->cand.p, c->cand.cap,
                                          c->counters.p, c->stream, &c->pending_gather, waves));
        c->gather_pending = true;
    } else {
        HIP_TRY(ffbf::launch_cand_compact(c->cstage.raw.p, c->cstage.raw.cap, c->cstage.counts.p, c->cstage.offsets.p,
                                          c->cstage.n_regions, c->cstage.region, c->cand.p, c->cand.cap,
                                          c->counters.p, c->n_cu, c->stream, waves));
    }
    return FFBF_OK;
}

int enqueue_scan(ffbf_ctx *c, bool defer_gather = false)
{
    RET_IF(enqueue_scan_begin(c));
    RET_IF(enqueue_scan_range(c, c->text, c->text_n, 0, true, defer_gather));
    RET_IF(record(c, EV_S3));
    return FFBF_OK;
}

// fresh: the scan just cleared the survivor index and count (no earlier
// pass since), so the pass need not clear them again.
int enqueue_pass(ffbf_ctx *c, bool fresh)
{
    c->scan_state_clean = false;
    RET_IF(record(c, EV_P0));
    if (!fresh) {
        HIP_TRY(ffbf::launch_fill32(c->table.p, 2ull * c->table_size, 0u, c->stream));
        RET_IF(clear_counters(c, ffbf::C_SURV, 1));
    }
    ffbf::PassArgs a{};
    a.pat_hash = c->pat_hash.p;
    a.off = c->off.p;
    a.n = c->n_pat;
    a.W = c->params.window;
    a.k = c->params.k;
    a.mask = c->mask;
    a.R = c->R.p;
    a.bitmap = c->bitmap.p;
    a.counters = c->counters.p;
    a.table = c->table.p;
    a.slot_chain = c->table.p + c->table_size;
    a.chain_next = c->chain_next.p;
    a.table_mask = c->table_size - 1u;
    a.method = c->split ? ffbf::METHOD_SPLIT : ffbf::METHOD_PAIR;
    a.split = c->smap;
    if (c->gather_pending) {
        c->gather_pending = false;
        HIP_TRY(ffbf::launch_pass_gather(a, c->pending_gather, c->n_cu, c->stream));
    } else {
        HIP_TRY(ffbf::launch_pattern_pass(a, c->n_cu, c->stream));
    }
    RET_IF(record(c, EV_P1));
    return FFBF_OK;
}

int enqueue_verify(ffbf_ctx *c, bool fresh)
{
    c->scan_state_clean = false;
    RET_IF(record(c, EV_V0));
    if (!fresh) RET_IF(clear_counters(c, ffbf::C_MATCH, 1));
    ffbf::VerifyArgs a{};
    a.text = c->text;
    a.n = c->text_n;
    a.cand = c->cand.p;
    a.cand_cap = c->cand.cap;
    a.bytes = c->bytes.p;
    a.off = c->off.p;
    a.pat_hash = c->pat_hash.p;
    a.table = c->table.p;
    a.slot_chain = c->table.p + c->table_size;
    a.chain_next = c->chain_next.p;
    a.table_mask = c->table_size - 1u;
    a.out = c->matches.p;
    a.out_cap = c->matches.cap;
    a.counters = c->counters.p;
    HIP_TRY(ffbf::launch_verify(a, c->n_cu, c->stream));
    RET_IF(record(c, EV_V1));
    return FFBF_OK;
}

int check_ctx(ffbf_ctx *c)
{
    if (!c) return fail(FFBF_EINVAL, "null context");
    return use_device(c);
}

int set_text(ffbf_ctx *c, const uint8_t *dev_text, uint64_t n)
{
    if (!c->built) return fail(FFBF_ESTATE, "scan before build (or filter load)");
    if (!dev_text && n) return fail(FFBF_EINVAL, "null text");
    c->text = dev_text;
    c->text_n = n;
    // room for twice the expected false-positive windows (fill^k each), at
    // least 1 per KiB and at most 1 per 16 bytes up front: an overflowing
    // list makes the path scan again
    const double fp = std::pow(f_fill(c), (double)c->params.k);
    const uint64_t fp_want = std::min<uint64_t>((uint64_t)(2.0 * fp * (double)n), n >> 4);
    const uint64_t want = std::max<uint64_t>(std::max<uint64_t>(65536, n >> 10), fp_want);
    RET_IF(ensure_cand(c, want));
    if (!c->matches.p) RET_IF(c->matches.ensure(65536, "matches"));
    c->scanned = c->passed = c->verified = false;
    c->sorted_valid = false;
    c->prefetched = 0;
    return FFBF_OK;
}

// Scan, growing the candidate list until it holds every candidate.
int do_scan(ffbf_ctx *c)
{
    RET_IF(enqueue_scan(c));
    RET_IF(sync_counters(c));
    if (c->h_counters[ffbf::C_CAND] > c->cand.cap) {
        RET_IF(ensure_cand(c, 2 * c->h_counters[ffbf::C_CAND]));
        RET_IF(enqueue_scan(c));
        RET_IF(sync_counters(c));
    }
    c->n_cand = c->h_counters[ffbf::C_CAND];
    c->scanned = true;
    c->passed = c->verified = false;
    return FFBF_OK;
}

// Sorts the verified matches on the device when they pack into 64-bit keys
// (ids < 2^24, offsets < 2^40); otherwise the host sorts them on copy-out.
// Grows the pinned landing area to at least n records (geometrically: an
// allocation of pinned memory is slow).
int ensure_host_matches(ffbf_ctx *c, uint64_t n)
{
    if (n <= c->h_match_cap && c->h_match) return FFBF_OK;
    const uint64_t want = std::max<uint64_t>({n, 2 * c->h_match_cap, kMatchPrefetch});
    if (c->h_match) (void)hipHostFree(c->h_match);
    c->h_match = nullptr;
    c->h_match_cap = 0;
    c->prefetched = 0;
    HIP_TRY(hipHostMalloc((void **)&c->h_match, want * sizeof(MatchRec), hipHostMallocDefault));
    c->h_match_cap = want;
    return FFBF_OK;
}

// Queues a copy of the match list's first kMatchPrefetch records (as many as
// the list holds room for) to the pinned landing area, after the stages
// already queued; valid once the stream has reached it.
int prefetch_matches(ffbf_ctx *c)
{
    c->prefetched = 0;
    const uint64_t k = std::min<uint64_t>(c->matches.cap, kMatchPrefetch);
    if (!k || !c->matches.p) return FFBF_OK;
    RET_IF(ensure_host_matches(c, k));
    HIP_TRY(hipMemcpyAsync(c->h_match, c->matches.p, k * sizeof(MatchRec), hipMemcpyDeviceToHost, c->stream));
    c->prefetched = k;
    return FFBF_OK;
}

int sort_on_device(ffbf_ctx *c)
{
    c->sorted_on_device = false;
    // a few thousand records sort faster on the host at copy-out than the
    // device radix sort's launches take
    if (c->n_match < ffbf::kDeviceSortMin || !ffbf::sort_packable(c->n_pat, c->text_n)) return FFBF_OK;
    const int end_bit = ffbf::sort_end_bit(c->n_pat);
    size_t tb = 0;
    HIP_TRY(ffbf::sort_temp_bytes(c->n_match, end_bit, &tb));
    if (2 * c->n_match > c->sort_keys.cap)
        RET_IF(c->sort_keys.ensure(std::max<uint64_t>(4 * c->n_match, 1 << 17), "sort keys"));
    if (tb > c->sort_temp.cap) RET_IF(c->sort_temp.ensure(std::max<uint64_t>(2 * tb, 1 << 20), "sort scratch"));
    tb = c->sort_temp.cap;
    HIP_TRY(ffbf::sort_matches(c->matches.p, c->n_match, c->sort_keys.p, c->sort_keys.p + c->n_match, c->sort_temp.p,
                               tb, end_bit, c->stream));
    c->sorted_on_device = true;
    c->prefetched = 0;  // the records were copied out before the sort reordered them
    return FFBF_OK;
}

int do_verify(ffbf_ctx *c)
{
    RET_IF(enqueue_verify(c, false));
    RET_IF(prefetch_matches(c));
    RET_IF(sync_counters(c));
    if (c->h_counters[ffbf::C_MATCH] > c->matches.cap) {
        RET_IF(c->matches.ensure(std::max<uint64_t>(2 * c->h_counters[ffbf::C_MATCH], 2 * c->matches.cap), "matches"));
        RET_IF(enqueue_verify(c, false));
        RET_IF(prefetch_matches(c));
        RET_IF(sync_counters(c));
    }
    c->n_match = c->h_counters[ffbf::C_MATCH];
    c->verified = true;
    c->sorted_valid = false;
    return sort_on_device(c);
}

int load_sorted_matches(ffbf_ctx *c)
{
    if (c->sorted_valid) return FFBF_OK;
    c->sorted.resize(c->n_match);
    if (c->n_match) {
        if (c->n_match > c->prefetched) {
            // through pinned memory: a copy to pageable memory is staged
            // and blocks longer
            RET_IF(ensure_host_matches(c, c->n_match));
            HIP_TRY(hipMemcpyAsync(c->h_match, c->matches.p, c->n_match * sizeof(MatchRec), hipMemcpyDeviceToHost,
                                   c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
        }
        std::memcpy(c->sorted.data(), c->h_match, c->n_match * sizeof(MatchRec));
    }
    if (!c->sorted_on_device)
        ffbf::sort_by_major_offset(c->sorted.data(), c->sorted.size(), [](const MatchRec &r) { return r.id; },
                                   c->sort_tmp);
    c->sorted_valid = true;
    return FFBF_OK;
}

float elapsed(ffbf_ctx *c, int a, int b)
{
    if (!c->ev_recorded[a] || !c->ev_recorded[b]) return 0.0f;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0f;
    }
    return ms;
}

}  // namespace ffbf

using namespace ffbf;

extern "C" {

const char *ffbf_last_error(void) { return ffbf::g_err.c_str(); }

const char *ffbf_version(void) { return "ffbf-amd 0.1 (gfx950)"; }

}  // extern "C"

namespace {

// Context set-up shared by both methods: sp null for the pair method.
int create_ctx(const ffbf_params *params, const ffbf_split_params *sp, int device, ffbf_ctx **out)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return fail(FFBF_EDEVICE, "no HIP device available");
    }
    if (device < 0 || device >= ndev) return fail(FFBF_EINVAL, "device out of range");
    ffbf_ctx *c = new (std::nothrow) ffbf_ctx();
    if (!c) return fail(FFBF_ENOMEM, "context allocation");
    c->params = *params;
    c->device = device;
    // the sbox as the kernels read it: 256 words, or for the split method
    // 256 (lo, hi) pairs with bytes >= 0x80 mapped to entry 0 (ffbf.cpp:127-130)
    std::vector<uint32_t> sb(params->sbox, params->sbox + 256);
    if (sp) {
        c->split = true;
        c->split_params = *sp;
        c->smap.s = sp->s;
        c->smap.cache = 1u << sp->log2_cache_bits;
        c->smap.cmask = c->smap.cache - 1u;
        c->smap.emask = (uint32_t)((1ull << sp->log2_ext_bits) - 1u);
        sb.assign(512, 0u);
        for (int b = 0; b < 256; b++) {
            const uint64_t v = sp->sbox[b >= 0x80 ? 0 : b];
            sb[2 * b] = (uint32_t)v;
            sb[2 * b + 1] = (uint32_t)(v >> 32);
        }
    }
    int st = FFBF_OK;
    auto init = [&]() -> int {
        HIP_TRY(hipSetDevice(device));
        HIP_TRY(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device));
        HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (int i = 0; i < EV_N; i++) HIP_TRY(hipEventCreate(&c->ev[i]));
        if (sp) {
            c->mask = 0xFFFFFFFFu;  // unused: indices are placed by smap
            c->f_words = ((1ull << sp->log2_cache_bits) + (1ull << sp->log2_ext_bits)) / 32u;
        } else {
            c->mask = params->log2_bits >= 32 ? 0xFFFFFFFFu : ((1u << params->log2_bits) - 1u);
            c->f_words = 1ull << (params->log2_bits - 5);
        }
        RET_IF(c->F.ensure(c->f_words, "forward filter"));
        RET_IF(c->R.ensure(c->f_words, "reverse filter"));
        RET_IF(c->sbox.ensure(sb.size(), "sbox"));
        RET_IF(c->counters.ensure(ffbf::C_COUNT, "counters"));
        HIP_TRY(hipHostMalloc((void **)&c->h_counters, ffbf::C_COUNT * sizeof(unsigned long long),
                              hipHostMallocDefault));
        HIP_TRY(hipMemcpy(c->sbox.p, sb.data(), sb.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(ffbf::launch_fill32(c->F.p, c->f_words, 0u, c->stream));
        HIP_TRY(ffbf::launch_fill32(c->R.p, c->f_words, 0u, c->stream));
        RET_IF(clear_counters(c, 0, ffbf::C_COUNT));
        HIP_TRY(hipStreamSynchronize(c->stream));
        return FFBF_OK;
    };
    st = init();
    if (st != FFBF_OK) {
        std::string msg = g_err;
        ffbf_destroy(c);
        g_err = msg;
        return st;
    }
    *out = c;
    return FFBF_OK;
}

}  // namespace

extern "C" {

int ffbf_create(const ffbf_params *params, int device, ffbf_ctx **out)
{
    if (!params || !out) return fail(FFBF_EINVAL, "null argument");
    *out = nullptr;
    if (params->window < 1 || params->window > 64) return fail(FFBF_EINVAL, "window must be 1..64");
    if (params->k < 1 || params->k > 5) return fail(FFBF_EINVAL, "k must be 1..5");
    if (params->log2_bits < 10 || params->log2_bits > 32) return fail(FFBF_EINVAL, "log2_bits must be 10..32");
    if (params->flags != 0) return fail(FFBF_EINVAL, "flags must be 0");
    return create_ctx(params, nullptr, device, out);
}

int ffbf_create_split(const ffbf_split_params *sp, int device, ffbf_ctx **out)
{
    if (!sp || !out) return fail(FFBF_EINVAL, "null argument");
    *out = nullptr;
    if (sp->window < 1 || sp->window > 64) return fail(FFBF_EINVAL, "window must be 1..64");
    if (sp->k < 1 || sp->k > (uint32_t)ffbf::kSplitMaxK) return fail(FFBF_EINVAL, "k must be 1..11");
    if (sp->s > sp->k) return fail(FFBF_EINVAL, "s must be 0..k");
    if (sp->log2_cache_bits < 5 || sp->log2_cache_bits > 31 || sp->log2_ext_bits < 5 || sp->log2_ext_bits > 31)
        return fail(FFBF_EINVAL, "log2_cache_bits and log2_ext_bits must be 5..31");
    if (sp->flags != 0) return fail(FFBF_EINVAL, "flags must be 0");
    ffbf_params p{};
    p.window = sp->window;
    p.k = sp->k;
    return create_ctx(&p, sp, device, out);
}

void ffbf_destroy(ffbf_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    ingest_release(c);
    c->F.release();
    c->R.release();
    c->sbox.release();
    c->fold.release();
    c->gfold.release();
    c->counters.release();
    c->bytes.release();
    c->off.release();
    c->pat_hash.release();
    c->bin_staging.release();
    c->bin_offs.release();
    c->bitmap.release();
    c->table.release();
    c->chain_next.release();
    c->owned_text.release();
    c->cand.release();
    c->cstage.release();
    c->d_zero.release();
    c->d_zero_build.release();
    c->matches.release();
    c->sort_keys.release();
    c->sort_temp.release();
    c->q_bytes.release();
    c->q_hits.release();
    c->q_off.release();
    if (c->h_counters) (void)hipHostFree(c->h_counters);
    if (c->h_match) (void)hipHostFree(c->h_match);
    for (int i = 0; i < EV_N; i++)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    (void)hipGetLastError();
    delete c;
}

int ffbf_set_patterns(ffbf_ctx *c, const uint8_t *bytes, const uint64_t *offsets, uint64_t n)
{
    RET_IF(check_ctx(c));
    if (!offsets) return fail(FFBF_EINVAL, "null offsets");
    if (n > 0xFFFFFFFEull) return fail(FFBF_EINVAL, "at most 2^32-2 patterns");
    if (offsets[0] != 0) return fail(FFBF_EINVAL, "offsets[0] must be 0");
    for (uint64_t i = 0; i < n; i++)
        if (offsets[i + 1] < offsets[i]) return fail(FFBF_EINVAL, "offsets must be non-decreasing");
    const uint64_t total = offsets[n];
    if (total && !bytes) return fail(FFBF_EINVAL, "null pattern bytes");
    uint64_t n_long = 0;
    for (uint64_t i = 0; i < n; i++) n_long += (offsets[i + 1] - offsets[i] >= c->params.window) ? 1u : 0u;
    RET_IF(c->bytes.ensure(total + 64u, "pattern bytes"));  // padding: the build reads aligned 16-byte units
    RET_IF(c->off.ensure(n + 1, "pattern offsets"));
    RET_IF(c->pat_hash.ensure(n, "pattern hashes"));
    RET_IF(c->bitmap.ensure((n + 31) / 32, "survivor bitmap"));
    RET_IF(c->chain_next.ensure(n, "survivor chains"));
    const uint32_t tl = std::max<uint32_t>(pow2_at_least(2 * n), 10);
    c->table_size = 1u << tl;
    RET_IF(c->table.ensure(2ull * c->table_size, "survivor index"));
    if (total) HIP_TRY(hipMemcpyAsync(c->bytes.p, bytes, total, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->off.p, offsets, (n + 1) * 8, hipMemcpyHostToDevice, c->stream));
    // prefix hash pairs only (F untouched until ffbf_build)
    if (c->split)
        HIP_TRY(ffbf::launch_prefix_hash_split(c->bytes.p, c->off.p, n, c->sbox.p, c->params.window, c->pat_hash.p,
                                               c->stream));
    else
        HIP_TRY(ffbf::launch_prefix_hash(c->bytes.p, c->off.p, n, c->sbox.p, c->params.window, c->pat_hash.p,
                                         c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->n_pat = n;
    c->n_long = n_long;
    c->patterns = true;
    // F was built from the previous set: a scan needs ffbf_build (or a
    // filter load) for this one first
    c->built = false;
    c->generation++;
    c->scanned = c->passed = c->verified = false;
    return FFBF_OK;
}

int ffbf_set_patterns_lines(ffbf_ctx *c, const char *buf, uint64_t len)
{
    if (!buf && len) return fail(FFBF_EINVAL, "null buffer");
    std::vector<uint64_t> off;
    std::vector<uint8_t> bytes;
    bytes.reserve(len);
    off.push_back(0);
    uint64_t start = 0;
    for (uint64_t i = 0; i < len; i++) {
        if (buf[i] == '\n') {
            bytes.insert(bytes.end(), buf + start, buf + i);
            off.push_back(bytes.size());
            start = i + 1;
        }
    }
    if (start < len) {
        bytes.insert(bytes.end(), buf + start, buf + len);
        off.push_back(bytes.size());
    }
    return ffbf_set_patterns(c, bytes.data(), off.data(), off.size() - 1);
}

int ffbf_build(ffbf_ctx *c)
{
    RET_IF(check_ctx(c));
    if (!c->patterns) return fail(FFBF_ESTATE, "build before patterns are loaded");
    RET_IF(record(c, EV_B0));
    const uint64_t slices = (c->f_words + (1u << (ffbf::kBinSliceLog2 - 5)) - 1u) >> (ffbf::kBinSliceLog2 - 5);
    // device-scope atomics run at ~27 G/s whatever F's size: below 2^20 bit
    // sets (~39 us of them) the clear and the atomics beat the binned
    // build's two passes
    const bool binned = c->build_method == FFBF_BUILD_BINNED ||
                        (c->build_method == FFBF_BUILD_AUTO && slices >= 2 &&
                         c->n_long * c->params.k >= (1ull << 20));
    // the scan mode and its folds, from the analytic fill (a rejected mode
    // request leaves the previous build in place)
    c->f_set_bits = -1;
    RET_IF(plan_mode(c));
    if (binned) {
        // bits sorted by 2^20-bit slice and set in LDS, each slice written
        // whole (no atomics, no clear), from the prefix hash pairs
        // ffbf_set_patterns computed
        const uint64_t nb = ffbf::build_bin_blocks(c->n_pat, c->params.k);
        RET_IF(c->bin_staging.ensure(std::max<uint64_t>(nb, 1) * ffbf::kBinEntries, "build staging"));
        RET_IF(c->bin_offs.ensure(std::max<uint64_t>(nb, 1) * (slices + 1u), "build slice offsets"));
        HIP_TRY(ffbf::launch_build_binned(c->off.p, c->pat_hash.p, c->n_pat, c->params.window, c->params.k, c->mask,
                                          c->split, c->smap, c->F.p, c->f_words, c->bin_staging.p, c->bin_offs.p,
                                          c->stream));
        RET_IF(run_fold_jobs(c));
    } else {
        // few bits: the atomics also set the folds' bits when that adds
        // fewer atomics than a fold launch costs (two launches fewer: no
        // separate clear of F, no fold pass)
        ffbf::BuildFolds fo{};
        const bool fused = c->n_fold_jobs > 0 && c->n_long * c->params.k * c->n_fold_jobs <= kFusedFoldAtomics;
        // one clear for F, the folds the atomics set, and the scan state (the
        // next scan then skips its own clear)
        std::vector<ffbf::ZeroSpan> zs;
        add_spans(zs, c->F.p, c->f_words);
        add_scan_spans(c, zs);
        if (fused) {
            // every fold is of F's first src_words words (a fold of a fold of
            // F, both powers of two, is the fold of F)
            fo.src_bits = c->fold_jobs[0].src_words * 32u;
            for (uint32_t i = 0; i < c->n_fold_jobs; i++) {
                fo.p[i] = c->fold_jobs[i].dst;
                fo.wmask[i] = c->fold_jobs[i].dst_words - 1u;
                add_spans(zs, c->fold_jobs[i].dst, c->fold_jobs[i].dst_words);
            }
        }
        RET_IF(upload_spans(c, zs, c->d_zero_build, c->build_zero_host));
        HIP_TRY(ffbf::launch_zero_spans(c->d_zero_build.p, (uint32_t)zs.size(), c->n_cu, c->stream));
        c->scan_state_clean = true;
        c->clean_key[0] = c->cstage.counts.p;
        c->clean_size[0] = c->cstage.n_regions;
        c->clean_key[1] = c->table.p;
        c->clean_size[1] = c->table_size;
        HIP_TRY(ffbf::launch_build_atomic(c->off.p, c->pat_hash.p, c->n_pat, c->params.window, c->params.k, c->mask,
                                          c->split, c->smap, c->F.p, fo, c->stream));
        if (!fused) RET_IF(run_fold_jobs(c));
    }
    RET_IF(record(c, EV_B1));
    c->built = true;
    c->scanned = c->passed = c->verified = false;
    return FFBF_OK;
}

int ffbf_set_timing(ffbf_ctx *c, int on)
{
    RET_IF(check_ctx(c));
    c->timing = on != 0;
    return FFBF_OK;
}

int ffbf_set_build_method(ffbf_ctx *c, int method)
{
    RET_IF(check_ctx(c));
    if (method < FFBF_BUILD_AUTO || method > FFBF_BUILD_BINNED) return fail(FFBF_EINVAL, "build method must be 0..2");
    c->build_method = method;
    return FFBF_OK;
}

int ffbf_set_scan_mode(ffbf_ctx *c, int mode, uint32_t fold_log2_bits)
{
    RET_IF(check_ctx(c));
    if (mode < -1 || mode > 4) return fail(FFBF_EINVAL, "mode must be -1..4");
    const int old_mode = c->mode_req;
    const uint32_t old_fold = c->fold_log2_req;
    c->mode_req = mode;
    c->fold_log2_req = fold_log2_bits;
    if (c->built) {
        const int st = choose_mode(c);
        if (st != FFBF_OK) {  // a rejected request leaves the previous one in force
            c->mode_req = old_mode;
            c->fold_log2_req = old_fold;
            return st;
        }
    }
    return FFBF_OK;
}

int ffbf_set_scan_tuning(ffbf_ctx *c, uint32_t probes, uint32_t block)
{
    RET_IF(check_ctx(c));
    if (probes > c->params.k) return fail(FFBF_EINVAL, "probes must be 0..k");
    if (block != 0 && block != 256 && block != 512 && block != 1024)
        return fail(FFBF_EINVAL, "block must be 0, 256, 512 or 1024");
    const uint32_t old_probes = c->probes_req, old_block = c->block_req;
    c->probes_req = probes;
    c->block_req = block;
    if (c->built) {
        const int st = choose_mode(c);
        if (st != FFBF_OK) {
            c->probes_req = old_probes;
            c->block_req = old_block;
            return st;
        }
    }
    return FFBF_OK;
}

int ffbf_scan_device(ffbf_ctx *c, const uint8_t *dev_text, uint64_t n)
{
    RET_IF(check_ctx(c));
    RET_IF(set_text(c, dev_text, n));
    return do_scan(c);
}

int ffbf_scan(ffbf_ctx *c, const uint8_t *text, uint64_t n)
{
    RET_IF(check_ctx(c));
    if (!text && n) return fail(FFBF_EINVAL, "null text");
    if (n > c->ingest_chunk) return scan_host(c, text, n);
    // one piece: nothing to overlap
    RET_IF(c->owned_text.ensure(std::max<uint64_t>(n, 1), "text"));
    if (n) HIP_TRY(hipMemcpyAsync(c->owned_text.p, text, n, hipMemcpyHostToDevice, c->stream));
    RET_IF(set_text(c, c->owned_text.p, n));
    return do_scan(c);
}

int ffbf_pattern_pass(ffbf_ctx *c, uint64_t *n_survivors)
{
    RET_IF(check_ctx(c));
    if (!c->scanned) return fail(FFBF_ESTATE, "pattern pass before scan");
    RET_IF(enqueue_pass(c, false));
    RET_IF(sync_counters(c));
    c->n_surv = c->h_counters[ffbf::C_SURV];
    c->passed = true;
    c->verified = false;
    if (n_survivors) *n_survivors = c->n_surv;
    return FFBF_OK;
}

int ffbf_verify(ffbf_ctx *c, uint64_t *n_matches)
{
    RET_IF(check_ctx(c));
    if (!c->passed) return fail(FFBF_ESTATE, "verify before pattern pass");
    RET_IF(do_verify(c));
    if (n_matches) *n_matches = c->n_match;
    return FFBF_OK;
}

int ffbf_run_device(ffbf_ctx *c, const uint8_t *dev_text, uint64_t n, uint64_t *n_matches)
{
    RET_IF(check_ctx(c));
    RET_IF(set_text(c, dev_text, n));
    RET_IF(enqueue_scan(c, true));
    RET_IF(enqueue_pass(c, true));
    RET_IF(enqueue_verify(c, true));
    RET_IF(prefetch_matches(c));
    RET_IF(sync_counters(c));
    if (c->h_counters[ffbf::C_CAND] > c->cand.cap) {
        // rerun the three stages with room for every candidate
        RET_IF(ensure_cand(c, 2 * c->h_counters[ffbf::C_CAND]));
        RET_IF(enqueue_scan(c, true));
        RET_IF(enqueue_pass(c, true));
        RET_IF(enqueue_verify(c, true));
        RET_IF(prefetch_matches(c));
        RET_IF(sync_counters(c));
    }
    c->n_cand = c->h_counters[ffbf::C_CAND];
    c->n_surv = c->h_counters[ffbf::C_SURV];
    c->scanned = c->passed = true;
    if (c->h_counters[ffbf::C_MATCH] > c->matches.cap) {
        RET_IF(c->matches.ensure(std::max<uint64_t>(2 * c->h_counters[ffbf::C_MATCH], 2 * c->matches.cap), "matches"));
        RET_IF(enqueue_verify(c, false));
        RET_IF(prefetch_matches(c));
        RET_IF(sync_counters(c));
    }
    c->n_match = c->h_counters[ffbf::C_MATCH];
    c->verified = true;
    c->sorted_valid = false;
    RET_IF(sort_on_device(c));
    if (n_matches) *n_matches = c->n_match;
    return FFBF_OK;
}

int ffbf_get_matches(ffbf_ctx *c, ffbf_match *out, uint64_t cap, uint64_t *n)
{
    RET_IF(check_ctx(c));
    if (!c->verified) return fail(FFBF_ESTATE, "no verified matches");
    if (n) *n = c->n_match;
    if (cap < c->n_match) return fail(FFBF_ECAPACITY, "match buffer too small");
    if (c->n_match && !out) return fail(FFBF_EINVAL, "null output");
    RET_IF(load_sorted_matches(c));
    if (c->n_match) std::memcpy(out, c->sorted.data(), c->n_match * sizeof(ffbf_match));
    return FFBF_OK;
}

int ffbf_match_text(ffbf_ctx *c, const uint8_t *text, uint64_t n, ffbf_match **out, uint64_t *n_out)
{
    RET_IF(check_ctx(c));
    if (!out || !n_out) return fail(FFBF_EINVAL, "null output");
    *out = nullptr;
    *n_out = 0;
    if (!text && n) return fail(FFBF_EINVAL, "null text");
    uint64_t nm = 0;
    if (n <= c->ingest_chunk) {
        // one piece: a single copy and the three stages back to back, one sync
        RET_IF(c->owned_text.ensure(std::max<uint64_t>(n, 1), "text"));
        if (n) HIP_TRY(hipMemcpyAsync(c->owned_text.p, text, n, hipMemcpyHostToDevice, c->stream));
        RET_IF(ffbf_run_device(c, c->owned_text.p, n, &nm));
    } else {
        RET_IF(scan_host(c, text, n));
        RET_IF(ffbf_pattern_pass(c, nullptr));
        RET_IF(ffbf_verify(c, &nm));
    }
    RET_IF(load_sorted_matches(c));
    ffbf_match *m = (ffbf_match *)std::malloc(std::max<uint64_t>(nm, 1) * sizeof(ffbf_match));
    if (!m) return fail(FFBF_ENOMEM, "match list allocation");
    if (nm) std::memcpy(m, c->sorted.data(), nm * sizeof(ffbf_match));
    *out = m;
    *n_out = nm;
    return FFBF_OK;
}

void ffbf_free_matches(ffbf_match *m) { std::free(m); }

int ffbf_query_lines(ffbf_ctx *c, const uint8_t *bytes, const uint64_t *offsets, uint64_t n_lines, uint8_t *hits)
{
    RET_IF(check_ctx(c));
    if (!c->built) return fail(FFBF_ESTATE, "query before build (or filter load)");
    if (n_lines == 0) return FFBF_OK;
    if (!offsets || !hits) return fail(FFBF_EINVAL, "null offsets or hits");
    if (offsets[0] != 0) return fail(FFBF_EINVAL, "offsets[0] must be 0");
    for (uint64_t i = 0; i < n_lines; i++)
        if (offsets[i + 1] < offsets[i]) return fail(FFBF_EINVAL, "offsets must be non-decreasing");
    const uint64_t total = offsets[n_lines];
    if (total && !bytes) return fail(FFBF_EINVAL, "null line bytes");
    RET_IF(c->q_bytes.ensure(std::max<uint64_t>(total, 1), "query lines"));
    RET_IF(c->q_off.ensure(n_lines + 1, "query offsets"));
    RET_IF(c->q_hits.ensure(n_lines, "query answers"));
    if (total) HIP_TRY(hipMemcpyAsync(c->q_bytes.p, bytes, total, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->q_off.p, offsets, (n_lines + 1) * 8, hipMemcpyHostToDevice, c->stream));
    ffbf::QueryArgs a{};
    a.bytes = c->q_bytes.p;
    a.off = c->q_off.p;
    a.n_lines = n_lines;
    a.F = c->F.p;
    a.sbox = c->sbox.p;
    a.W = c->params.window;
    a.k = c->params.k;
    a.mask = c->mask;
    a.method = c->split ? ffbf::METHOD_SPLIT : ffbf::METHOD_PAIR;
    a.split = c->smap;
    a.hits = c->q_hits.p;
    HIP_TRY(ffbf::launch_query_lines(a, c->n_cu, c->stream));
    HIP_TRY(hipMemcpyAsync(hits, c->q_hits.p, n_lines, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FFBF_OK;
}

int ffbf_query_line(ffbf_ctx *c, const uint8_t *line, uint64_t len, int *hit)
{
    if (!hit) return fail(FFBF_EINVAL, "null output");
    *hit = 0;
    if (!line && len) return fail(FFBF_EINVAL, "null line");
    const uint64_t off[2] = {0, len};
    uint8_t h = 0;
    RET_IF(ffbf_query_lines(c, line, off, 1, &h));
    *hit = h;
    return FFBF_OK;
}

int ffbf_get_filter(ffbf_ctx *c, int which, uint32_t *out, uint64_t n_words)
{
    RET_IF(check_ctx(c));
    if (!out) return fail(FFBF_EINVAL, "null output");
    const uint32_t *src = nullptr;
    uint64_t words = c->f_words;
    if (which == FFBF_FILTER_FORWARD) {
        src = c->F.p;
    } else if (which == FFBF_FILTER_REVERSE) {
        src = c->R.p;
    } else if (which == FFBF_FILTER_FOLD || which == FFBF_FILTER_L2_FOLD) {
        if (!c->built) return fail(FFBF_ESTATE, "no filter built or loaded");
        const bool lds = which == FFBF_FILTER_FOLD;
        if (lds ? (c->mode != ffbf::MODE_FOLD && c->mode != ffbf::MODE_HYBRID) : c->gfold_words == 0)
            return fail(FFBF_ESTATE, "the scan mode in force has no such fold");
        src = lds ? c->fold.p : c->gfold.p;
        words = lds ? c->pre_words : c->gfold_words;
    } else {
        return fail(FFBF_EINVAL, "which must be one of FFBF_FILTER_FORWARD .. FFBF_FILTER_L2_FOLD");
    }
    if (n_words < words) return fail(FFBF_ECAPACITY, "filter buffer too small");
    HIP_TRY(hipMemcpyAsync(out, src, words * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FFBF_OK;
}

int ffbf_get_candidates(ffbf_ctx *c, uint64_t *out, uint64_t cap, uint64_t *n)
{
    RET_IF(check_ctx(c));
    if (!c->scanned) return fail(FFBF_ESTATE, "no scan yet");
    if (n) *n = c->n_cand;
    if (cap < c->n_cand) return fail(FFBF_ECAPACITY, "candidate buffer too small");
    if (c->n_cand && !out) return fail(FFBF_EINVAL, "null output");
    std::vector<Cand> h(c->n_cand);
    if (c->n_cand) {
        HIP_TRY(hipMemcpyAsync(h.data(), c->cand.p, c->n_cand * sizeof(Cand), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    for (uint64_t i = 0; i < c->n_cand; i++) out[i] = h[i].offset;
    std::sort(out, out + c->n_cand);
    return FFBF_OK;
}

int ffbf_get_survivors(ffbf_ctx *c, uint32_t *bitmap, uint64_t n_words, uint64_t *n_survivors)
{
    RET_IF(check_ctx(c));
    if (!c->passed) return fail(FFBF_ESTATE, "no pattern pass yet");
    if (n_survivors) *n_survivors = c->n_surv;
    const uint64_t need = (c->n_pat + 31) / 32;
    if (n_words < need) return fail(FFBF_ECAPACITY, "bitmap buffer too small");
    if (need && !bitmap) return fail(FFBF_EINVAL, "null output");
    if (need) {
        HIP_TRY(hipMemcpyAsync(bitmap, c->bitmap.p, need * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return FFBF_OK;
}

int ffbf_filter_save(ffbf_ctx *c, uint32_t *deltas, uint64_t cap, uint64_t *n)
{
    RET_IF(check_ctx(c));
    if (!c->built) return fail(FFBF_ESTATE, "no filter built");
    std::vector<uint32_t> h(c->f_words);
    HIP_TRY(hipMemcpyAsync(h.data(), c->F.p, c->f_words * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    uint64_t cnt = 0;
    for (uint32_t w : h) cnt += (uint64_t)__builtin_popcount(w);
    if (n) *n = cnt;
    if (cap < cnt) return fail(FFBF_ECAPACITY, "delta buffer too small");
    if (cnt && !deltas) return fail(FFBF_EINVAL, "null output");
    uint32_t prev = 0;
    uint64_t o = 0;
    for (uint64_t w = 0; w < c->f_words; w++) {
        for (uint32_t m = h[w]; m; m &= m - 1) {
            const uint32_t bit = (uint32_t)(w * 32u) + (uint32_t)__builtin_ctz(m);
            deltas[o++] = bit - prev;
            prev = bit;
        }
    }
    return FFBF_OK;
}

int ffbf_filter_load(ffbf_ctx *c, const uint32_t *deltas, uint64_t n)
{
    RET_IF(check_ctx(c));
    if (!deltas && n) return fail(FFBF_EINVAL, "null deltas");
    std::vector<uint32_t> h(c->f_words, 0u);
    uint32_t prev = 0;
    for (uint64_t i = 0; i < n; i++) {
        prev += deltas[i];
        if ((uint64_t)prev >= c->f_words * 32u) return fail(FFBF_EINVAL, "bit index beyond the filter size");
        h[prev >> 5] |= 1u << (prev & 31u);
    }
    HIP_TRY(hipMemcpyAsync(c->F.p, h.data(), c->f_words * 4, hipMemcpyHostToDevice, c->stream));
    uint64_t ones = 0;
    for (uint32_t w : h) ones += (uint64_t)__builtin_popcount(w);
    c->f_set_bits = (int64_t)ones;
    RET_IF(choose_mode(c));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->built = true;
    c->scanned = c->passed = c->verified = false;
    return FFBF_OK;
}

int ffbf_get_timings(ffbf_ctx *c, ffbf_timings *out)
{
    RET_IF(check_ctx(c));
    if (!out) return fail(FFBF_EINVAL, "null output");
    HIP_TRY(hipStreamSynchronize(c->stream));
    out->build_ms = elapsed(c, EV_B0, EV_B1);
    out->scan_ms = elapsed(c, EV_S0, EV_S3);
    out->scan_kernel_ms = elapsed(c, EV_S1, EV_S2);
    out->pattern_pass_ms = elapsed(c, EV_P0, EV_P1);
    out->verify_ms = elapsed(c, EV_V0, EV_V1);
    out->scan_mode = c->mode;
    out->fold_log2_bits = c->mode == ffbf::MODE_GLOBAL ? 0u : pow2_at_least((uint64_t)c->pre_words * 32u);
    out->probes = ffbf::mode_in_lds(c->mode) ? std::min(c->probes, 3u) : 1u;
    out->block = scan_block(c);
    out->ingest_ms = (float)c->ingest_ms;
    out->l2_fold_log2_bits = c->gfold_words ? pow2_at_least((uint64_t)c->gfold_words * 32u) : 0u;
    return FFBF_OK;
}

int ffbf_synchronize(ffbf_ctx *c)
{
    RET_IF(check_ctx(c));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FFBF_OK;
}

void *ffbf_get_stream(ffbf_ctx *c) { return c ? (void *)c->stream : nullptr; }

int ffbf_device_alloc(int device, uint64_t bytes, void **out)
{
    if (!out) return fail(FFBF_EINVAL, "null output");
    *out = nullptr;
    HIP_TRY(hipSetDevice(device));
    hipError_t e = hipMalloc(out, std::max<uint64_t>(bytes, 1));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(FFBF_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return FFBF_OK;
}

int ffbf_device_free(void *p)
{
    if (p) HIP_TRY(hipFree(p));
    return FFBF_OK;
}

int ffbf_device_upload(void *dst, const void *src, uint64_t bytes)
{
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return FFBF_OK;
}

int ffbf_device_download(void *dst, const void *src, uint64_t bytes)
{
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return FFBF_OK;
}

int ffbf_device_copy(void *dst, const void *src, uint64_t bytes)
{
    if ((bytes % 16u) || ((uintptr_t)dst % 16u) || ((uintptr_t)src % 16u))
        return fail(FFBF_EINVAL, "device copy needs 16-byte aligned pointers and size");
    HIP_TRY(ffbf::launch_copy(dst, src, bytes, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    return FFBF_OK;
}

int ffbf_host_alloc(uint64_t bytes, void **out)
{
    if (!out) return fail(FFBF_EINVAL, "null output");
    *out = nullptr;
    hipError_t e = hipHostMalloc(out, std::max<uint64_t>(bytes, 1), hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(FFBF_ENOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    return FFBF_OK;
}

int ffbf_host_free(void *ptr)
{
    if (ptr) HIP_TRY(hipHostFree(ptr));
    return FFBF_OK;
}

int ffbf_device_count(int *out)
{
    if (!out) return fail(FFBF_EINVAL, "null output");
    *out = 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *out = n;
    return FFBF_OK;
}

namespace {

// FFBF_PROBE_UPLOAD: 64 pieces of a pinned buffer, piece i starting i mod 16
// bytes into its 1/64 share and 16 bytes shorter than it, uploaded back to
// back into a device buffer by k_upload, once to warm and `steps` times timed.
int upload_bench(uint32_t log2_bytes, uint32_t steps, float *ms, double *bytes_per_s)
{
    if (log2_bytes < 16 || log2_bytes > 31) return fail(FFBF_EINVAL, "upload bench: log2_bytes must be 16..31");
    const uint64_t bytes = 1ull << log2_bytes, share = bytes / ffbf::kUploadMax;
    uint8_t *h = nullptr, *d = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto done = [&](int st) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (h) (void)hipHostFree(h);
        if (d) (void)hipFree(d);
        return st;
    };
    if (hipHostMalloc((void **)&h, bytes, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&d, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return done(fail(FFBF_ENOMEM, "upload bench buffers"));
    }
    std::memset(h, 0x5A, bytes);
    ffbf::UploadArgs ua{};
    ua.dst = d;
    uint64_t at = 0, total = 0;
    for (uint32_t i = 0; i < ffbf::kUploadMax; i++) {
        ua.src[i] = (uint64_t)(uintptr_t)(h + (uint64_t)i * share + (i & 15u));
        ua.len[i] = share - 16u;
        ua.dst_off[i] = at;
        ua.seg_start[i] = ua.total_segs;
        ua.total_segs += ffbf::upload_segs(at, ua.len[i]);
        at += ua.len[i];
        total += ua.len[i];
    }
    ua.n = ffbf::kUploadMax;
    ua.seg_start[ua.n] = ua.total_segs;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return done(fail(FFBF_EDEVICE, "event creation"));
    bool ok = ffbf::launch_upload(ua, nullptr) == hipSuccess && hipEventRecord(e0, nullptr) == hipSuccess;
    for (uint32_t s = 0; ok && s < steps; s++) ok = ffbf::launch_upload(ua, nullptr) == hipSuccess;
    ok = ok && hipEventRecord(e1, nullptr) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    if (!ok) return done(fail(FFBF_EDEVICE, std::string("upload bench: ") + hipGetErrorString(hipGetLastError())));
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t / (float)steps;
    if (bytes_per_s) *bytes_per_s = (double)total * (double)steps / ((double)t * 1e-3);
    return done(FFBF_OK);
}

}  // namespace

int ffbf_probe_bench(int device, uint32_t log2_bytes, uint32_t steps, uint32_t kind, float *ms, double *ops_per_s)
{
    if (log2_bytes < 4 || log2_bytes > 34 || steps == 0) return fail(FFBF_EINVAL, "log2_bytes must be 4..34, steps > 0");
    if (kind > FFBF_PROBE_UPLOAD) return fail(FFBF_EINVAL, "unknown probe kind");
    if (kind == FFBF_PROBE_UPLOAD) {
        HIP_TRY(hipSetDevice(device));
        return upload_bench(log2_bytes, steps, ms, ops_per_s);
    }
    HIP_TRY(hipSetDevice(device));
    int n_cu = 0;
    HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device));
    const uint32_t blocks = (uint32_t)n_cu * 8u;
    DevBuf<uint32_t> arr, out;
    RET_IF(arr.ensure(1ull << (log2_bytes - 2), "gather array"));
    RET_IF(out.ensure((uint64_t)blocks * 256u, "gather output"));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto done = [&](int st) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        arr.release();
        out.release();
        return st;
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return done(fail(FFBF_EDEVICE, "event creation"));
    if (ffbf::launch_fill32(arr.p, 1ull << (log2_bytes - 2), 0x5A5A5A5Au, nullptr) != hipSuccess ||
        ffbf::launch_gather(arr.p, log2_bytes - 2, steps, out.p, blocks, kind, nullptr) != hipSuccess ||
        hipEventRecord(e0, nullptr) != hipSuccess ||
        ffbf::launch_gather(arr.p, log2_bytes - 2, steps, out.p, blocks, kind, nullptr) != hipSuccess ||
        hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
        return done(fail(FFBF_EDEVICE, std::string("probe bench: ") + hipGetErrorString(hipGetLastError())));
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t;
    // random kinds: accesses per second; the streaming pass: bytes read and
    // written per second
    const double ops = kind == FFBF_PROBE_STREAM ? 2.0 * (double)(1ull << log2_bytes) * (double)steps
                                                 : (double)blocks * 256.0 * 16.0 * (double)steps;
    if (ops_per_s) *ops_per_s = ops / ((double)t * 1e-3);
    return done(FFBF_OK);
}

int ffbf_gather_bench(int device, uint32_t log2_bytes, uint32_t steps, float *ms, double *probes_per_s)
{
    return ffbf_probe_bench(device, log2_bytes, steps, FFBF_PROBE_LOAD, ms, probes_per_s);
}

}  // extern "C"
