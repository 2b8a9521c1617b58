// ffbf_ctx.h -- the matcher context behind the C ABI and the helpers its
// implementation files share (internal).
#pragma once

#include <algorithm>
#include <string>
#include <vector>

#include "ffbf.h"
#include "ffbf_device.h"

namespace ffbf {

// Records msg as the calling thread's last error; returns code.
int fail(int code, const std::string &msg);

#define HIP_TRY(expr)                                                                                      \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return ::ffbf::fail(FFBF_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define RET_IF(st)                      \
    do {                                \
        int st_ = (st);                 \
        if (st_ != FFBF_OK) return st_; \
    } while (0)

// A device array that only grows.
template <class T>
struct DevBuf {
    T *p = nullptr;
    uint64_t cap = 0;  // elements

    int ensure(uint64_t n, const char *what)
    {
        if (n <= cap && p) return FFBF_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const uint64_t want = std::max<uint64_t>(n, 1);
        hipError_t e = hipMalloc((void **)&p, want * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(FFBF_ENOMEM, std::string("hipMalloc ") + what + " (" + std::to_string(want * sizeof(T)) +
                                         " bytes): " + hipGetErrorString(e));
        }
        cap = want;
        return FFBF_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Stage events: build B0-B1; scan S0 (clears) S1 (kernel) S2 (compaction) S3;
// pattern pass P0-P1; verification V0-V1.
enum Ev { EV_B0, EV_B1, EV_S0, EV_S1, EV_S2, EV_S3, EV_P0, EV_P1, EV_V0, EV_V1, EV_N };

// Where a scan launch stages its candidates before they are compacted into
// the candidate list: one region per wave slot plus a spill area (ScanArgs),
// the regions' fill counts and the compaction's offsets.  Sized from the
// list's capacity; the counts are zeroed before every scan.
struct CandStage {
    DevBuf<Cand> raw;
    DevBuf<uint32_t> counts;
    DevBuf<uint64_t> offsets;
    uint32_t n_regions = 0, region = 0;
    uint64_t list_cap = 0;  // list capacity the layout was made for

    int ensure(int n_cu, uint64_t cap)
    {
        if (raw.p && cap == list_cap) return FFBF_OK;
        const uint64_t list_cap_new = cap;
        const uint32_t nr = (uint32_t)n_cu * kRegionWavesPerCU;
        // a launch uses at most half the slots: regions alone then hold a
        // full list; the spill area takes half a list more
        const uint32_t reg = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(4u, 2u * list_cap_new / nr), 4096u);
        RET_IF(raw.ensure((uint64_t)nr * reg + list_cap_new / 2u + 1u, "candidate staging"));
        RET_IF(counts.ensure(nr, "candidate region counts"));
        RET_IF(offsets.ensure((uint64_t)nr + 2u, "candidate offsets"));
        n_regions = nr;
        region = reg;
        list_cap = list_cap_new;
        return FFBF_OK;
    }
    // the staging fields of a scan launch whose waves use slots [wave0, end)
    void fill(ScanArgs &a, uint32_t wave0, uint32_t end) const
    {
        a.cand = raw.p;
        a.cand_cap = raw.cap;
        a.wave_counts = counts.p;
        a.region = region;
        a.wave0 = wave0;
        a.region_end = end;
        a.spill_base = (uint64_t)n_regions * region;
    }
    void release()
    {
        raw.release();
        counts.release();
        offsets.release();
        n_regions = region = 0;
        list_cap = 0;
    }
};

uint32_t pow2_at_least(uint64_t n);

}  // namespace ffbf

struct ffbf_ctx {
    ffbf_params params{};  // window and k hold for both methods
    // the CPU binary's split filter (ffbf_create_split)
    bool split = false;
    ffbf_split_params split_params{};
    ffbf::SplitMap smap{};
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    uint32_t mask = 0;
    uint64_t f_words = 0;

    ffbf::DevBuf<uint32_t> F, R, sbox, fold, gfold;
    ffbf::DevBuf<unsigned long long> counters;
    unsigned long long *h_counters = nullptr;  // pinned

    // bumped whenever F, the pattern set or the scan mode changes: a batch
    // that borrows the context rebuilds its descriptors when it sees a new value
    uint64_t generation = 0;

    // scan mode and tuning
    int mode_req = -1;
    uint32_t fold_log2_req = 0;
    uint32_t probes_req = 0, block_req = 0;
    int build_method = FFBF_BUILD_AUTO;
    uint32_t mode = ffbf::MODE_GLOBAL;
    uint32_t pre_words = 0;   // prefilter size in words (F itself, LDS fold or global fold)
    uint32_t gfold_words = 0; // size of the L2-resident fold behind the LDS fold (mode 4)
    // The OR-folds the chosen mode needs, in order: each folds
    // src[0..src_words) onto dst[0..dst_words) (powers of two), once F is
    // complete -- or the atomic build sets their bits beside F's.
    struct FoldJob {
        const uint32_t *src;
        uint64_t src_words;
        uint32_t *dst;
        uint32_t dst_words;
    };
    FoldJob fold_jobs[2] = {};
    uint32_t n_fold_jobs = 0;
    // Set when the atomic build's clear also covered the scan state (R, the
    // region counts, the scan's counters, the survivor index) of the buffers
    // in clean_key / clean_size: the next scan need not clear it again.
    // Whatever writes that state (a scan, pattern pass, verification, an
    // allocation-free run, a batch over this filter) clears the flag.
    bool scan_state_clean = false;
    // The fused run path leaves the candidate gather of its scan to the
    // pattern pass's launch (k_pass_gather): set between the two.
    bool gather_pending = false;
    ffbf::GatherArgs pending_gather{};
    const void *clean_key[2] = {nullptr, nullptr};  // region counts and survivor index cleared
    uint64_t clean_size[2] = {0, 0};
    // the atomic build's clear list (F, the folds it sets, the scan state), as last uploaded
    std::vector<ffbf::ZeroSpan> build_zero_host;
    ffbf::DevBuf<ffbf::ZeroSpan> d_zero_build;
    uint32_t probes = 1;      // LDS prefilter bits tested before queueing
    double pre_fill = 0.0;    // expected fraction of prefilter bits set
    uint64_t n_long = 0;      // patterns with length >= W (each sets k bits of F)
    int64_t f_set_bits = -1;  // exact number of set bits of F when known (filter load)

    // pattern set
    ffbf::DevBuf<uint8_t> bytes;
    ffbf::DevBuf<uint64_t> off;
    ffbf::DevBuf<uint2> pat_hash;
    // the atomic-free build's scratch: bit indices sorted by slice per block,
    // and each block's slice offsets
    ffbf::DevBuf<uint32_t> bin_staging, bin_offs;
    ffbf::DevBuf<uint32_t> bitmap, table, chain_next;
    uint64_t n_pat = 0;
    uint32_t table_size = 0;

    // scan / results
    const uint8_t *text = nullptr;
    uint64_t text_n = 0;
    ffbf::DevBuf<uint8_t> owned_text;
    ffbf::DevBuf<ffbf::Cand> cand;  // compacted candidate list
    ffbf::CandStage cstage;         // per-launch staging before compaction
    // everything a scan clears first (R, region counts, scan counters,
    // survivor index), cleared in one launch; rebuilt when a buffer moves
    ffbf::DevBuf<ffbf::ZeroSpan> d_zero;
    uint32_t n_zero = 0;
    const void *zero_key[2] = {nullptr, nullptr};  // region counts and index the list was made for
    uint32_t zero_table_size = 0;
    ffbf::DevBuf<ffbf::MatchRec> matches;
    ffbf::DevBuf<uint64_t> sort_keys;  // 2 x capacity
    ffbf::DevBuf<uint8_t> sort_temp;
    bool sorted_on_device = false;
    uint64_t n_cand = 0, n_surv = 0, n_match = 0;
    std::vector<ffbf::MatchRec> sorted;
    std::vector<ffbf::MatchRec> sort_tmp;  // the host sort's second buffer
    bool sorted_valid = false;
    // pinned landing area of the match list on its way to the host: the
    // fused run copies the first kMatchPrefetch records there before its one
    // synchronisation, so a short list needs no second round trip
    ffbf::MatchRec *h_match = nullptr;
    uint64_t h_match_cap = 0;  // records
    uint64_t prefetched = 0;   // leading records of the current list already in h_match

    bool patterns = false, built = false, scanned = false, passed = false, verified = false;
    hipEvent_t ev[ffbf::EV_N] = {};
    bool ev_recorded[ffbf::EV_N] = {};
    bool timing = true;  // record the stage events (ffbf_set_timing)

    // line queries (ffbf_query_lines): the lines, their offsets and answers
    ffbf::DevBuf<uint8_t> q_bytes, q_hits;
    ffbf::DevBuf<uint64_t> q_off;

    // text ingest from host memory and files (ffbf_ingest.cpp)
    hipStream_t copy_stream = nullptr;
    uint8_t *staging[4] = {};         // pinned pieces of the text on their way to the device
    hipEvent_t ev_staged[4] = {};     // H2D copy of each staging buffer done
    bool staged_pending[4] = {};
    uint64_t staging_bytes = 0;       // size of each allocated staging buffer
    uint32_t staging_depth = 0;       // staging buffers allocated
    uint64_t ingest_chunk = 64ull << 20;
    uint32_t ingest_depth = 4;
    uint32_t ingest_threads = 8;      // host threads copying or reading a piece
    double ingest_ms = 0.0;           // last ingest, host clock, first byte to last candidate
};

namespace ffbf {
// Context internals shared by the implementation files (ffbf_capi.cpp).
int check_ctx(ffbf_ctx *c);
int set_text(ffbf_ctx *c, const uint8_t *dev_text, uint64_t n);
int record(ffbf_ctx *c, int e);
int sync_counters(ffbf_ctx *c);
int enqueue_scan_begin(ffbf_ctx *c);
// defer_gather: the candidate gather is left to the next enqueue_pass's launch.
int enqueue_scan_range(ffbf_ctx *c, const uint8_t *text, uint64_t n, uint64_t base, bool mark_end,
                       bool defer_gather = false);
int do_scan(ffbf_ctx *c);
// Grows the candidate list (and its staging) to hold at least n entries.
int ensure_cand(ffbf_ctx *c, uint64_t n);
// Text ingest (ffbf_ingest.cpp): host memory through the staging pipeline,
// and release of the pipeline's buffers and stream.
int scan_host(ffbf_ctx *c, const uint8_t *text, uint64_t n);
void ingest_release_staging(ffbf_ctx *c);
void ingest_release(ffbf_ctx *c);
// The stream a batch object queues its work on (ffbf_batch.cpp).
hipStream_t batch_stream(ffbf_batch *b);
// Sizes a batch's descriptor, chunk and candidate buffers for batches of up
// to max_texts texts and max_bytes bytes, so that a stream of such batches
// never reallocates (each move costs a device-wide hipMalloc and a graph
// re-capture).
int batch_reserve(ffbf_batch *b, uint32_t max_texts, uint64_t max_bytes);
// Stage launchers shared with the batch path (ffbf_capi.cpp).
int enqueue_pass(ffbf_ctx *c, bool fresh);
// The filter fields of a scan launch over context c (both hash methods):
// filters, sbox, prefilter and its geometry, probes; not the text, staging
// or counters.
void filter_scan_args(const ffbf_ctx *c, ScanArgs &a);

// An LDS fold of fill `fill` that lets more than 5% of windows through its P
// up-front bits: testing F's first bit right away (one round trip per 16
// windows of a lane) beats queueing them all (a round trip per drain of 64).
inline bool fold_passes_many(double fill, uint32_t P)
{
    double pass = 1.0;
    for (uint32_t i = 0; i < P; i++) pass *= fill;
    return pass > 0.05;
}
int use_device(ffbf_ctx *c);
// Workgroup size of the context's scan launches (its scan mode and tuning).
uint32_t scan_block(ffbf_ctx *c);
// Expected fraction of F's bits that are set (exact after a filter load).
double f_fill(ffbf_ctx *c);
}  // namespace ffbf
