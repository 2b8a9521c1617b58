// ffbf_device.h -- records and launchers shared by the HIP kernels
// (ffbf_kernels.hip) and the host context (ffbf_capi.cpp).  Internal: not part
// of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ffbf {

// One candidate window: offset in the text and its hash pair (h1, h2), so the
// pattern pass / verification never rehash the text.
struct Cand {
    uint64_t offset;
    uint32_t h1, h2;
};
static_assert(sizeof(Cand) == 16, "Cand is 16 bytes");

// Same layout as ffbf_match in ffbf.h.
struct MatchRec {
    uint32_t id, reserved;
    uint64_t offset;
};
static_assert(sizeof(MatchRec) == 16, "MatchRec is 16 bytes");

// Device counters (unsigned long long each).
// C_CAND: candidates found (beyond the list capacity when some were lost, so
// that the caller grows the list and scans again); C_CVALID: leading entries
// of the compacted list that hold candidates (what verification reads).
enum Counter : int { C_CAND = 0, C_SURV = 1, C_MATCH = 2, C_SPILL = 4, C_CVALID = 5, C_COUNT = 8 };

// Candidate staging of a scan launch: wave w of the launch appends to its own
// region of `region` entries at (wave0 + w) * region of the raw candidate
// buffer, with no atomics, and records how many it kept; a wave whose region
// is full appends to the shared spill area after all regions (one counter
// atomic per wave flush).  A compaction pass then appends every region and
// the spill to the contiguous candidate list.
constexpr uint32_t kRegionWavesPerCU = 64;  // region slots per CU: two launches of up to 32 waves

// How the scan tests a window before the full forward probe.
enum ScanMode : uint32_t {
    MODE_GLOBAL = 0,  // first forward bit read from F itself (L2/MALL), the rest for survivors
    MODE_FOLD = 1,    // OR-fold of F held in LDS tested first (exact superset), then F
    MODE_LDS = 2,     // the whole of F held in LDS: the LDS test is exact
    MODE_GFOLD = 3,   // OR-fold of F in global memory, small enough to stay in L2, then F
    MODE_HYBRID = 4,  // LDS fold first, then the L2-resident fold's first bit before queueing;
                      // the queue drains through the L2 fold's other bits, then F
};

// Modes whose scan holds a prefilter in each workgroup's LDS.
__host__ __device__ inline bool mode_in_lds(uint32_t m) { return m == MODE_FOLD || m == MODE_LDS || m == MODE_HYBRID; }

// Hash method of a context.  PAIR: the GPU reference's two 32-bit rolling
// hashes and one power-of-two filter (greptest.cu).  SPLIT: the CPU binary's
// 64-bit hash, filter in a cache part and an external part, hashing
// restarted at every newline and bytes >= 0x80 hashed as 0 (ffbf.cpp).
enum HashMethod : uint32_t { METHOD_PAIR = 0, METHOD_SPLIT = 1 };

// Placement of the split method's bit indices (ffbf.cpp:154-158,
// 180-184): function j < s keeps x & cmask (cache part, [0, cache)); the
// others (x & emask) + cache (external part).
struct SplitMap {
    uint32_t s, cmask, emask, cache;
};
constexpr int kSplitMaxK = 11;  // getHashValue has 11 functions

struct ScanArgs {
    const uint8_t *text;
    uint64_t n;
    const uint32_t *F;
    uint32_t *R;
    const uint32_t *pre;  // prefilter words: F itself or a fold of it
    const uint32_t *sbox;
    Cand *cand;
    uint64_t cand_cap;
    unsigned long long *counters;
    uint32_t W, k, mask;
    uint32_t pre_words;  // power of two (>= 4 when copied to LDS)
    uint32_t pre_mask;   // pre_words*32 - 1
    uint32_t pre_log2w;  // log2(pre_words)
    uint32_t P;          // forward bits tested against the prefilter before queueing (1..k; 1 for global)
    uint32_t pre_exact;  // prefilter is F itself: no further F probes
    uint32_t lds;        // prefilter is copied to LDS (batched scan: per filter)
    uint32_t f_first;    // LDS fold passing many windows: F's first forward bit is tested right
                         // after it, before queueing (one round trip per 16 windows, not per 64 queued)
    uint32_t r_lds;      // whole F in LDS and small: R is gathered in LDS right after F (pre_words
                         // words) and ORed into global R when the workgroup is done with the filter
    uint64_t off_base;   // added to every recorded offset (a text scanned in pieces)
    // candidate staging (cand/cand_cap above are the raw buffer)
    uint32_t *wave_counts;  // entries kept per region slot
    uint32_t region;        // entries per wave region
    uint32_t wave0;         // region slot of the launch's first wave
    uint32_t region_end;    // slots [wave0, region_end) are the launch's; waves beyond only spill
    uint64_t spill_base;    // first entry after all region slots
    uint32_t method;        // HashMethod; SPLIT: sbox is 256 (lo, hi) pairs, mask unused
    SplitMap split;
    // MODE_HYBRID: the L2-resident fold behind the LDS fold (f_first tests its
    // first bit instead of F's; the drain tests its other bits, then all of F)
    const uint32_t *pre2;
    uint32_t pre2_mask;
};

// Batched scan: many texts (concatenated in one device buffer), each routed
// to one of several resident filters, walked by persistent workgroups that
// pull chunks of wave-iterations (kIterUnits 16-byte units) from a global counter.
struct BatchText {
    uint64_t start;  // offset of the text in the batch buffer
    uint64_t n;      // bytes
    uint32_t filter;
    uint32_t reserved;
};

struct BatchChunk {
    uint32_t text;
    uint32_t filter;
    uint32_t it0;   // first wave-iteration (kIterUnits 16-byte units) of the text
    uint32_t n_it;  // iterations in the chunk
};

struct BatchMatchRec {
    uint32_t text, id;
    uint64_t offset;
};
static_assert(sizeof(BatchMatchRec) == 16, "BatchMatchRec is 16 bytes");

// Bits of a batched candidate's offset field that hold the offset in its
// text; the text id sits above them.
constexpr int kBatchOffsetBits = 40;

// Scan geometry: a wave-iteration loads 64 consecutive 16-byte units of text,
// one per lane, and owns the windows that start in the first kIterUnits of
// them; the top lanes' units only complete the windows of the lanes below.
constexpr uint32_t kIterUnits = 62;
// Two units per lane (32 window starts): a wave-iteration owns the windows of
// the first 126 of 128 units; lane 63's pair only completes lane 62's windows.
constexpr uint32_t kIterUnits2 = 126;

struct BatchArgs {
    const uint8_t *texts;
    const BatchText *tdesc;
    const BatchChunk *chunks;
    const ScanArgs *filters;  // per filter; cand/counters are the batch's
    const uint32_t *sbox;     // split method: 256 (lo, hi) pairs
    unsigned long long *ctl;  // [0] next chunk, [1] number of chunks, [2] index of the first, [3] number of texts
    uint32_t W;
    uint32_t method;          // HashMethod of every filter of the batch
};

struct VerifyArgs {
    const uint8_t *text;
    uint64_t n;
    const Cand *cand;
    uint64_t cand_cap;
    const uint8_t *bytes;
    const uint64_t *off;
    const uint2 *pat_hash;
    const uint32_t *table;       // table_size slots: 0 empty, else representative id + 1
    const uint32_t *slot_chain;  // table_size heads: further ids with the slot's key (id + 1)
    const uint32_t *chain_next;  // per pattern: next id + 1 in its chain
    uint32_t table_mask;
    MatchRec *out;
    uint64_t out_cap;
    unsigned long long *counters;
};

struct BatchVerifyArgs {
    const uint8_t *texts;
    const BatchText *tdesc;
    const Cand *cand;
    uint64_t cand_cap;
    const VerifyArgs *filters;  // per filter: pattern set and survivor index (text fields unused)
    BatchMatchRec *out;
    uint64_t out_cap;
    unsigned long long *counters;  // C_CVALID read, C_MATCH appended
    const unsigned long long *ctl;  // batch control words; [3] = number of texts
};

struct PassArgs {
    const uint2 *pat_hash;
    const uint64_t *off;
    uint64_t n;
    uint32_t W, k, mask;
    const uint32_t *R;
    uint32_t *bitmap;   // ceil(n/32) words, fully rewritten
    unsigned long long *counters;
    uint32_t *table;
    uint32_t *slot_chain;
    uint32_t *chain_next;
    uint32_t table_mask;
    uint32_t method;  // HashMethod
    SplitMap split;
};

// Line queries (Filter::operator[]): line i is bytes[off[i] .. off[i+1]).
struct QueryArgs {
    const uint8_t *bytes;
    const uint64_t *off;
    uint64_t n_lines;
    const uint32_t *F;
    const uint32_t *sbox;  // split method: 256 (lo, hi) pairs, entries 0x80..0xFF = entry 0
    uint32_t W, k, mask;
    uint32_t method;  // HashMethod
    SplitMap split;
    uint8_t *hits;    // per line: 1 if some window has its k forward bits set in F
};

// A run of u32 words to clear (a few thousand words each, so that many
// workgroups share a large buffer).
struct ZeroSpan {
    uint32_t *p;
    uint32_t words;
};

// Pieces of host text, each in one device-mapped host allocation (the GPU
// reads it at the host address), copied back to back into a device buffer
// by one launch.  The work is cut into segments of 63 16-byte destination
// chunks (one wave each): piece p has segments [seg_start[p],
// seg_start[p + 1]).
constexpr uint32_t kUploadMax = 64;  // pieces per launch (1.8 KiB of kernel arguments)
struct UploadArgs {
    uint8_t *dst;
    uint32_t n;           // pieces in this launch
    uint32_t total_segs;  // seg_start[n]
    uint64_t src[kUploadMax];
    uint64_t dst_off[kUploadMax];
    uint64_t len[kUploadMax];
    uint32_t seg_start[kUploadMax + 1];
};
// Segments of a piece of n > 0 bytes at destination offset dst_off.
inline uint32_t upload_segs(uint64_t dst_off, uint64_t n)
{
    const uint64_t chunks = ((dst_off + n - 1u) >> 4) - (dst_off >> 4) + 1u;
    return (uint32_t)((chunks + 62u) / 63u);
}

// Launchers (hipError_t of the launch).
hipError_t launch_fill32(uint32_t *p, uint64_t n_words, uint32_t value, hipStream_t s);
hipError_t launch_copy(void *dst, const void *src, uint64_t bytes, hipStream_t s);
hipError_t launch_upload(const UploadArgs &a, hipStream_t s);
// Prefix hash pairs of the patterns' first W bytes ((0, 0) for shorter ones).
hipError_t launch_prefix_hash(const uint8_t *bytes, const uint64_t *off, uint64_t n, const uint32_t *sbox, uint32_t W,
                              uint2 *pat_hash, hipStream_t s);
// The split method's: 64-bit prefix hashes from a 256-entry (lo, hi) sbox.
hipError_t launch_prefix_hash_split(const uint8_t *bytes, const uint64_t *off, uint64_t n, const uint32_t *sbox64,
                                    uint32_t W, uint2 *pat_hash, hipStream_t s);
// Up to two OR-folds of F that the atomic build sets beside F (null p: none).
// Bit x of F with x < src_bits is bit x mod 32 (wmask + 1) of each: the fold
// of F's first src_bits bits onto wmask + 1 words (a power of two).
struct BuildFolds {
    uint32_t *p[2];
    uint32_t wmask[2];
    uint64_t src_bits;
};

// Atomic build: the k forward bits of every pattern at least W long, set by
// atomicOr into a cleared F (and cleared folds) from the prefix hash pairs
// (placed by m when split).
hipError_t launch_build_atomic(const uint64_t *off, const uint2 *pat_hash, uint64_t n, uint32_t W, uint32_t k,
                               uint32_t mask, bool split, const SplitMap &m, uint32_t *F, const BuildFolds &folds,
                               hipStream_t s);
// Atomic-free build (filters of more than one 2^20-bit slice): bit indices of
// kBinEntries / k patterns per block sorted by slice into `staging`
// (build_bin_blocks(n, k) * kBinEntries words) with per-block slice offsets in
// `offs` (blocks * (slices + 1) words), then each slice set in LDS and written
// whole to F.  Reads the prefix hash pairs; F need not be cleared first.
constexpr uint32_t kBinSliceLog2 = 20;
constexpr uint32_t kBinEntries = 6144;
inline uint32_t build_bin_patterns(uint32_t k) { return kBinEntries / (k ? k : 1u); }
uint64_t build_bin_blocks(uint64_t n, uint32_t k);
hipError_t launch_build_binned(const uint64_t *off, const uint2 *pat_hash, uint64_t n, uint32_t W, uint32_t k,
                               uint32_t mask, bool split, SplitMap m, uint32_t *F, uint64_t f_words, uint32_t *staging,
                               uint32_t *offs, hipStream_t s);
hipError_t launch_fold(const uint32_t *F, uint64_t f_words, uint32_t *fold, uint32_t fold_words, hipStream_t s);
// waves (when given): the waves launched, whose region slots a.wave0 .. +waves
// the launch may fill
hipError_t launch_scan(const ScanArgs &a, uint32_t mode, uint32_t block, int n_cu, hipStream_t s,
                       uint32_t *waves = nullptr);
// Random-access microbenchmark (kind: FFBF_PROBE_* of ffbf.h).
hipError_t launch_gather(uint32_t *arr, uint32_t log2_words, uint32_t steps, uint32_t *out, uint32_t blocks,
                         uint32_t kind, hipStream_t s);
hipError_t launch_pattern_pass(const PassArgs &a, int n_cu, hipStream_t s);
// Pattern passes of nf filters in one launch (wave_start[nf+1], device
// arrays); filters whose active flag is 0 get an empty survivor set.
hipError_t launch_pattern_pass_multi(const PassArgs *filters, const uint64_t *wave_start, uint32_t nf,
                                     const uint32_t *active, uint64_t total_waves, int n_cu, hipStream_t s,
                                     bool split = false);
// Clears every span of the list in one launch.
hipError_t launch_zero_spans(const ZeroSpan *spans, uint32_t n, int n_cu, hipStream_t s);
// Appends the regions and spill of the raw candidate buffer to the list
// out[counters[C_CAND]..], updates C_CAND (beyond out_cap when anything was
// lost) and zeroes the region counts and C_SPILL for the next launch.
// n_used (at most n_regions): only region slots [0, n_used) can hold
// entries (the launch's waves); the spill follows all n_regions slots.
hipError_t launch_cand_compact(const Cand *raw, uint64_t raw_cap, uint32_t *wave_counts, uint64_t *offsets,
                               uint32_t n_regions, uint32_t region, Cand *out, uint64_t out_cap,
                               unsigned long long *counters, int n_cu, hipStream_t s, uint32_t n_used = UINT32_MAX);
// Where the candidate gather copies from and to (the offsets' results).
struct GatherArgs {
    const Cand *raw;
    const uint64_t *offsets;
    uint32_t n;  // regions
    uint32_t region;
    uint64_t spill_base;
    Cand *out;
    uint64_t out_cap;
};
// The compaction's first half (offsets, counters, cleared counts); *g
// receives the gather it leaves to launch_pass_gather.
hipError_t launch_cand_offsets(const Cand *raw, uint64_t raw_cap, uint32_t *wave_counts, uint64_t *offsets,
                               uint32_t n_regions, uint32_t region, Cand *out, uint64_t out_cap,
                               unsigned long long *counters, hipStream_t s, GatherArgs *g, uint32_t n_used = UINT32_MAX);
// The pattern pass and the candidate gather g in one launch.
hipError_t launch_pass_gather(const PassArgs &a, const GatherArgs &g, int n_cu, hipStream_t s);
hipError_t launch_verify(const VerifyArgs &a, int n_cu, hipStream_t s);
hipError_t launch_query_lines(const QueryArgs &a, int n_cu, hipStream_t s);

// units_per_lane 2 (32 window starts per lane) applies to compile-time W of
// the pair method; chunks of the LDS class are then counted in wave-iterations
// of kIterUnits2 units
hipError_t launch_scan_batch(const BatchArgs &a, uint32_t block, uint32_t probes, uint32_t units_per_lane,
                             uint32_t pre_words_max, int n_cu, hipStream_t s);
// the same over filters whose prefilter is in global memory (256-thread blocks, no LDS prefilter)
hipError_t launch_scan_batch_global(const BatchArgs &a, int n_cu, hipStream_t s);
hipError_t launch_verify_batch(const BatchVerifyArgs &a, int n_cu, hipStream_t s);
hipError_t sort_batch_matches(BatchMatchRec *m, uint64_t n, uint64_t *keys_a, uint64_t *keys_b, void *temp,
                              size_t temp_bytes, int id_shift, int text_shift, int end_bit, hipStream_t s);

// Match lists shorter than this are sorted on the host at copy-out.
constexpr uint64_t kDeviceSortMin = 4096;
// Match records a fused run brings to the host speculatively, with its counters.
constexpr uint64_t kMatchPrefetch = 4096;

// Device sort of matches by (id, offset) through packed 64-bit keys.
bool sort_packable(uint64_t n_patterns, uint64_t text_bytes);
int sort_end_bit(uint64_t n_patterns);
hipError_t sort_temp_bytes(uint64_t n, int end_bit, size_t *bytes);
hipError_t sort_matches(MatchRec *m, uint64_t n, uint64_t *keys_a, uint64_t *keys_b, void *temp, size_t temp_bytes,
                        int end_bit, hipStream_t s);

// Largest fold (in 32-bit words) the scan kernel can hold in LDS on this
// device (split: beside the split method's 2 KiB sbox).
uint32_t max_fold_words(int device, bool split = false);
// Whether a templated fast scan exists for this window length.
bool scan_has_fast_path(uint32_t W);

}  // namespace ffbf
