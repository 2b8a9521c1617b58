/*
 * ffbf.h -- C ABI of the MI355X feed-forward Bloom filter (FFBF) matcher.
 *
 * The flow is the reference's: load a pattern set, build the forward filter
 * F from each pattern's W-byte prefix window, scan a text (every offset whose
 * k forward bits are set in F becomes a candidate and feeds k reverse bits
 * forward into R), run the pattern pass against R, then verify the surviving
 * (pattern, candidate) pairs exactly.  The result is the list of
 * (pattern id, text offset) matches, pattern id being the 0-based position of
 * the pattern in the loaded set (its line number for newline-separated input).
 *
 * Reference interfaces replaced (efficient/ffbf):
 *   setup_bloom_search        src/gpgpu/greptest.cu:455-474   -> ffbf_build
 *   executeGrepOverlap        src/gpgpu/greptest.cu:505-576   -> ffbf_scan_file / ffbf_scan / ffbf_scan_device
 *   executePatternFiltering   src/gpgpu/greptest.cu:578-596   -> ffbf_pattern_pass
 *   bvDump                    src/gpgpu/greptest.cu:341-351   -> ffbf_get_filter
 *   positions bitmap copy-out src/gpgpu/greptest.cu:697-704   -> ffbf_get_candidates
 *   patterns bitmap copy-out  src/gpgpu/greptest.cu:705-707   -> ffbf_get_survivors
 *   grep -Ff hand-off         src/rabin-karp/README:49-52     -> ffbf_verify / ffbf_get_matches
 *   Filter::operator<<        src/rabin-karp/ffbf.cpp:268-283 -> ffbf_set_patterns + ffbf_build
 *   Filter::processFile       src/rabin-karp/ffbf.cpp:310-392 -> ffbf_scan / ffbf_scan_fd
 *   Filter::filterPhrases     src/rabin-karp/ffbf.cpp:394-485 -> ffbf_pattern_pass + ffbf_write_survivors
 *   printpositions            src/gpgpu/greptest.cu:353-398   -> ffbf_write_candidate_lines
 *   printpatterns             src/gpgpu/greptest.cu:400-426   -> ffbf_write_survivors
 *   Filter::save/loadFilterToFile  src/rabin-karp/ffbf.cpp:487-527 -> ffbf_filter_save / ffbf_filter_load
 *   Filter<hash_rot_sbox_pre_2,19> src/rabin-karp/ffbf.cpp:88-185  -> ffbf_create_split
 *   Filter::operator[]        src/rabin-karp/ffbf.cpp:285-308 -> ffbf_query_line / ffbf_query_lines
 *
 * Conventions: every call returns FFBF_OK (0) or a negative status and never
 * exits or throws; ffbf_last_error() gives the calling thread's last message.
 * A context owns all of its device memory (filters stay resident between
 * scans).  Host pointers passed in are borrowed for the duration of the call;
 * a device text passed to ffbf_scan_device is borrowed until the next scan.
 * A context may be used by one thread at a time; distinct contexts are
 * independent.  Filters are uint32 words, bit b in word b>>5 under mask
 * 1u<<(b&31) (greptest.cu:80-89).
 */
#ifndef FFBF_H
#define FFBF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFBF_OK 0
#define FFBF_EINVAL (-1)    /* bad argument */
#define FFBF_EDEVICE (-2)   /* HIP runtime or kernel failure */
#define FFBF_ENOMEM (-3)    /* device or host allocation failed */
#define FFBF_ESTATE (-4)    /* call out of order (e.g. scan before build) */
#define FFBF_ECAPACITY (-5) /* caller buffer too small; required size reported */
#define FFBF_ECANCELED (-6) /* request cancelled before it completed */

#define FFBF_FILTER_FORWARD 0 /* F: built from patterns */
#define FFBF_FILTER_REVERSE 1 /* R: fed forward by the last scan */
#define FFBF_FILTER_FOLD 2    /* the scan's OR-fold of F held in LDS (modes 1 and 4) */
#define FFBF_FILTER_L2_FOLD 3 /* the scan's L2-resident OR-fold of F (modes 3 and 4) */

/* Method parameters.  The reference fixes them at compile time: HASH_LEN 19,
 * BLOOMBITS 2^28, four hash functions and its sbox table
 * (greptest.cu:50-54, 115-120; sbox.h:2-67). */
typedef struct ffbf_params {
    uint32_t window;    /* W, bytes per hashed window, 1..64 */
    uint32_t k;         /* hash functions per filter, 1..5 */
    uint32_t log2_bits; /* filter size 2^log2_bits bits, 10..32 */
    uint32_t flags;     /* must be 0 */
    uint32_t sbox[256]; /* substitution table of the rolling hash */
} ffbf_params;

/* The CPU binary's filter (src/rabin-karp/ffbf.cpp, Filter<H, LEN> with
 * hash_rot_sbox_pre_2): a 64-bit rolling hash whose low and high halves are
 * h1 and h2 (hash_rot_sbox_pre_2.h:26-27, 41-50); k hash functions
 * (getHashValue, ffbf.cpp:135-159; reverse getHashValue_2, :161-185), the
 * first s of them addressing a cache part of 2^c bits and the rest an
 * external part of 2^e bits placed after it (:154-158); the text hashed line
 * by line, restarting after every '\n' (:342-358), so windows holding a
 * newline are never candidates; bytes >= 0x80 hashed as 0 (:124-133).  The
 * reference's compile-time values are window 19, k 5, s 2, c 24, e 28 and the
 * table of hash_rot_sbox_pre_2.h:55-120.  Filters hold 2^c + 2^e bits. */
typedef struct ffbf_split_params {
    uint32_t window;          /* LEN, 1..64 */
    uint32_t k;               /* K, 1..11 */
    uint32_t s;               /* S, 0..k: functions addressing the cache part */
    uint32_t log2_cache_bits; /* c, 5..31 */
    uint32_t log2_ext_bits;   /* e, 5..31 */
    uint32_t flags;           /* must be 0 */
    uint64_t sbox[256];       /* 64-bit substitution table */
} ffbf_split_params;

typedef struct ffbf_match {
    uint32_t pattern_id;
    uint32_t reserved;
    uint64_t offset;
} ffbf_match;

/* Device time in milliseconds of the most recent run of each stage,
 * measured with HIP events on the context's stream. */
typedef struct ffbf_timings {
    float build_ms;        /* filter build (clear F, hash prefixes, set bits) */
    float scan_ms;         /* clears + scan with feed-forward + candidate compaction (ffbf_run_device:
                              its offsets only; the gather shares the pattern pass's launch) */
    float scan_kernel_ms;  /* the scan kernel alone */
    float pattern_pass_ms; /* pattern pass (+ survivor index; + the candidate gather in ffbf_run_device) */
    float verify_ms;       /* exact verification */
    uint32_t scan_mode;      /* 0 F probed in HBM/L2, 1 LDS fold, 2 whole F in LDS, 3 L2-resident fold,
                                4 LDS fold then L2-resident fold */
    uint32_t fold_log2_bits; /* prefilter size (0 in mode 0) */
    uint32_t probes;         /* forward bits tested against the prefilter before queueing */
    uint32_t block;          /* scan workgroup size */
    float ingest_ms;         /* last scan of host memory or a file: first byte read to last
                                candidate recorded (host clock, includes the overlapped copies) */
    uint32_t l2_fold_log2_bits; /* modes 3 and 4: size of the L2-resident fold (else 0) */
} ffbf_timings;

typedef struct ffbf_ctx ffbf_ctx;

const char *ffbf_last_error(void);
const char *ffbf_version(void);

/* Creates a context on HIP device `device`, allocating F and R. */
int ffbf_create(const ffbf_params *params, int device, ffbf_ctx **out);

/* Creates a context of the CPU binary's split filter (Filter<H, LEN>
 * constructor, ffbf.cpp:88-108).  Every entry point below works on it
 * unchanged (patterns, build, scan, pattern pass, verification, outputs,
 * persistence over its 2^c + 2^e bits, batches and servers); the filters of
 * one batch or server share one hash method, and split filters there share
 * W, k and the table.  Scan modes: -1 automatic, 0 (F probed directly; the
 * cache part first) or 1 (an LDS fold of the cache part, of at most 2^20
 * bits; needs s >= 1 and a cache part of at least 2^7 bits). */
int ffbf_create_split(const ffbf_split_params *params, int device, ffbf_ctx **out);
void ffbf_destroy(ffbf_ctx *ctx);

/* Loads a pattern set: pattern i is bytes[offsets[i] .. offsets[i+1]);
 * offsets has n_patterns+1 entries.  Copied to the device (resident).
 * Patterns shorter than W are kept (their ids count) but never match,
 * as ffbf.cpp:271-275 skips them.  Loading a set discards the built filter:
 * scans return FFBF_ESTATE until ffbf_build or ffbf_filter_load. */
int ffbf_set_patterns(ffbf_ctx *ctx, const uint8_t *bytes, const uint64_t *offsets, uint64_t n_patterns);

/* Same, from newline-separated text (one pattern per line, id = 0-based line
 * number as ffbf.cpp:404-446; a final line without '\n' counts). */
int ffbf_set_patterns_lines(ffbf_ctx *ctx, const char *buf, uint64_t len);

/* Builds F from the loaded patterns (greptest.cu:455-474, :96-123). */
int ffbf_build(ffbf_ctx *ctx);

/* Scans a host text: clears R and scans every offset in [0, n-W]
 * (executeGrepOverlap, greptest.cu:505-576).  A text longer than one ingest
 * piece (ffbf_set_ingest) moves into device memory owned by the context
 * through pinned staging buffers, piece by piece; each piece's copy overlaps
 * the scan of the windows already complete (a window straddling two pieces
 * is scanned once both have landed).  A shorter text is copied in one go. */
int ffbf_scan(ffbf_ctx *ctx, const uint8_t *text, uint64_t n);

/* The same for everything readable from a file descriptor, from its current
 * position to the end (Filter::processFile(int fd), ffbf.cpp:310-392, which
 * the reference's CLI runs on stdin).  Regular files are read in parallel
 * slices; pipes and sockets sequentially, the device text growing as needed. */
int ffbf_scan_fd(ffbf_ctx *ctx, int fd);

/* The same for a file by name (executeGrepOverlap(filename, ...),
 * greptest.cu:505-576). */
int ffbf_scan_file(ffbf_ctx *ctx, const char *path);

/* ffbf_scan_file, then the pattern pass and verification; n_matches may be
 * NULL (matches are then read with ffbf_get_matches). */
int ffbf_run_file(ffbf_ctx *ctx, const char *path, uint64_t *n_matches);

/* Ingest tuning for ffbf_scan, ffbf_scan_fd and ffbf_scan_file: bytes per
 * staging buffer (1 MiB..4 GiB; default 64 MiB), staging buffers (2..4;
 * default 4) and host threads filling each buffer (1..64; default 8).  0
 * keeps a value. */
int ffbf_set_ingest(ffbf_ctx *ctx, uint64_t chunk_bytes, uint32_t depth, uint32_t threads);

/* Scans a text already in device memory of the context's device. */
int ffbf_scan_device(ffbf_ctx *ctx, const uint8_t *dev_text, uint64_t n);

/* Scan, pattern pass and verification of a device-resident text queued back
 * to back on the context's stream with one synchronisation at the end (the
 * whole of greptest.cu:665-677 plus the exact step).  n_matches may be NULL;
 * the matches are then read with ffbf_get_matches. */
int ffbf_run_device(ffbf_ctx *ctx, const uint8_t *dev_text, uint64_t n, uint64_t *n_matches);

/* ffbf_run_device without a single device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * around HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The context must have been created, given patterns and built
 * beforehand (that is where its resident state was allocated).  The call's
 * own buffers (scan counters, candidate staging and list, match list, sort
 * storage) live in a fixed 32 MiB device array of the library, one copy per
 * GPU, reused by every call on that GPU; calls on one GPU are serialised by
 * the library (they would race on that array).  Limits of that array: at most
 * 262,144 candidate windows and 262,144 matches per call (FFBF_ECAPACITY
 * beyond, the text being left to ffbf_run_device), at most 2^24 patterns and
 * 2^40 text bytes (FFBF_EINVAL).  dev_matches: device memory of cap entries;
 * the matches land there sorted by (pattern id, offset), the same list
 * ffbf_run_device + ffbf_get_matches give; *n_matches receives the total
 * (FFBF_ECAPACITY if cap is smaller, nothing copied).  R and the survivor set
 * are updated as by ffbf_run_device; the context's candidate and match
 * getters do not see this call's results. */
int ffbf_run_device_noalloc(ffbf_ctx *ctx, const uint8_t *dev_text, uint64_t n, ffbf_match *dev_matches, uint64_t cap,
                            uint64_t *n_matches);

/* Pattern pass against R (executePatternFiltering, greptest.cu:578-596).
 * n_survivors may be NULL. */
int ffbf_pattern_pass(ffbf_ctx *ctx, uint64_t *n_survivors);

/* Exact verification of candidates x survivors; n_matches may be NULL. */
int ffbf_verify(ffbf_ctx *ctx, uint64_t *n_matches);

/* Copies the verified matches, sorted by (pattern id, offset).  *n receives
 * the total; returns FFBF_ECAPACITY if cap < total (nothing copied). */
int ffbf_get_matches(ffbf_ctx *ctx, ffbf_match *out, uint64_t cap, uint64_t *n);

/* One call for the whole path on a host text: scan, pattern pass, verify.
 * *out is allocated by the library (free with ffbf_free_matches). */
int ffbf_match_text(ffbf_ctx *ctx, const uint8_t *text, uint64_t n, ffbf_match **out, uint64_t *n_out);
void ffbf_free_matches(ffbf_match *matches);

/* The surviving patterns of the last pattern pass, one per line in id order,
 * written to patterns_path, and their 0-based ids, one per line, to
 * index_path (NULL: not written) -- the filtered phrases and "__index_" files
 * of Filter::filterPhrases (ffbf.cpp:394-485, named at ffbf.cpp:537-541;
 * printpatterns, greptest.cu:400-426, writes the first only). */
int ffbf_write_survivors(ffbf_ctx *ctx, const char *patterns_path, const char *index_path);

/* Every line of the last scanned text that holds a candidate window start,
 * once each, in text order, each followed by '\n' (printpositions,
 * greptest.cu:353-398): the filtered corpus the reference hands to grep -Ff.
 * A candidate that starts on a '\n' byte prints nothing; a last line without
 * '\n' ends at the end of the text.  text/n: the caller's host copy of the
 * scanned text, or NULL/0 to read it back from the device. */
int ffbf_write_candidate_lines(ffbf_ctx *ctx, const uint8_t *text, uint64_t n, const char *out_path);

/* Line query (Filter::operator[], ffbf.cpp:285-308): *hit = 1 iff the line
 * is at least W bytes long and one of its windows -- hashed from a reset
 * over the line, not restarted inside it (for the split method bytes >= 0x80
 * hash as 0, ffbf.cpp:124-133) -- has all k forward bits set in F.  Nothing is
 * fed forward; R and the results of the last scan are untouched.  Needs a
 * built (or loaded) filter. */
int ffbf_query_line(ffbf_ctx *ctx, const uint8_t *line, uint64_t len, int *hit);

/* The same for n_lines lines at once, line i being bytes[offsets[i] ..
 * offsets[i+1]) (offsets has n_lines+1 entries, offsets[0] = 0); hits[i]
 * receives 0 or 1.  One launch for all lines (one wave per line). */
int ffbf_query_lines(ffbf_ctx *ctx, const uint8_t *bytes, const uint64_t *offsets, uint64_t n_lines, uint8_t *hits);

/* Filter dump (bvDump, greptest.cu:341-351): n_words must be >= 2^log2_bits/32
 * for F and R.  The folds (no counterpart in the reference: the scan's
 * prefilters, for inspection and tests) have 2^fold_log2_bits / 32 and
 * 2^l2_fold_log2_bits / 32 words (ffbf_get_timings); FFBF_ESTATE when the
 * scan mode in force has no such fold. */
int ffbf_get_filter(ffbf_ctx *ctx, int which, uint32_t *out, uint64_t n_words);

/* Candidate offsets of the last scan, ascending (the reference's positions
 * bitmap, greptest.cu:246, in list form). */
int ffbf_get_candidates(ffbf_ctx *ctx, uint64_t *out, uint64_t cap, uint64_t *n);

/* Survivor bitmap of the last pattern pass (greptest.cu:284): bit i of word
 * i>>5 set iff pattern i survived; n_words >= ceil(n_patterns/32). */
int ffbf_get_survivors(ffbf_ctx *ctx, uint32_t *bitmap, uint64_t n_words, uint64_t *n_survivors);

/* F persisted as the reference's cache file body (ffbf.cpp:508-527): one
 * uint32 delta per set bit.  ffbf_filter_load replaces F (ffbf.cpp:487-506). */
int ffbf_filter_save(ffbf_ctx *ctx, uint32_t *deltas, uint64_t cap, uint64_t *n);
int ffbf_filter_load(ffbf_ctx *ctx, const uint32_t *deltas, uint64_t n);

/* How the scan tests windows before F in HBM (results never depend on it):
 * mode -1 picks automatically after each build; 0 first probe in F;
 * 1 OR-fold of F held in LDS, fold_log2_bits bits (0 = largest that fits);
 * 2 whole filter in LDS (only when it fits); 3 OR-fold of F in global memory
 * sized to stay in L2 (fold_log2_bits, 0 = the largest that fits one XCD's
 * L2, 2^25 bits on MI355X); 4 (F larger than L2) the LDS fold of mode 1
 * (fold_log2_bits), then mode 3's fold, then F.  A fold bit is the OR of
 * all F bits congruent to it, so the fold never rejects a true candidate.
 * Takes effect at the next scan. */
int ffbf_set_scan_mode(ffbf_ctx *ctx, int mode, uint32_t fold_log2_bits);

/* Launch tuning of the scan (results never depend on it): probes = forward
 * bits tested against an LDS prefilter before a window is queued (0 auto,
 * else 1..k; more than 3 acts as 3); block = workgroup size of the LDS
 * prefilter modes (0 auto = 1024, or 512); the global modes always use 256.
 * ffbf_get_timings reports the effective values. */
int ffbf_set_scan_tuning(ffbf_ctx *ctx, uint32_t probes, uint32_t block);

/* How ffbf_build sets F's bits (F is the same either way); no counterpart in
 * the reference, whose grepSetup sets one bit per atomic (greptest.cu:96-123).
 * FFBF_BUILD_ATOMIC: F cleared, then one device-scope atomicOr per bit.
 * FFBF_BUILD_BINNED: bit indices counting-sorted by 2^20-bit slice, each
 * slice set in LDS and written whole (no atomics on F, no clear).
 * FFBF_BUILD_AUTO (0): binned for filters of two slices or more with at
 * least 2^20 bit sets (where the atomics would dominate), else atomic.
 * Takes effect at the next ffbf_build. */
enum { FFBF_BUILD_AUTO = 0, FFBF_BUILD_ATOMIC = 1, FFBF_BUILD_BINNED = 2 };
int ffbf_set_build_method(ffbf_ctx *ctx, int method);

/* Stage timing events on (1, the default) or off (0).  Each event recorded
 * between two dependent kernels costs ~4-5 us of device time, about ten per
 * build and run; with timing off ffbf_get_timings reports 0 for the stage
 * times (the mode, fold and probe fields stay valid).  The reference times
 * its stages with host clocks around synchronised calls (ffbf.cpp:37-51). */
int ffbf_set_timing(ffbf_ctx *ctx, int on);

int ffbf_get_timings(ffbf_ctx *ctx, ffbf_timings *out);
int ffbf_synchronize(ffbf_ctx *ctx);

/* The context's HIP stream (a hipStream_t), for callers that order their own
 * device work against the matcher's. */
void *ffbf_get_stream(ffbf_ctx *ctx);

/* ---- Batched scans: many texts, many resident filters, one launch ----
 * The reference scans one text against one filter per run
 * (greptest.cu:505-576, ffbf.cpp:310-392); a batch binds several built
 * contexts (same device, window, k and sbox) and scans a set of device texts,
 * each routed to one of them, in one persistent launch replayed from a
 * captured hipGraph.  Each text's match list equals what ffbf_run_device
 * gives for it alone.  A filter's R after a batch run is the union over the
 * texts routed to it.  While a batch call runs its contexts must not be used
 * directly.  Between calls they may be rebuilt, given new patterns or another
 * scan mode: the next batch call picks the change up (a context whose new
 * pattern set is not yet built makes it fail with FFBF_ESTATE).  A context
 * must outlive every batch that holds it. */
typedef struct ffbf_batch ffbf_batch;

typedef struct ffbf_batch_match {
    uint32_t text_id;
    uint32_t pattern_id;
    uint64_t offset; /* within the text */
} ffbf_batch_match;

typedef struct ffbf_batch_tuning {
    uint32_t chunk_iters;    /* wave-iterations per work chunk (0 = 32); one is 992 bytes of text, or
                                2016 with 32 bytes per lane */
    uint32_t probes;         /* LDS prefilter bits tested before queueing, 1..min(k,3) (0 = auto) */
    uint32_t block;          /* workgroup size, 512 or 1024 (0 = 1024) */
    uint32_t use_graph;      /* 1: replay the captured hipGraph; 0: launch stage by stage */
    uint32_t bytes_per_lane; /* text bytes (window starts) per lane in the LDS-prefilter scan, 16 or 32
                                (0 = 32); 32 applies to W 19, 24 and 32 of the pair method, others use 16 */
} ffbf_batch_tuning;

int ffbf_batch_create(ffbf_ctx *const *filters, uint32_t n_filters, ffbf_batch **out);
void ffbf_batch_destroy(ffbf_batch *batch);
/* Texts are dev_texts[offsets[i] .. offsets[i+1]) (offsets has n_texts+1
 * entries), text i routed to filter filter_of_text[i] (index into the
 * create-time array).  The device buffer is borrowed until the next call. */
int ffbf_batch_set_texts(ffbf_batch *batch, const uint8_t *dev_texts, const uint64_t *offsets,
                         const uint32_t *filter_of_text, uint32_t n_texts);
int ffbf_batch_run(ffbf_batch *batch, uint64_t *n_matches);
/* The same texts through the single-text path, one ffbf_run_device each
 * (the per-request launch baseline); results land in the same match list. */
int ffbf_batch_run_per_text(ffbf_batch *batch, uint64_t *n_matches);
/* Matches sorted by (text id, pattern id, offset). */
int ffbf_batch_get_matches(ffbf_batch *batch, ffbf_batch_match *out, uint64_t cap, uint64_t *n);
int ffbf_batch_set_tuning(ffbf_batch *batch, const ffbf_batch_tuning *tuning);
int ffbf_batch_get_tuning(ffbf_batch *batch, ffbf_batch_tuning *tuning);
/* Times every (block, probes, chunk size) candidate on the current texts and
 * keeps the fastest; results never depend on the tuning. */
int ffbf_batch_autotune(ffbf_batch *batch, ffbf_batch_tuning *best, float *best_ms);
/* Device time of the last run (whole graph) and of its scan kernel (only
 * when the run was not a graph replay; else -1). */
int ffbf_batch_timings(ffbf_batch *batch, float *run_ms, float *scan_ms);
/* Timing events around each run on (1, the default) or off (0; about 10 us
 * less per run, ffbf_batch_timings then reports 0).  The autotuner times
 * its candidates either way.  A server's batches run without them. */
int ffbf_batch_set_timing(ffbf_batch *batch, int on);

/* ---- Serving: continuous batching over resident filters ----
 * The reference has no request loop.  A server keeps `slots` copies of the
 * given built contexts resident (each copy with its own R and survivor
 * index), and one worker per slot: a free slot admits, in arrival order,
 * every pending request that fits its text buffer and the HBM budget, stages
 * the texts, scans them as one batch on its own stream and completes them;
 * slots overlap on the device.  A request longer than a slot's buffer is
 * scanned in pieces of that size overlapping by the filter's longest pattern
 * minus one byte, each match reported once.  A request's matches equal what
 * ffbf_match_text gives for its text alone.  The contexts passed in are only
 * read at creation. */
typedef struct ffbf_server ffbf_server;

typedef struct ffbf_server_config {
    uint64_t hbm_budget_bytes;   /* device bytes of request data in flight (0 = 4 GiB) */
    uint64_t max_batch_bytes;    /* text bytes per batch = per slot buffer (0 = min(budget / slots, 128 MiB)) */
    uint32_t slots;              /* batches in flight at once (0 = 4) */
    uint32_t max_batch_requests; /* requests per batch (0 = 512) */
    uint32_t upload;             /* how request text reaches the device: FFBF_UPLOAD_* */
    uint32_t reserved;           /* 0 */
} ffbf_server_config;

/* FFBF_UPLOAD_COPY: hipMemcpyAsync per piece (the DMA engine).
 * FFBF_UPLOAD_KERNEL: a copy kernel on the slot's stream reads text that
 * lies wholly in one device-mapped host allocation (hipHostMalloc,
 * ffbf_host_alloc) over the link; other text still goes by
 * hipMemcpyAsync.  FFBF_UPLOAD_DEFAULT (0) is FFBF_UPLOAD_COPY. */
enum { FFBF_UPLOAD_DEFAULT = 0, FFBF_UPLOAD_COPY = 1, FFBF_UPLOAD_KERNEL = 2 };

typedef struct ffbf_server_stats {
    uint64_t submitted, completed, rejected, batches, bytes, pending, in_flight_bytes;
    uint64_t kernel_upload_bytes; /* request bytes the copy kernel uploaded */
} ffbf_server_stats;

int ffbf_server_create(ffbf_ctx *const *filters, uint32_t n_filters, const ffbf_server_config *config,
                       ffbf_server **out);
/* Finishes every request already submitted, then stops.  Call after the
 * last ffbf_server_wait. */
void ffbf_server_destroy(ffbf_server *server);
/* Queues a scan of text[0..n) (host memory, borrowed until the request
 * completes) against filter `filter`.  FFBF_ECAPACITY if a piece of it can
 * never be admitted (a slot's buffer, with its candidate reserve, larger
 * than the HBM budget); FFBF_EINVAL if it needs pieces and the filter's
 * longest pattern exceeds half a slot. */
int ffbf_server_submit(ffbf_server *server, uint32_t filter, const uint8_t *text, uint64_t n, uint64_t *request_id);
/* ffbf_server_submit for n requests at once (stops at the first error). */
int ffbf_server_submit_many(ffbf_server *server, uint32_t n, const uint32_t *filters, const uint8_t *const *texts,
                            const uint64_t *sizes, uint64_t *request_ids);
/* Cancels a request: its pieces not yet on the device are dropped and it
 * completes with FFBF_ECANCELED (at once, or when its pieces already being
 * scanned are done; its text is borrowed until then).  FFBF_ESTATE if it had
 * completed already (ffbf_server_wait still returns its result). */
int ffbf_server_cancel(ffbf_server *server, uint64_t request_id);
int ffbf_server_poll(ffbf_server *server, uint64_t request_id, int *done);
/* Blocks until the request completes; *matches (sorted by pattern id, then
 * offset) is allocated by the library (ffbf_free_matches); latency_ms is
 * submit-to-completion.  The request id is released.  One wait per request:
 * a second concurrent wait on the same id returns FFBF_EINVAL. */
int ffbf_server_wait(ffbf_server *server, uint64_t request_id, ffbf_match **matches, uint64_t *n, double *latency_ms);
int ffbf_server_get_stats(ffbf_server *server, ffbf_server_stats *out);

/* Plain device-memory helpers for callers without their own HIP code. */
int ffbf_device_alloc(int device, uint64_t bytes, void **out);
int ffbf_device_free(void *ptr);
int ffbf_device_upload(void *dst, const void *src, uint64_t bytes);
int ffbf_device_download(void *dst, const void *src, uint64_t bytes);
int ffbf_device_count(int *out);
/* Page-locked host memory (texts submitted from it upload at full link speed). */
int ffbf_host_alloc(uint64_t bytes, void **out);
int ffbf_host_free(void *ptr);
/* Plain copy kernel over 16-byte aligned device buffers (synchronous); the
 * reference point for HBM byte counters. */
int ffbf_device_copy(void *dst, const void *src, uint64_t bytes);

/* Random-gather microbenchmark: 16 independent uint32 loads per lane per
 * step at hashed addresses of a 2^log2_bytes-byte device array, over the
 * whole device.  Reports the time of one launch and the gather rate. */
int ffbf_gather_bench(int device, uint32_t log2_bytes, uint32_t steps, float *ms, double *probes_per_s);

/* The same access pattern with another kind of access: plain loads (what
 * ffbf_gather_bench runs), non-temporal loads, relaxed atomic loads at agent
 * or system scope, or device-scope atomicOr with the result unused (the
 * filter build's and the feed-forward's operation); or, for comparison,
 * `steps` streaming passes that read and write back the whole array.
 * FFBF_PROBE_UPLOAD times `steps` uploads of a 2^log2_bytes pinned host
 * buffer (log2_bytes 16..31) into device memory by the server's upload
 * kernel, as 64 pieces at staggered source alignments packed back to back.
 * Reports one launch's time and accesses per second (bytes read and written
 * per second for the streaming passes, bytes uploaded per second for the
 * upload).  No counterpart in the reference: it measures the rates that
 * bound the scan, the build and the serving loop's uploads. */
enum {
    FFBF_PROBE_LOAD = 0,
    FFBF_PROBE_LOAD_NONTEMPORAL = 1,
    FFBF_PROBE_LOAD_AGENT = 2,
    FFBF_PROBE_LOAD_SYSTEM = 3,
    FFBF_PROBE_ATOMIC_OR = 4,
    FFBF_PROBE_STREAM = 5,
    FFBF_PROBE_UPLOAD = 6
};
int ffbf_probe_bench(int device, uint32_t log2_bytes, uint32_t steps, uint32_t kind, float *ms, double *ops_per_s);

#ifdef __cplusplus
}
#endif
#endif
