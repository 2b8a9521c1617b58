// ffbf_ingest.cpp -- text ingest from host memory, file descriptors and files,
// and the reference's text-format outputs.
//
// Ingest (executeGrepOverlap, greptest.cu:505-576; Filter::processFile,
// ffbf.cpp:310-392): the text goes through a ring of pinned staging buffers
// (filled by parallel memcpy or pread, or by read() for pipes) into one
// device buffer on a copy stream.  As soon as a piece has landed, every window
// whose W bytes are all on the device is scanned on the context's stream,
// which waits only for that piece's copy; each scan launch starts at the first
// window the previous one did not cover, so windows that straddle two pieces
// are scanned exactly once (the reference's per-chunk halo can read bytes
// whose copy has not finished, greptest.cu:542-566).  Reading the next piece,
// its H2D copy and the scan of the previous pieces overlap.
//
// Outputs: the surviving patterns with their 0-based line numbers
// (filterPhrases, ffbf.cpp:394-485; printpatterns, greptest.cu:400-426) and
// the lines of the text that hold a candidate window (printpositions,
// greptest.cu:353-398), the two files the reference hands to grep -Ff.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "ffbf_ctx.h"

using namespace ffbf;

namespace {

// A byte stream read piece by piece into pinned staging memory.
struct Source {
    virtual ~Source() = default;
    // Fills dst with up to cap bytes; returns the count (0 at the end of the
    // input) or -1 after recording an error.
    virtual int64_t read(uint8_t *dst, uint64_t cap) = 0;
};

// Runs fn(lo, hi) over [0, n) split into up to `threads` slices of whole
// 64 KiB blocks, on worker threads when there is enough to share.
template <class Fn>
void parallel_slices(uint64_t n, uint32_t threads, Fn fn)
{
    const uint64_t block = 64u << 10;
    const uint64_t blocks = (n + block - 1) / block;
    const uint64_t t = std::min<uint64_t>(std::max<uint32_t>(threads, 1u), blocks);
    if (t <= 1 || n < (4u << 20)) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(t);
    const uint64_t per = (blocks + t - 1) / t * block;
    for (uint64_t lo = 0; lo < n; lo += per) pool.emplace_back(fn, lo, std::min(n, lo + per));
    for (auto &th : pool) th.join();
}

struct MemSource : Source {
    const uint8_t *p;
    uint64_t n, pos = 0;
    uint32_t threads;
    MemSource(const uint8_t *p_, uint64_t n_, uint32_t t) : p(p_), n(n_), threads(t) {}
    int64_t read(uint8_t *dst, uint64_t cap) override
    {
        const uint64_t m = std::min(cap, n - pos);
        const uint8_t *src = p + pos;
        parallel_slices(m, threads, [&](uint64_t lo, uint64_t hi) { std::memcpy(dst + lo, src + lo, hi - lo); });
        pos += m;
        return (int64_t)m;
    }
};

// A regular file of known size: slices of each piece are read in parallel.
struct FileSource : Source {
    int fd;
    uint64_t pos, end;
    uint32_t threads;
    FileSource(int fd_, uint64_t start, uint64_t size, uint32_t t) : fd(fd_), pos(start), end(size), threads(t) {}
    int64_t read(uint8_t *dst, uint64_t cap) override
    {
        const uint64_t m = std::min(cap, end - pos);
        bool bad = false;
        int err = 0;
        parallel_slices(m, threads, [&](uint64_t lo, uint64_t hi) {
            uint64_t at = lo;
            while (at < hi) {
                const ssize_t r = ::pread(fd, dst + at, hi - at, (off_t)(pos + at));
                if (r < 0 && errno == EINTR) continue;
                if (r <= 0) {
                    __atomic_store_n(&bad, true, __ATOMIC_RELAXED);
                    if (r < 0) __atomic_store_n(&err, errno, __ATOMIC_RELAXED);
                    return;
                }
                at += (uint64_t)r;
            }
        });
        if (bad) {
            fail(FFBF_EINVAL, err ? std::string("reading the text file: ") + std::strerror(err)
                                  : std::string("text file shrank while it was read"));
            return -1;
        }
        pos += m;
        return (int64_t)m;
    }
};

// Pipes, sockets, terminals: sequential reads until the piece is full.
struct StreamSource : Source {
    int fd;
    explicit StreamSource(int fd_) : fd(fd_) {}
    int64_t read(uint8_t *dst, uint64_t cap) override
    {
        uint64_t got = 0;
        while (got < cap) {
            const ssize_t r = ::read(fd, dst + got, cap - got);
            if (r < 0 && errno == EINTR) continue;
            if (r < 0) {
                fail(FFBF_EINVAL, std::string("reading the text: ") + std::strerror(errno));
                return -1;
            }
            if (r == 0) break;
            got += (uint64_t)r;
        }
        return (int64_t)got;
    }
};

int ensure_staging(ffbf_ctx *c)
{
    if (!c->copy_stream) HIP_TRY(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    if (c->staging_bytes == c->ingest_chunk && c->staging_depth == c->ingest_depth) return FFBF_OK;
    HIP_TRY(hipStreamSynchronize(c->copy_stream));
    ingest_release_staging(c);
    for (uint32_t i = 0; i < c->ingest_depth; i++) {
        HIP_TRY(hipHostMalloc((void **)&c->staging[i], c->ingest_chunk, hipHostMallocDefault));
        HIP_TRY(hipEventCreateWithFlags(&c->ev_staged[i], hipEventDisableTiming));
        c->staged_pending[i] = false;
    }
    c->staging_bytes = c->ingest_chunk;
    c->staging_depth = c->ingest_depth;
    return FFBF_OK;
}

// Grows the owned device text to hold `need` bytes, keeping its first `keep`.
int grow_text(ffbf_ctx *c, uint64_t need, uint64_t keep)
{
    if (need <= c->owned_text.cap && c->owned_text.p) return FFBF_OK;
    // scans and copies in flight read or write the old buffer
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->copy_stream) HIP_TRY(hipStreamSynchronize(c->copy_stream));
    DevBuf<uint8_t> nb;
    RET_IF(nb.ensure(std::max<uint64_t>(need, 2 * c->owned_text.cap), "text"));
    if (keep) HIP_TRY(hipMemcpy(nb.p, c->owned_text.p, keep, hipMemcpyDeviceToDevice));
    c->owned_text.release();
    c->owned_text = nb;
    return FFBF_OK;
}

// Waits for every staged copy so the staging ring can be refilled.
int drain_copies(ffbf_ctx *c)
{
    if (c->copy_stream) HIP_TRY(hipStreamSynchronize(c->copy_stream));
    for (uint32_t i = 0; i < c->staging_depth; i++) c->staged_pending[i] = false;
    return FFBF_OK;
}

// Reads the whole source into the owned device text, scanning as it lands.
// size_hint: the exact size when known (0 otherwise).
int ingest(ffbf_ctx *c, Source &src, uint64_t size_hint)
{
    if (!c->built) return fail(FFBF_ESTATE, "scan before build (or filter load)");
    const auto t0 = std::chrono::steady_clock::now();
    RET_IF(ensure_staging(c));
    const uint64_t W = c->params.window;
    RET_IF(grow_text(c, std::max<uint64_t>(size_hint, 1), 0));
    RET_IF(set_text(c, c->owned_text.p, size_hint));
    RET_IF(enqueue_scan_begin(c));
    uint64_t copied = 0, scanned = 0;  // bytes on the device; first window not yet scanned
    int st = FFBF_OK;
    for (uint32_t i = 0;; i++) {
        const uint32_t s = i % c->staging_depth;
        if (c->staged_pending[s]) {
            HIP_TRY(hipEventSynchronize(c->ev_staged[s]));
            c->staged_pending[s] = false;
        }
        const int64_t got = src.read(c->staging[s], c->staging_bytes);
        if (got < 0) {
            st = FFBF_EINVAL;
            break;
        }
        if (got == 0) break;
        if (copied + (uint64_t)got > c->owned_text.cap) {
            st = grow_text(c, copied + (uint64_t)got, copied);
            if (st != FFBF_OK) break;
        }
        HIP_TRY(hipMemcpyAsync(c->owned_text.p + copied, c->staging[s], (size_t)got, hipMemcpyHostToDevice,
                               c->copy_stream));
        HIP_TRY(hipEventRecord(c->ev_staged[s], c->copy_stream));
        c->staged_pending[s] = true;
        copied += (uint64_t)got;
        if (copied >= W && copied - W + 1 > scanned) {
            // every window starting before copied-W+1 has all its bytes here
            const uint64_t end = copied - W + 1;
            HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_staged[s], 0));
            RET_IF(enqueue_scan_range(c, c->owned_text.p + scanned, end - scanned + W - 1, scanned, false));
            scanned = end;
        }
    }
    if (st != FFBF_OK) {
        (void)hipStreamSynchronize(c->stream);
        (void)drain_copies(c);
        c->text = nullptr;
        c->text_n = 0;
        return st;
    }
    RET_IF(record(c, EV_S2));
    RET_IF(record(c, EV_S3));
    c->text = c->owned_text.p;
    c->text_n = copied;
    RET_IF(sync_counters(c));
    RET_IF(drain_copies(c));
    if (c->h_counters[C_CAND] > c->cand.cap) {
        // rescan the resident text with room for every candidate
        RET_IF(ensure_cand(c, 2 * c->h_counters[C_CAND]));
        RET_IF(do_scan(c));
    } else {
        c->n_cand = c->h_counters[C_CAND];
        c->scanned = true;
        c->passed = c->verified = false;
    }
    c->ingest_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return FFBF_OK;
}

int scan_fd(ffbf_ctx *c, int fd)
{
    struct stat sb {};
    if (fstat(fd, &sb) == 0 && S_ISREG(sb.st_mode)) {
        const off_t here = lseek(fd, 0, SEEK_CUR);
        const uint64_t start = here > 0 ? (uint64_t)here : 0u;
        const uint64_t size = (uint64_t)sb.st_size > start ? (uint64_t)sb.st_size - start : 0u;
        FileSource src(fd, start, (uint64_t)sb.st_size, c->ingest_threads);
        RET_IF(ingest(c, src, size));
        (void)lseek(fd, (off_t)(start + c->text_n), SEEK_SET);
        return FFBF_OK;
    }
    StreamSource src(fd);
    return ingest(c, src, 0);
}

// Bytes of the scanned text on the host: the caller's copy, or read back.
int host_text(ffbf_ctx *c, const uint8_t *text, uint64_t n, std::vector<uint8_t> &buf, const uint8_t **out,
              uint64_t *out_n)
{
    if (text) {
        if (n != c->text_n) return fail(FFBF_EINVAL, "text length differs from the scanned text");
        *out = text;
        *out_n = n;
        return FFBF_OK;
    }
    buf.resize(c->text_n);
    if (c->text_n) {
        HIP_TRY(hipMemcpyAsync(buf.data(), c->text, c->text_n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    *out = buf.data();
    *out_n = c->text_n;
    return FFBF_OK;
}

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
};

int open_out(File &out, const char *path)
{
    if (!path) return fail(FFBF_EINVAL, "null output path");
    out.f = std::fopen(path, "wb");
    if (!out.f) return fail(FFBF_EINVAL, std::string("cannot open ") + path + ": " + std::strerror(errno));
    return FFBF_OK;
}

int close_out(File &out, const char *path)
{
    const bool bad = std::ferror(out.f) != 0;
    const int r = std::fclose(out.f);
    out.f = nullptr;
    if (bad || r != 0) return fail(FFBF_EINVAL, std::string("writing ") + path + " failed");
    return FFBF_OK;
}

}  // namespace

namespace ffbf {

void ingest_release_staging(ffbf_ctx *c)
{
    for (uint32_t i = 0; i < 4; i++) {
        if (c->staging[i]) (void)hipHostFree(c->staging[i]);
        if (c->ev_staged[i]) (void)hipEventDestroy(c->ev_staged[i]);
        c->staging[i] = nullptr;
        c->ev_staged[i] = nullptr;
        c->staged_pending[i] = false;
    }
    c->staging_bytes = 0;
    c->staging_depth = 0;
}

void ingest_release(ffbf_ctx *c)
{
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    ingest_release_staging(c);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    c->copy_stream = nullptr;
}

int scan_host(ffbf_ctx *c, const uint8_t *text, uint64_t n)
{
    MemSource src(text, n, c->ingest_threads);
    return ingest(c, src, n);
}

}  // namespace ffbf

extern "C" {

int ffbf_set_ingest(ffbf_ctx *c, uint64_t chunk_bytes, uint32_t depth, uint32_t threads)
{
    RET_IF(check_ctx(c));
    if (chunk_bytes && (chunk_bytes < (1u << 20) || chunk_bytes > (1ull << 32)))
        return fail(FFBF_EINVAL, "chunk_bytes must be 1 MiB..4 GiB (0 keeps it)");
    if (depth && (depth < 2 || depth > 4)) return fail(FFBF_EINVAL, "depth must be 2..4 (0 keeps it)");
    if (threads > 64) return fail(FFBF_EINVAL, "threads must be 0..64 (0 keeps it)");
    if (chunk_bytes) c->ingest_chunk = chunk_bytes;
    if (depth) c->ingest_depth = depth;
    if (threads) c->ingest_threads = threads;
    return FFBF_OK;
}

int ffbf_scan_fd(ffbf_ctx *c, int fd)
{
    RET_IF(check_ctx(c));
    if (fd < 0) return fail(FFBF_EINVAL, "bad file descriptor");
    return scan_fd(c, fd);
}

int ffbf_scan_file(ffbf_ctx *c, const char *path)
{
    RET_IF(check_ctx(c));
    if (!path) return fail(FFBF_EINVAL, "null path");
    const int fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (fd < 0) return fail(FFBF_EINVAL, std::string("cannot open ") + path + ": " + std::strerror(errno));
    const int st = scan_fd(c, fd);
    ::close(fd);
    return st;
}

int ffbf_run_file(ffbf_ctx *c, const char *path, uint64_t *n_matches)
{
    RET_IF(ffbf_scan_file(c, path));
    RET_IF(ffbf_pattern_pass(c, nullptr));
    return ffbf_verify(c, n_matches);
}

int ffbf_write_survivors(ffbf_ctx *c, const char *patterns_path, const char *index_path)
{
    RET_IF(check_ctx(c));
    if (!c->passed) return fail(FFBF_ESTATE, "no pattern pass yet");
    const uint64_t words = (c->n_pat + 31) / 32;
    std::vector<uint32_t> bits(words);
    std::vector<uint64_t> off(c->n_pat + 1);
    if (words) HIP_TRY(hipMemcpyAsync(bits.data(), c->bitmap.p, words * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(off.data(), c->off.p, (c->n_pat + 1) * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<uint8_t> bytes(off[c->n_pat]);
    if (!bytes.empty()) {
        HIP_TRY(hipMemcpyAsync(bytes.data(), c->bytes.p, bytes.size(), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    File pf, xf;
    RET_IF(open_out(pf, patterns_path));
    if (index_path) RET_IF(open_out(xf, index_path));
    for (uint64_t w = 0; w < words; w++) {
        for (uint32_t m = bits[w]; m; m &= m - 1) {
            const uint64_t id = w * 32u + (uint64_t)__builtin_ctz(m);
            std::fwrite(bytes.data() + off[id], 1, off[id + 1] - off[id], pf.f);
            std::fputc('\n', pf.f);
            if (xf.f) std::fprintf(xf.f, "%llu\n", (unsigned long long)id);
        }
    }
    RET_IF(close_out(pf, patterns_path));
    if (xf.f) RET_IF(close_out(xf, index_path));
    return FFBF_OK;
}

int ffbf_write_candidate_lines(ffbf_ctx *c, const uint8_t *text, uint64_t n, const char *out_path)
{
    RET_IF(check_ctx(c));
    if (!c->scanned) return fail(FFBF_ESTATE, "no scan yet");
    std::vector<uint8_t> buf;
    const uint8_t *t = nullptr;
    uint64_t tn = 0;
    RET_IF(host_text(c, text, n, buf, &t, &tn));
    std::vector<uint64_t> cand(c->n_cand);
    uint64_t got = 0;
    RET_IF(ffbf_get_candidates(c, cand.data(), cand.size(), &got));  // sorted
    File out;
    RET_IF(open_out(out, out_path));
    // a candidate at offset o prints its line (the bytes between the '\n'
    // before o and the next '\n' at or after o) unless that line was printed
    // already or o is itself a line break
    int64_t prev_end = -1;
    for (uint64_t i = 0; i < got; i++) {
        const uint64_t o = cand[i];
        if ((int64_t)o <= prev_end || t[o] == '\n') continue;
        const void *pl = o ? memrchr(t, '\n', o) : nullptr;
        const uint64_t start = pl ? (uint64_t)((const uint8_t *)pl - t) + 1u : 0u;
        const void *nl = std::memchr(t + o, '\n', tn - o);
        const uint64_t end = nl ? (uint64_t)((const uint8_t *)nl - t) : tn;
        std::fwrite(t + start, 1, end - start, out.f);
        std::fputc('\n', out.f);
        prev_end = (int64_t)end;
    }
    return close_out(out, out_path);
}

}  // extern "C"
