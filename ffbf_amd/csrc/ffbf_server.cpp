// ffbf_server.cpp -- continuous-batching request loop over resident filters:
// the device side of the serving loop (ffbf_sched.cpp holds the scheduling).
//
// A server owns `slots` workers.  Each slot has its own copy of every filter
// (pattern set, F and prefilters copied device to device; its own R and
// survivor index), a batch object on its own stream and a device text buffer
// of max_batch_bytes.  The scheduler hands a free slot, in arrival order,
// every pending piece of request text that still fits its buffer and the
// server's HBM budget; the slot copies the pieces into its buffer straight
// from the requests' memory (pieces adjacent in host memory in one copy),
// runs one batched scan (ffbf_batch_run) and
// returns each piece's matches.  Requests that arrive while slots are busy
// wait and are admitted by the next slot to free up; several slots (four by
// default) work at once, so one batch's copies overlap another's scan.  A
// request longer than a slot's buffer is scanned in overlapping pieces (the
// overlap is the longest pattern of its filter minus one byte), each match
// reported by the piece owning its offset.  A request's matches are those of
// ffbf_run_device on its text alone: verification is per piece and piece
// borders never cut a match.
#include <atomic>
#include <cstring>
#include <memory>
#include <vector>

#include "ffbf_ctx.h"
#include "ffbf_sched.h"

using namespace ffbf;

namespace {

struct Slot {
    std::vector<ffbf_ctx *> filters;  // owned replicas
    ffbf_batch *batch = nullptr;
    uint8_t *d_text = nullptr;
    uint64_t cap = 0;
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> route;
    std::vector<ffbf_batch_match> bm;
    UploadArgs upload{};  // the copy kernel's arguments for the current launch
};

// Device-to-device copy of a built context: same parameters, patterns,
// F and prefilters; fresh R and survivor index.  max_len: the longest
// pattern of the set.
int clone_ctx(const ffbf_ctx *src, ffbf_ctx **out, uint64_t *max_len)
{
    RET_IF(src->split ? ffbf_create_split(&src->split_params, src->device, out)
                      : ffbf_create(&src->params, src->device, out));
    ffbf_ctx *c = *out;
    c->mode_req = src->mode_req;
    c->fold_log2_req = src->fold_log2_req;
    c->probes_req = src->probes_req;
    c->block_req = src->block_req;
    // pattern set
    std::vector<uint64_t> off(src->n_pat + 1);
    HIP_TRY(hipMemcpy(off.data(), src->off.p, off.size() * 8, hipMemcpyDeviceToHost));
    *max_len = 0;
    for (uint64_t i = 0; i < src->n_pat; i++) *max_len = std::max(*max_len, off[i + 1] - off[i]);
    std::vector<uint8_t> bytes(off.back());
    if (!bytes.empty()) HIP_TRY(hipMemcpy(bytes.data(), src->bytes.p, bytes.size(), hipMemcpyDeviceToHost));
    RET_IF(ffbf_set_patterns(c, bytes.data(), off.data(), src->n_pat));
    // F as built (or loaded) by the source, then the same mode choice.  The
    // copy is ordered on the clone's stream before the folds computed from
    // it (a device-to-device hipMemcpy may return before it completes).
    HIP_TRY(hipStreamSynchronize(src->stream));
    HIP_TRY(hipMemcpyAsync(c->F.p, src->F.p, src->f_words * 4, hipMemcpyDeviceToDevice, c->stream));
    c->f_set_bits = src->f_set_bits;
    c->built = true;
    c->mode_req = src->mode_req;
    RET_IF(ffbf_set_scan_mode(c, src->mode_req, src->fold_log2_req));
    return ffbf_synchronize(c);
}

// Whether the device may read text[0, n) at its host address: the bytes lie
// wholly in one host allocation mapped to the device at the same address
// (hipHostMalloc memory).  Anything else (pageable memory, device memory,
// a range running past its allocation) answers false.
bool device_readable(int device, const uint8_t *text, uint64_t n)
{
    if (!text || !n) return false;
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, text) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (at.type != hipMemoryTypeHost || at.devicePointer != (const void *)text || at.device != device) return false;
    void *base = nullptr;
    size_t size = 0;
    if (hipPointerGetAttribute(&base, HIP_POINTER_ATTRIBUTE_RANGE_START_ADDR, (hipDeviceptr_t)text) != hipSuccess ||
        hipPointerGetAttribute(&size, HIP_POINTER_ATTRIBUTE_RANGE_SIZE, (hipDeviceptr_t)text) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    const uintptr_t b = (uintptr_t)base, t = (uintptr_t)text;
    return b && t >= b && n <= size && t - b <= size - n;
}

class GpuExecutor : public SchedExecutor {
  public:
    std::vector<Slot> slots;
    int device = 0;
    bool upload_kernel = false;
    std::atomic<uint64_t> kernel_bytes{0};

    ~GpuExecutor() override
    {
        for (Slot &sl : slots) {
            if (sl.batch) ffbf_batch_destroy(sl.batch);
            for (ffbf_ctx *c : sl.filters) ffbf_destroy(c);
            if (sl.d_text) (void)hipFree(sl.d_text);
        }
    }

    // Copies the pieces back to back into the slot's buffer on the batch's
    // stream, scans them as one batch, and returns each piece's matches.
    int run(uint32_t k, std::vector<SchedPiece *> &work, std::string &err) override
    {
        Slot &sl = slots[k];
        (void)hipSetDevice(device);
        sl.offsets.assign(work.size() + 1, 0);
        sl.route.resize(work.size());
        const hipStream_t stream = batch_stream(sl.batch);
        // with FFBF_UPLOAD_KERNEL, pieces in device-mapped host memory go by
        // the copy kernel, up to kUploadMax per launch; the rest (all pieces
        // otherwise) by the DMA engine
        UploadArgs &ua = sl.upload;
        ua.dst = sl.d_text;
        ua.n = 0;
        ua.total_segs = 0;
        uint64_t by_kernel = 0;
        auto flush = [&]() -> bool {
            if (!ua.n) return true;
            ua.seg_start[ua.n] = ua.total_segs;
            const hipError_t e = launch_upload(ua, stream);
            ua.n = 0;
            ua.total_segs = 0;
            if (e == hipSuccess) return true;
            err = std::string("text upload kernel: ") + hipGetErrorString(e);
            return false;
        };
        // DMA copies of pieces adjacent in host memory (a caller's receive
        // ring, consecutive requests of one buffer) are merged: they are
        // adjacent in the slot's buffer too, and each copy command has a
        // fixed cost
        const uint8_t *run_src = nullptr;
        uint64_t run_at = 0, run_n = 0;
        auto flush_run = [&]() -> bool {
            if (!run_n) return true;
            const hipError_t e = hipMemcpyAsync(sl.d_text + run_at, run_src, run_n, hipMemcpyHostToDevice, stream);
            run_n = 0;
            if (e == hipSuccess) return true;
            err = std::string("text upload: ") + hipGetErrorString(e);
            return false;
        };
        for (size_t i = 0; i < work.size(); i++) {
            const uint64_t n = work[i]->n, at = sl.offsets[i];
            if (n && upload_kernel && work[i]->device_readable) {
                ua.src[ua.n] = (uint64_t)(uintptr_t)work[i]->text;
                ua.dst_off[ua.n] = at;
                ua.len[ua.n] = n;
                ua.seg_start[ua.n] = ua.total_segs;
                ua.total_segs += upload_segs(at, n);
                by_kernel += n;
                if (++ua.n == kUploadMax && !flush()) return FFBF_EDEVICE;
            } else if (n) {
                if (run_n && run_src + run_n == work[i]->text && run_at + run_n == at) {
                    run_n += n;
                } else {
                    if (!flush_run()) return FFBF_EDEVICE;
                    run_src = work[i]->text;
                    run_at = at;
                    run_n = n;
                }
            }
            sl.offsets[i + 1] = at + n;
            sl.route[i] = work[i]->filter;
        }
        if (!flush() || !flush_run()) return FFBF_EDEVICE;
        kernel_bytes += by_kernel;
        uint64_t nm = 0;
        int st = ffbf_batch_set_texts(sl.batch, sl.d_text, sl.offsets.data(), sl.route.data(), (uint32_t)work.size());
        if (st == FFBF_OK) st = ffbf_batch_run(sl.batch, &nm);
        if (st == FFBF_OK) {
            sl.bm.resize(nm);
            st = ffbf_batch_get_matches(sl.batch, sl.bm.data(), nm, &nm);
        }
        if (st != FFBF_OK) {
            err = ffbf_last_error();
            return st;
        }
        for (uint64_t i = 0; i < nm; i++) {
            const ffbf_batch_match &m = sl.bm[i];
            ffbf_match x;
            x.pattern_id = m.pattern_id;
            x.reserved = 0;
            x.offset = m.offset;
            work[m.text_id]->out.push_back(x);
        }
        return FFBF_OK;
    }
};

}  // namespace

struct ffbf_server {
    ffbf_server_config cfg{};
    std::vector<uint64_t> overlap;  // per filter: its longest pattern - 1 (at least W - 1)
    std::unique_ptr<GpuExecutor> exec;
    std::unique_ptr<Scheduler> sched;
};

extern "C" {

int ffbf_server_create(ffbf_ctx *const *filters, uint32_t n_filters, const ffbf_server_config *config,
                       ffbf_server **out)
{
    if (!filters || !n_filters || !out) return fail(FFBF_EINVAL, "need at least one filter");
    *out = nullptr;
    for (uint32_t f = 0; f < n_filters; f++) {
        if (!filters[f]) return fail(FFBF_EINVAL, "null filter");
        // the slots' replicas copy F as built: the sources must have one
        if (!filters[f]->built || !filters[f]->patterns)
            return fail(FFBF_ESTATE, "every filter must have patterns and a built filter");
    }
    std::unique_ptr<ffbf_server> s(new (std::nothrow) ffbf_server());
    if (!s) return fail(FFBF_ENOMEM, "server allocation");
    if (config) s->cfg = *config;
    if (!s->cfg.hbm_budget_bytes) s->cfg.hbm_budget_bytes = 4ull << 30;
    // four slots of at most 128 MiB: a batch's scan overlaps the next
    // batches' copies, which bound the loop (measured: 89% of the pinned
    // host-to-device rate, against 76% with two slots of budget/2)
    if (!s->cfg.slots) s->cfg.slots = 4;
    if (!s->cfg.max_batch_requests) s->cfg.max_batch_requests = 512;
    if (!s->cfg.max_batch_bytes)
        s->cfg.max_batch_bytes = std::min<uint64_t>(s->cfg.hbm_budget_bytes / s->cfg.slots, 128ull << 20);
    if (s->cfg.upload > FFBF_UPLOAD_KERNEL) return fail(FFBF_EINVAL, "unknown upload method");
    // the copy kernel addresses a piece with one 32-bit buffer range
    if (s->cfg.upload == FFBF_UPLOAD_KERNEL && s->cfg.max_batch_bytes > (1ull << 31))
        return fail(FFBF_EINVAL, "the upload kernel needs max_batch_bytes of at most 2 GiB");
    s->exec.reset(new (std::nothrow) GpuExecutor());
    if (!s->exec) return fail(FFBF_ENOMEM, "server allocation");
    s->exec->device = filters[0]->device;
    s->exec->upload_kernel = s->cfg.upload == FFBF_UPLOAD_KERNEL;
    s->exec->slots.resize(s->cfg.slots);
    s->overlap.assign(n_filters, 0);
    for (uint32_t k = 0; k < s->cfg.slots; k++) {
        Slot &sl = s->exec->slots[k];
        for (uint32_t f = 0; f < n_filters; f++) {
            ffbf_ctx *c = nullptr;
            uint64_t max_len = 0;
            const int st = clone_ctx(filters[f], &c, &max_len);
            if (c) sl.filters.push_back(c);
            RET_IF(st);  // the executor's destructor releases what was made
            const uint64_t W = filters[f]->params.window;
            s->overlap[f] = std::max<uint64_t>(max_len, W) - 1u;
        }
        RET_IF(ffbf_batch_create(sl.filters.data(), n_filters, &sl.batch));
        RET_IF(ffbf_batch_set_timing(sl.batch, 0));  // ~10 us of events per batch nobody reads
        RET_IF(batch_reserve(sl.batch, s->cfg.max_batch_requests, s->cfg.max_batch_bytes));
        sl.cap = s->cfg.max_batch_bytes;
        if (hipMalloc((void **)&sl.d_text, std::max<uint64_t>(sl.cap, 16)) != hipSuccess) {
            (void)hipGetLastError();
            sl.d_text = nullptr;
            return fail(FFBF_ENOMEM, "server slot buffers");
        }
        // one batch now, a zero text per filter: the slot's graph is captured
        // and instantiated here (its buffers never move, so requests replay
        // it) instead of in the first request's latency
        const uint64_t wn = std::min<uint64_t>(4096, sl.cap / n_filters) & ~15ull;
        if (wn >= 64) {
            HIP_TRY(launch_fill32((uint32_t *)sl.d_text, wn * n_filters / 4u, 0u, batch_stream(sl.batch)));
            std::vector<uint64_t> woff(n_filters + 1);
            std::vector<uint32_t> wroute(n_filters);
            for (uint32_t f = 0; f <= n_filters; f++) woff[f] = (uint64_t)f * wn;
            for (uint32_t f = 0; f < n_filters; f++) wroute[f] = f;
            uint64_t nm = 0;
            RET_IF(ffbf_batch_set_texts(sl.batch, sl.d_text, woff.data(), wroute.data(), n_filters));
            RET_IF(ffbf_batch_run(sl.batch, &nm));
        }
    }
    SchedConfig sc;
    sc.hbm_budget_bytes = s->cfg.hbm_budget_bytes;
    sc.slot_bytes = s->cfg.max_batch_bytes;
    sc.slots = s->cfg.slots;
    sc.max_batch_requests = s->cfg.max_batch_requests;
    s->sched.reset(new (std::nothrow) Scheduler(sc, s->exec.get()));
    if (!s->sched) return fail(FFBF_ENOMEM, "server allocation");
    s->sched->start();
    *out = s.release();
    return FFBF_OK;
}

void ffbf_server_destroy(ffbf_server *s)
{
    if (!s) return;
    s->sched.reset();  // finishes every submitted request, joins the slots
    s->exec.reset();
    delete s;
}

int ffbf_server_submit(ffbf_server *s, uint32_t filter, const uint8_t *text, uint64_t n, uint64_t *request_id)
{
    if (!s || !request_id) return fail(FFBF_EINVAL, "null argument");
    if (filter >= s->overlap.size()) return fail(FFBF_EINVAL, "filter id out of range");
    if (!text && n) return fail(FFBF_EINVAL, "null text");
    std::string err;
    const bool readable = s->exec->upload_kernel && device_readable(s->exec->device, text, n);
    const int st = s->sched->submit(filter, text, n, s->overlap[filter], request_id, err, readable);
    return st == FFBF_OK ? st : fail(st, err);
}

int ffbf_server_submit_many(ffbf_server *s, uint32_t n, const uint32_t *filters, const uint8_t *const *texts,
                            const uint64_t *sizes, uint64_t *request_ids)
{
    if (!s || (n && (!filters || !texts || !sizes || !request_ids))) return fail(FFBF_EINVAL, "null argument");
    for (uint32_t i = 0; i < n; i++) RET_IF(ffbf_server_submit(s, filters[i], texts[i], sizes[i], &request_ids[i]));
    return FFBF_OK;
}

int ffbf_server_cancel(ffbf_server *s, uint64_t request_id)
{
    if (!s) return fail(FFBF_EINVAL, "null argument");
    std::string err;
    const int st = s->sched->cancel(request_id, err);
    return st == FFBF_OK ? st : fail(st, err);
}

int ffbf_server_poll(ffbf_server *s, uint64_t request_id, int *done)
{
    if (!s || !done) return fail(FFBF_EINVAL, "null argument");
    std::string err;
    const int st = s->sched->poll(request_id, done, err);
    return st == FFBF_OK ? st : fail(st, err);
}

int ffbf_server_wait(ffbf_server *s, uint64_t request_id, ffbf_match **matches, uint64_t *n, double *latency_ms)
{
    if (!s || !matches || !n) return fail(FFBF_EINVAL, "null argument");
    *matches = nullptr;
    *n = 0;
    std::vector<ffbf_match> got;
    std::string err;
    const int st = s->sched->wait(request_id, got, latency_ms, err);
    if (st != FFBF_OK) return fail(st, err);
    ffbf_match *m = (ffbf_match *)std::malloc(std::max<size_t>(got.size(), 1) * sizeof(ffbf_match));
    if (!m) return fail(FFBF_ENOMEM, "match list allocation");
    if (!got.empty()) std::memcpy(m, got.data(), got.size() * sizeof(ffbf_match));
    *matches = m;
    *n = got.size();
    return FFBF_OK;
}

int ffbf_server_get_stats(ffbf_server *s, ffbf_server_stats *out)
{
    if (!s || !out) return fail(FFBF_EINVAL, "null argument");
    s->sched->stats(out);
    out->kernel_upload_bytes = s->exec->kernel_bytes.load();
    return FFBF_OK;
}

}  // extern "C"
