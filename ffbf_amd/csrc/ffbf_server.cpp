// ffbf_server.cpp -- continuous-batching request loop over resident filters.
//
// A server owns `slots` workers.  Each slot has its own copy of every filter
// (pattern set, F and prefilters copied device to device; its own R and
// survivor index), a batch object on its own stream and a device text buffer
// of max_batch_bytes.  A free slot takes, in arrival order, every pending
// request that still fits its text buffer and the server's HBM budget, copies
// the texts into its buffer straight from the requests' memory, runs one
// batched scan (ffbf_batch_run) and hands every request its own matches.  Requests that
// arrive while slots are busy wait in the queue and are admitted by the next
// slot to free up; several slots (four by default) work at once, so one
// batch's copies overlap another's scan.  A request's matches are those of
// ffbf_run_device on its text alone: verification is per text.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ffbf_ctx.h"

using namespace ffbf;
using clk = std::chrono::steady_clock;

namespace {

struct Request {
    uint64_t id = 0;
    uint32_t filter = 0;
    const uint8_t *text = nullptr;
    uint64_t n = 0;
    uint64_t footprint = 0;
    clk::time_point t_submit, t_done;
    int status = FFBF_OK;
    std::string error;
    std::vector<ffbf_match> matches;
    bool done = false;
    bool claimed = false;  // a wait holds it (one waiter per request)
};

struct Slot {
    std::vector<ffbf_ctx *> filters;  // owned replicas
    ffbf_batch *batch = nullptr;
    uint8_t *d_text = nullptr;
    uint64_t cap = 0;
    std::thread th;
};

// Device-to-device copy of a built context: same parameters, patterns,
// F and prefilters; fresh R and survivor index.
int clone_ctx(const ffbf_ctx *src, ffbf_ctx **out)
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

}  // namespace

struct ffbf_server {
    ffbf_server_config cfg{};
    uint32_t n_filters = 0;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::unordered_map<uint64_t, Request *> reqs;
    uint64_t next_id = 1;
    uint64_t in_flight = 0;  // footprint admitted and not finished
    bool stop = false;
    ffbf_server_stats stats{};
};

namespace {

void serve_slot(ffbf_server *s, Slot *slot)
{
    (void)hipSetDevice(slot->filters[0]->device);
    std::vector<Request *> work;
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> route;
    std::vector<ffbf_batch_match> bm;
    for (;;) {
        work.clear();
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
            if (s->stop && s->pending.empty()) return;
            uint64_t bytes = 0;
            while (!s->pending.empty() && work.size() < s->cfg.max_batch_requests) {
                Request *r = s->pending.front();
                if (bytes + r->n > slot->cap) break;
                if (s->in_flight + r->footprint > s->cfg.hbm_budget_bytes) break;
                s->pending.pop_front();
                s->in_flight += r->footprint;
                bytes += r->n;
                work.push_back(r);
            }
            if (work.empty()) {
                // the budget is held by other slots: wait until the oldest
                // request fits (every request fits an empty budget)
                s->cv_work.wait(lk, [&] {
                    return s->stop || s->pending.empty() ||
                           s->in_flight + s->pending.front()->footprint <= s->cfg.hbm_budget_bytes;
                });
                continue;
            }
        }
        // copy the texts back to back into the slot's device buffer, on the
        // batch's stream (straight from the caller's memory: pinned memory
        // goes at full link speed), then scan them as one batch
        offsets.assign(work.size() + 1, 0);
        route.resize(work.size());
        int st = FFBF_OK;
        std::string err;
        uint64_t nm = 0;
        const hipStream_t stream = batch_stream(slot->batch);
        for (size_t i = 0; i < work.size(); i++) {
            if (work[i]->n && st == FFBF_OK &&
                hipMemcpyAsync(slot->d_text + offsets[i], work[i]->text, work[i]->n, hipMemcpyHostToDevice, stream) !=
                    hipSuccess) {
                st = FFBF_EDEVICE;
                err = std::string("text upload: ") + hipGetErrorString(hipGetLastError());
            }
            offsets[i + 1] = offsets[i] + work[i]->n;
            route[i] = work[i]->filter;
        }
        if (st == FFBF_OK) st = ffbf_batch_set_texts(slot->batch, slot->d_text, offsets.data(), route.data(), (uint32_t)work.size());
        if (st == FFBF_OK) st = ffbf_batch_run(slot->batch, &nm);
        if (st == FFBF_OK) {
            bm.resize(nm);
            st = ffbf_batch_get_matches(slot->batch, bm.data(), nm, &nm);
        }
        if (st != FFBF_OK && err.empty()) err = ffbf_last_error();
        if (st == FFBF_OK) {
            for (const ffbf_batch_match &m : bm) {
                ffbf_match x;
                x.pattern_id = m.pattern_id;
                x.reserved = 0;
                x.offset = m.offset;
                work[m.text_id]->matches.push_back(x);
            }
        }
        const auto now = clk::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (Request *r : work) {
                r->status = st;
                r->error = err;
                r->t_done = now;
                r->done = true;
                s->in_flight -= r->footprint;
                s->stats.completed++;
                s->stats.bytes += r->n;
            }
            s->stats.batches++;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

void shutdown(ffbf_server *s)
{
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (Slot &sl : s->slots)
        if (sl.th.joinable()) sl.th.join();
    for (Slot &sl : s->slots) {
        if (sl.batch) ffbf_batch_destroy(sl.batch);
        for (ffbf_ctx *c : sl.filters) ffbf_destroy(c);
        if (sl.d_text) (void)hipFree(sl.d_text);
    }
    for (auto &kv : s->reqs) delete kv.second;
    s->reqs.clear();
}

}  // namespace

extern "C" {

int ffbf_server_create(ffbf_ctx *const *filters, uint32_t n_filters, const ffbf_server_config *config,
                       ffbf_server **out)
{
    if (!filters || !n_filters || !out) return fail(FFBF_EINVAL, "need at least one filter");
    *out = nullptr;
    ffbf_server *s = new (std::nothrow) ffbf_server();
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
    s->n_filters = n_filters;
    s->slots.resize(s->cfg.slots);
    int st = FFBF_OK;
    for (uint32_t k = 0; k < s->cfg.slots && st == FFBF_OK; k++) {
        Slot &sl = s->slots[k];
        for (uint32_t f = 0; f < n_filters && st == FFBF_OK; f++) {
            ffbf_ctx *c = nullptr;
            st = clone_ctx(filters[f], &c);
            if (c) sl.filters.push_back(c);
        }
        if (st == FFBF_OK) st = ffbf_batch_create(sl.filters.data(), n_filters, &sl.batch);
        if (st == FFBF_OK) st = batch_reserve(sl.batch, s->cfg.max_batch_requests, s->cfg.max_batch_bytes);
        if (st == FFBF_OK) {
            sl.cap = s->cfg.max_batch_bytes;
            if (hipMalloc((void **)&sl.d_text, std::max<uint64_t>(sl.cap, 16)) != hipSuccess) {
                (void)hipGetLastError();
                st = fail(FFBF_ENOMEM, "server slot buffers");
            }
        }
    }
    if (st != FFBF_OK) {
        const std::string msg = ffbf_last_error();
        shutdown(s);
        delete s;
        fail(st, msg);
        return st;
    }
    for (Slot &sl : s->slots) sl.th = std::thread(serve_slot, s, &sl);
    *out = s;
    return FFBF_OK;
}

void ffbf_server_destroy(ffbf_server *s)
{
    if (!s) return;
    shutdown(s);
    delete s;
}

int ffbf_server_submit(ffbf_server *s, uint32_t filter, const uint8_t *text, uint64_t n, uint64_t *request_id)
{
    if (!s || !request_id) return fail(FFBF_EINVAL, "null argument");
    if (filter >= s->n_filters) return fail(FFBF_EINVAL, "filter id out of range");
    if (!text && n) return fail(FFBF_EINVAL, "null text");
    // device bytes the request holds while in flight: its text, plus the
    // candidate records a scan of it may need (16 B per KiB reserved)
    const uint64_t footprint = n + (n >> 6) + 4096;
    if (n > s->cfg.max_batch_bytes || footprint > s->cfg.hbm_budget_bytes) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.rejected++;
        return fail(FFBF_ECAPACITY, "request larger than a batch slot or the HBM budget");
    }
    Request *r = new (std::nothrow) Request();
    if (!r) return fail(FFBF_ENOMEM, "request allocation");
    r->filter = filter;
    r->text = text;
    r->n = n;
    r->footprint = footprint;
    r->t_submit = clk::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        s->reqs[r->id] = r;
        s->pending.push_back(r);
        s->stats.submitted++;
    }
    s->cv_work.notify_one();
    *request_id = r->id;
    return FFBF_OK;
}

int ffbf_server_submit_many(ffbf_server *s, uint32_t n, const uint32_t *filters, const uint8_t *const *texts,
                            const uint64_t *sizes, uint64_t *request_ids)
{
    if (!s || (n && (!filters || !texts || !sizes || !request_ids))) return fail(FFBF_EINVAL, "null argument");
    for (uint32_t i = 0; i < n; i++) RET_IF(ffbf_server_submit(s, filters[i], texts[i], sizes[i], &request_ids[i]));
    return FFBF_OK;
}

int ffbf_server_poll(ffbf_server *s, uint64_t request_id, int *done)
{
    if (!s || !done) return fail(FFBF_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request_id);
    if (it == s->reqs.end()) return fail(FFBF_EINVAL, "unknown request");
    *done = it->second->done ? 1 : 0;
    return FFBF_OK;
}

int ffbf_server_wait(ffbf_server *s, uint64_t request_id, ffbf_match **matches, uint64_t *n, double *latency_ms)
{
    if (!s || !matches || !n) return fail(FFBF_EINVAL, "null argument");
    *matches = nullptr;
    *n = 0;
    Request *r = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(request_id);
        if (it == s->reqs.end()) return fail(FFBF_EINVAL, "unknown request");
        r = it->second;
        if (r->claimed) return fail(FFBF_EINVAL, "request already being waited for");
        r->claimed = true;
        s->cv_done.wait(lk, [&] { return r->done; });
        s->reqs.erase(request_id);
    }
    const int st = r->status;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (st == FFBF_OK) {
        ffbf_match *m = (ffbf_match *)std::malloc(std::max<size_t>(r->matches.size(), 1) * sizeof(ffbf_match));
        if (!m) {
            delete r;
            return fail(FFBF_ENOMEM, "match list allocation");
        }
        if (!r->matches.empty()) std::memcpy(m, r->matches.data(), r->matches.size() * sizeof(ffbf_match));
        *matches = m;
        *n = r->matches.size();
    } else {
        fail(st, r->error);
    }
    delete r;
    return st;
}

int ffbf_server_get_stats(ffbf_server *s, ffbf_server_stats *out)
{
    if (!s || !out) return fail(FFBF_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->pending = s->pending.size();
    out->in_flight_bytes = s->in_flight;
    return FFBF_OK;
}

}  // extern "C"
