// ffbf_sched.cpp -- the serving loop's scheduler (host only, no HIP): pieces,
// admission by slot capacity and HBM budget, cancellation, completion.
#include "ffbf_sched.h"

#include <algorithm>

#include "ffbf_hostsort.h"

namespace ffbf {

Scheduler::Scheduler(const SchedConfig &cfg, SchedExecutor *exec) : cfg_(cfg), exec_(exec) {}

Scheduler::~Scheduler() { shutdown(); }

void Scheduler::start()
{
    for (uint32_t k = 0; k < cfg_.slots; k++) threads_.emplace_back(&Scheduler::slot_loop, this, k);
}

void Scheduler::shutdown()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (std::thread &t : threads_)
        if (t.joinable()) t.join();
    threads_.clear();
    for (auto &kv : reqs_) delete kv.second;
    reqs_.clear();
}

int Scheduler::submit(uint32_t filter, const uint8_t *text, uint64_t n, uint64_t overlap, uint64_t *id,
                      std::string &err, bool device_readable)
{
    const uint64_t cap = cfg_.slot_bytes;
    if (n > cap && (overlap == 0 || 2 * overlap > cap)) {
        err = "request longer than a slot needs an overlap of 1 to slot_bytes/2 bytes";
        return FFBF_EINVAL;
    }
    if (piece_footprint(std::min(n, cap)) > cfg_.hbm_budget_bytes) {
        std::lock_guard<std::mutex> lk(mu_);
        stats_.rejected++;
        err = "request piece larger than the HBM budget";
        return FFBF_ECAPACITY;
    }
    SchedRequest *r = new (std::nothrow) SchedRequest();
    if (!r) {
        err = "request allocation";
        return FFBF_ENOMEM;
    }
    r->filter = filter;
    r->n = n;
    // pieces: [start, start + min(cap, n - start)); every one but the last
    // owns its first cap - overlap window starts, the next one starting there
    const uint64_t stride = n > cap ? cap - overlap : cap;
    uint64_t start = 0;
    for (;;) {
        SchedPiece p;
        p.req = r;
        p.filter = filter;
        p.start = start;
        p.n = std::min(cap, n - start);
        p.text = text ? text + start : nullptr;
        const bool last = start + p.n >= n;
        p.own = last ? p.n : stride;
        p.footprint = piece_footprint(p.n);
        p.device_readable = device_readable;
        r->pieces.push_back(std::move(p));
        if (last) break;
        start += stride;
    }
    r->left = (uint32_t)r->pieces.size();
    r->t_submit = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> lk(mu_);
        r->id = next_id_++;
        reqs_[r->id] = r;
        for (SchedPiece &p : r->pieces) pending_.push_back(&p);
        stats_.submitted++;
    }
    cv_work_.notify_all();
    *id = r->id;
    return FFBF_OK;
}

// Called with mu_ held once r's last piece is back (or dropped).
void Scheduler::finish_locked(SchedRequest *r)
{
    if (r->canceled) {
        r->status = FFBF_ECANCELED;
        r->error = "request cancelled";
        r->matches.clear();
    } else if (r->status == FFBF_OK) {
        std::vector<ffbf_match> tmp;
        sort_by_major_offset(r->matches.data(), r->matches.size(), [](const ffbf_match &m) { return m.pattern_id; },
                             tmp);
    } else {
        r->matches.clear();
    }
    r->t_done = std::chrono::steady_clock::now();
    r->done = true;
    stats_.completed++;
    if (r->status == FFBF_OK) stats_.bytes += r->n;
}

int Scheduler::cancel(uint64_t id, std::string &err)
{
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) {
        err = "unknown request";
        return FFBF_EINVAL;
    }
    SchedRequest *r = it->second;
    if (r->done) {
        err = "request already completed";
        return FFBF_ESTATE;
    }
    r->canceled = true;
    const size_t before = pending_.size();
    pending_.erase(std::remove_if(pending_.begin(), pending_.end(), [&](SchedPiece *p) { return p->req == r; }),
                   pending_.end());
    r->left -= (uint32_t)(before - pending_.size());
    if (r->left == 0) finish_locked(r);
    lk.unlock();
    cv_done_.notify_all();
    cv_work_.notify_all();
    return FFBF_OK;
}

int Scheduler::poll(uint64_t id, int *done, std::string &err)
{
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) {
        err = "unknown request";
        return FFBF_EINVAL;
    }
    *done = it->second->done ? 1 : 0;
    return FFBF_OK;
}

int Scheduler::wait(uint64_t id, std::vector<ffbf_match> &matches, double *latency_ms, std::string &err)
{
    SchedRequest *r = nullptr;
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = reqs_.find(id);
        if (it == reqs_.end()) {
            err = "unknown request";
            return FFBF_EINVAL;
        }
        r = it->second;
        if (r->claimed) {
            err = "request already being waited for";
            return FFBF_EINVAL;
        }
        r->claimed = true;
        cv_done_.wait(lk, [&] { return r->done; });
        reqs_.erase(id);
    }
    const int st = r->status;
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (st == FFBF_OK)
        matches.swap(r->matches);
    else
        err = r->error;
    delete r;
    return st;
}

void Scheduler::stats(ffbf_server_stats *out)
{
    std::lock_guard<std::mutex> lk(mu_);
    *out = stats_;
    out->pending = pending_.size();
    out->in_flight_bytes = in_flight_;
}

void Scheduler::slot_loop(uint32_t slot)
{
    std::vector<SchedPiece *> work;
    for (;;) {
        work.clear();
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_work_.wait(lk, [&] { return stop_ || !pending_.empty(); });
            if (stop_ && pending_.empty()) return;
            uint64_t bytes = 0;
            while (!pending_.empty() && work.size() < cfg_.max_batch_requests) {
                SchedPiece *p = pending_.front();
                if (bytes + p->n > cfg_.slot_bytes) break;
                if (in_flight_ + p->footprint > cfg_.hbm_budget_bytes) break;
                pending_.pop_front();
                in_flight_ += p->footprint;
                bytes += p->n;
                p->req->running++;
                work.push_back(p);
            }
            if (work.empty()) {
                // the budget is held by other slots: wait until the oldest
                // piece fits (every piece fits an empty budget, and each
                // completion notifies)
                cv_work_.wait(lk, [&] {
                    return pending_.empty() || in_flight_ + pending_.front()->footprint <= cfg_.hbm_budget_bytes;
                });
                continue;
            }
        }
        std::string err;
        for (SchedPiece *p : work) p->out.clear();
        const int st = exec_->run(slot, work, err);
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (SchedPiece *p : work) {
                SchedRequest *r = p->req;
                if (st != FFBF_OK && r->status == FFBF_OK) {
                    r->status = st;
                    r->error = err;
                }
                if (st == FFBF_OK && !r->canceled)
                    for (const ffbf_match &m : p->out)
                        if (m.offset < p->own) {
                            ffbf_match g = m;
                            g.offset += p->start;
                            r->matches.push_back(g);
                        }
                p->out.clear();
                p->out.shrink_to_fit();
                in_flight_ -= p->footprint;
                r->running--;
                if (--r->left == 0) finish_locked(r);
            }
            stats_.batches++;
        }
        cv_done_.notify_all();
        cv_work_.notify_all();
    }
}

}  // namespace ffbf
