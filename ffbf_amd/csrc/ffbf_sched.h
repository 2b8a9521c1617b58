// ffbf_sched.h -- the serving loop's scheduler, independent of the device
// (internal; host C++ only, no HIP).
//
// Requests are cut into pieces that fit one slot's text buffer: a request of
// at most `slot_bytes` is one piece; a longer one becomes pieces of
// slot_bytes with `overlap` bytes shared by consecutive pieces (the longest
// pattern of the request's filter minus one), each piece owning the window
// starts below its stride, so that every match of the whole text is found by
// exactly one piece.  `slots` worker threads each take, in arrival order,
// every pending piece that still fits their buffer and the HBM budget, hand
// them to the executor as one batch and complete them; a request completes
// when its last piece does, its matches sorted by (pattern id, offset).
// Cancelling drops a request's pending pieces; it then completes (with
// FFBF_ECANCELED) once none of its pieces is on the device.
#pragma once

#include <stdint.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ffbf.h"

namespace ffbf {

struct SchedRequest;

// A contiguous piece of a request's text: bytes [start, start + n) of it.
struct SchedPiece {
    SchedRequest *req = nullptr;
    uint32_t filter = 0;
    const uint8_t *text = nullptr;  // req's text + start
    uint64_t start = 0, n = 0;
    uint64_t own = 0;        // window starts this piece reports: local offsets below own
    uint64_t footprint = 0;  // device bytes held while in flight
    bool device_readable = false;  // the text lies in device-mapped host memory
    std::vector<ffbf_match> out;  // filled by the executor: matches at local offsets
};

// Runs pieces on the device.  run() is called from the slot's own thread
// and fills every piece's `out` (any order); it returns FFBF_OK or a status,
// with a message in err.
class SchedExecutor {
  public:
    virtual ~SchedExecutor() {}
    virtual int run(uint32_t slot, std::vector<SchedPiece *> &work, std::string &err) = 0;
};

struct SchedConfig {
    uint64_t hbm_budget_bytes = 0;  // device bytes of pieces in flight
    uint64_t slot_bytes = 0;        // text bytes per batch
    uint32_t slots = 0;
    uint32_t max_batch_requests = 0;  // pieces per batch
};

struct SchedRequest {
    uint64_t id = 0;
    uint32_t filter = 0;
    uint64_t n = 0;
    std::chrono::steady_clock::time_point t_submit, t_done;
    std::vector<SchedPiece> pieces;
    uint32_t left = 0;     // pieces not yet completed
    uint32_t running = 0;  // pieces handed to the executor and not yet back
    int status = FFBF_OK;
    std::string error;
    std::vector<ffbf_match> matches;
    bool done = false, claimed = false, canceled = false;
};

// Device bytes a piece of n text bytes holds while in flight: the text and
// the candidate records a scan of it may need (16 bytes per KiB), plus 4 KiB.
inline uint64_t piece_footprint(uint64_t n) { return n + (n >> 6) + 4096u; }

class Scheduler {
  public:
    // The executor must outlive the scheduler; start() spawns the slots.
    Scheduler(const SchedConfig &cfg, SchedExecutor *exec);
    ~Scheduler();
    void start();
    // Finishes everything submitted, then joins the slots.
    void shutdown();

    // overlap: bytes consecutive pieces share (at most slot_bytes / 2).
    // device_readable: the device may read the text at its host address
    // (passed on to every piece).
    int submit(uint32_t filter, const uint8_t *text, uint64_t n, uint64_t overlap, uint64_t *id, std::string &err,
               bool device_readable = false);
    // FFBF_OK: the request will complete with FFBF_ECANCELED; FFBF_ESTATE: it
    // had completed already (its result stands); FFBF_EINVAL: unknown id.
    int cancel(uint64_t id, std::string &err);
    int poll(uint64_t id, int *done, std::string &err);
    // Blocks until the request completes; releases its id.
    int wait(uint64_t id, std::vector<ffbf_match> &matches, double *latency_ms, std::string &err);
    void stats(ffbf_server_stats *out);
    const SchedConfig &config() const { return cfg_; }

  private:
    void slot_loop(uint32_t slot);
    void finish_locked(SchedRequest *r);

    SchedConfig cfg_;
    SchedExecutor *exec_;
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<SchedPiece *> pending_;
    std::unordered_map<uint64_t, SchedRequest *> reqs_;
    uint64_t next_id_ = 1;
    uint64_t in_flight_ = 0;
    bool stop_ = false;
    ffbf_server_stats stats_{};
};

}  // namespace ffbf
