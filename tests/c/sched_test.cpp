// The serving scheduler (ffbf_amd/csrc/ffbf_sched.cpp) against a fake
// executor on the CPU: a naive matcher stands in for the device, and the
// executor checks the admission rules (slot bytes, pieces per batch, HBM
// budget in flight) and can fail batches or hold them until released.
// Built with g++ under AddressSanitizer + UndefinedBehaviorSanitizer, and
// under ThreadSanitizer, by tests/test_sanitize.py.  Exits 0 when every check
// passes.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ffbf_sched.h"

using namespace ffbf;

static std::atomic<int> g_failures{0};
#define CHECK(c)                                                               \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            g_failures++;                                                      \
        }                                                                      \
    } while (0)

// Every (pattern id, offset) with text[o .. o+|p|) == p, sorted.
static std::vector<ffbf_match> naive(const std::vector<std::string> &pats, const uint8_t *t, uint64_t n)
{
    std::vector<ffbf_match> out;
    for (uint32_t id = 0; id < pats.size(); id++) {
        const std::string &p = pats[id];
        for (uint64_t o = 0; o + p.size() <= n; o++)
            if (std::memcmp(t + o, p.data(), p.size()) == 0) out.push_back(ffbf_match{id, 0, o});
    }
    return out;
}

struct FakeExecutor : SchedExecutor {
    SchedConfig cfg;
    std::vector<std::vector<std::string>> filters;
    std::atomic<uint64_t> in_flight{0}, max_in_flight{0}, batches{0};
    std::atomic<int> fail_batch{-1};  // batch number that returns FFBF_ENOMEM
    // gate: while closed, run() waits (pieces stay on the "device")
    std::mutex gm;
    std::condition_variable gcv;
    bool gate_open = true;
    std::atomic<int> waiting{0};

    int run(uint32_t slot, std::vector<SchedPiece *> &work, std::string &err) override
    {
        CHECK(slot < cfg.slots);
        CHECK(!work.empty() && work.size() <= cfg.max_batch_requests);
        uint64_t bytes = 0, fp = 0;
        for (SchedPiece *p : work) {
            bytes += p->n;
            fp += p->footprint;
            CHECK(p->out.empty());
            CHECK(p->own <= p->n);
        }
        CHECK(bytes <= cfg.slot_bytes);
        const uint64_t now = in_flight += fp;
        CHECK(now <= cfg.hbm_budget_bytes);
        uint64_t m = max_in_flight.load();
        while (now > m && !max_in_flight.compare_exchange_weak(m, now)) {
        }
        {
            std::unique_lock<std::mutex> lk(gm);
            waiting++;
            gcv.wait(lk, [&] { return gate_open; });
            waiting--;
        }
        const int b = (int)batches++;
        int st = FFBF_OK;
        if (b == fail_batch.load()) {
            err = "injected allocation failure";
            st = FFBF_ENOMEM;
        } else {
            for (SchedPiece *p : work) p->out = naive(filters[p->filter], p->text, p->n);
        }
        in_flight -= fp;
        return st;
    }
    void close_gate()
    {
        std::lock_guard<std::mutex> lk(gm);
        gate_open = false;
    }
    void open_gate()
    {
        {
            std::lock_guard<std::mutex> lk(gm);
            gate_open = true;
        }
        gcv.notify_all();
    }
};

static std::vector<std::string> make_patterns(std::mt19937_64 &rng, int n, int lmin, int lmax)
{
    std::vector<std::string> v;
    for (int i = 0; i < n; i++) {
        std::string s((size_t)(lmin + rng() % (uint64_t)(lmax - lmin + 1)), 'a');
        for (char &c : s) c = "ab"[rng() % 2];
        v.push_back(s);
    }
    return v;
}

static std::vector<uint8_t> make_text(std::mt19937_64 &rng, uint64_t n)
{
    std::vector<uint8_t> t(n);
    for (uint8_t &c : t) c = (uint8_t)"abc"[rng() % 3];
    return t;
}

static uint64_t overlap_of(const std::vector<std::string> &pats)
{
    size_t m = 1;
    for (const std::string &p : pats) m = std::max(m, p.size());
    return m - 1;
}

static bool same(const std::vector<ffbf_match> &a, const std::vector<ffbf_match> &b)
{
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); i++)
        if (a[i].pattern_id != b[i].pattern_id || a[i].offset != b[i].offset) return false;
    return true;
}

// Many requests of 0..3.5 slots, several filters, from two submitting threads:
// every result equals the naive matcher on the whole text; the budget holds.
static void test_pieces_and_budget()
{
    std::mt19937_64 rng(1);
    FakeExecutor ex;
    ex.cfg.hbm_budget_bytes = 3 * piece_footprint(4096);
    ex.cfg.slot_bytes = 4096;
    ex.cfg.slots = 4;
    ex.cfg.max_batch_requests = 5;
    for (int f = 0; f < 3; f++) ex.filters.push_back(make_patterns(rng, 12, 3 + f, 9 + 3 * f));
    const int N = 120;
    std::vector<std::vector<uint8_t>> texts(N);
    std::vector<uint32_t> filt(N);
    for (int i = 0; i < N; i++) {
        const uint64_t n = i % 10 == 0 ? 0 : rng() % (uint64_t)(3.5 * 4096);
        texts[i] = make_text(rng, n);
        filt[i] = (uint32_t)(rng() % 3);
    }
    // the exact piece borders: a request of exactly one slot, one byte more,
    // and a length that leaves a last piece just past the overlap
    texts[1] = make_text(rng, 4096);
    texts[2] = make_text(rng, 4097);
    texts[3] = make_text(rng, 4096 + (4096 - overlap_of(ex.filters[filt[3]])) + 1);
    std::vector<uint64_t> ids(N);
    {
        Scheduler s(ex.cfg, &ex);
        s.start();
        std::thread t2([&] {
            for (int i = 1; i < N; i += 2) {
                std::string err;
                CHECK(s.submit(filt[i], texts[i].data(), texts[i].size(), overlap_of(ex.filters[filt[i]]), &ids[i],
                               err) == FFBF_OK);
            }
        });
        for (int i = 0; i < N; i += 2) {
            std::string err;
            CHECK(s.submit(filt[i], texts[i].data(), texts[i].size(), overlap_of(ex.filters[filt[i]]), &ids[i], err) ==
                  FFBF_OK);
        }
        t2.join();
        for (int i = 0; i < N; i++) {
            std::vector<ffbf_match> got;
            std::string err;
            double lat = -1;
            CHECK(s.wait(ids[i], got, &lat, err) == FFBF_OK);
            CHECK(lat >= 0);
            CHECK(same(got, naive(ex.filters[filt[i]], texts[i].data(), texts[i].size())));
        }
        ffbf_server_stats st;
        s.stats(&st);
        CHECK(st.submitted == (uint64_t)N && st.completed == (uint64_t)N && st.pending == 0 && st.in_flight_bytes == 0);
        std::string err;
        uint64_t id;
        // a request whose piece exceeds the budget is refused up front
        Scheduler tiny(SchedConfig{piece_footprint(4096) - 1, 4096, 1, 4}, &ex);
        CHECK(tiny.submit(0, texts[1].data(), 4096, 10, &id, err) == FFBF_ECAPACITY);
        ffbf_server_stats ts;
        tiny.stats(&ts);
        CHECK(ts.rejected == 1);
        // longer than a slot needs an overlap of 1 .. slot/2
        CHECK(s.submit(0, texts[3].data(), texts[3].size(), 0, &id, err) == FFBF_EINVAL);
        CHECK(s.submit(0, texts[3].data(), texts[3].size(), 2049, &id, err) == FFBF_EINVAL);
        std::vector<ffbf_match> none;
        CHECK(s.wait(987654, none, nullptr, err) == FFBF_EINVAL);
    }
    CHECK(ex.max_in_flight <= ex.cfg.hbm_budget_bytes);
    CHECK(ex.in_flight == 0);
}

// A failed batch fails exactly its requests; the loop goes on.
static void test_executor_failure()
{
    std::mt19937_64 rng(2);
    FakeExecutor ex;
    ex.cfg = SchedConfig{1 << 20, 4096, 1, 2};
    ex.filters.push_back(make_patterns(rng, 8, 3, 6));
    ex.fail_batch = 1;
    std::vector<std::vector<uint8_t>> texts;
    for (int i = 0; i < 6; i++) texts.push_back(make_text(rng, 1500));
    Scheduler s(ex.cfg, &ex);
    std::vector<uint64_t> ids(6);
    std::string err;
    // all six queued before the slot starts: batches {0,1}, {2,3}, {4,5}
    for (int i = 0; i < 6; i++) CHECK(s.submit(0, texts[i].data(), texts[i].size(), 5, &ids[i], err) == FFBF_OK);
    s.start();
    // one slot, two pieces per batch: batch 1 holds requests 2 and 3
    for (int i = 0; i < 6; i++) {
        std::vector<ffbf_match> got;
        const int st = s.wait(ids[i], got, nullptr, err);
        if (i == 2 || i == 3) {
            CHECK(st == FFBF_ENOMEM);
            CHECK(err.find("injected") != std::string::npos);
        } else {
            CHECK(st == FFBF_OK);
            CHECK(same(got, naive(ex.filters[0], texts[i].data(), texts[i].size())));
        }
    }
}

// Cancel before, during and after a request runs.
static void test_cancel()
{
    std::mt19937_64 rng(3);
    FakeExecutor ex;
    ex.cfg = SchedConfig{1 << 20, 4096, 1, 1};
    ex.filters.push_back(make_patterns(rng, 8, 3, 6));
    std::vector<uint8_t> a = make_text(rng, 3000), b = make_text(rng, 3000), c = make_text(rng, 3000);
    std::vector<uint8_t> big = make_text(rng, 3 * 4096);
    Scheduler s(ex.cfg, &ex);
    ex.close_gate();
    s.start();
    std::string err;
    uint64_t ia, ib, ic, ibig;
    CHECK(s.submit(0, a.data(), a.size(), 5, &ia, err) == FFBF_OK);
    while (ex.waiting.load() == 0) std::this_thread::yield();  // a is on the "device"
    CHECK(s.submit(0, b.data(), b.size(), 5, &ib, err) == FFBF_OK);
    CHECK(s.submit(0, big.data(), big.size(), 5, &ibig, err) == FFBF_OK);
    CHECK(s.submit(0, c.data(), c.size(), 5, &ic, err) == FFBF_OK);
    int done = 0;
    // b never started: cancelled at once
    CHECK(s.cancel(ib, err) == FFBF_OK);
    CHECK(s.poll(ib, &done, err) == FFBF_OK && done == 1);
    std::vector<ffbf_match> got;
    CHECK(s.wait(ib, got, nullptr, err) == FFBF_ECANCELED);
    // a is running: it completes cancelled once its batch is back
    CHECK(s.cancel(ia, err) == FFBF_OK);
    CHECK(s.poll(ia, &done, err) == FFBF_OK && done == 0);
    CHECK(s.cancel(987654, err) == FFBF_EINVAL);
    ex.open_gate();
    CHECK(s.wait(ia, got, nullptr, err) == FFBF_ECANCELED);
    // the multi-piece request and c run normally
    CHECK(s.wait(ibig, got, nullptr, err) == FFBF_OK);
    CHECK(same(got, naive(ex.filters[0], big.data(), big.size())));
    while (s.poll(ic, &done, err) == FFBF_OK && !done) std::this_thread::yield();
    CHECK(s.cancel(ic, err) == FFBF_ESTATE);  // too late: the result stands
    CHECK(s.wait(ic, got, nullptr, err) == FFBF_OK);
    CHECK(same(got, naive(ex.filters[0], c.data(), c.size())));
    // a multi-piece request cancelled while its first piece runs
    ex.close_gate();
    uint64_t ib2;
    CHECK(s.submit(0, big.data(), big.size(), 5, &ib2, err) == FFBF_OK);
    while (ex.waiting.load() == 0) std::this_thread::yield();
    CHECK(s.cancel(ib2, err) == FFBF_OK);
    ex.open_gate();
    CHECK(s.wait(ib2, got, nullptr, err) == FFBF_ECANCELED);
    ffbf_server_stats st;
    s.stats(&st);
    CHECK(st.pending == 0 && st.in_flight_bytes == 0);
}

// Destroying the scheduler finishes what was submitted.
static void test_shutdown_drains()
{
    std::mt19937_64 rng(4);
    FakeExecutor ex;
    ex.cfg = SchedConfig{1 << 20, 2048, 2, 3};
    ex.filters.push_back(make_patterns(rng, 5, 2, 4));
    std::vector<std::vector<uint8_t>> texts;
    for (int i = 0; i < 20; i++) texts.push_back(make_text(rng, 1000 + 300 * (uint64_t)i));
    {
        Scheduler s(ex.cfg, &ex);
        s.start();
        std::string err;
        uint64_t id;
        for (auto &t : texts) CHECK(s.submit(0, t.data(), t.size(), 3, &id, err) == FFBF_OK);
    }
    CHECK(ex.in_flight == 0);
    uint64_t pieces = 0;
    for (auto &t : texts) pieces += t.size() <= 2048 ? 1 : 1 + (t.size() - 2048 + 2044) / 2045;
    CHECK(ex.batches.load() >= pieces / 3);
}

int main()
{
    test_pieces_and_budget();
    test_executor_failure();
    test_cancel();
    test_shutdown_drains();
    if (g_failures) {
        std::fprintf(stderr, "%d checks failed\n", g_failures.load());
        return 1;
    }
    std::printf("sched_test: all checks passed\n");
    return 0;
}
