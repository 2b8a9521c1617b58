"""The serving workload as benchmarked, every request against the oracle.

8 resident filters of 10k..2M patterns (lengths 24-80, sets alternating DNA
and printable ASCII), W=24, k=3, filter bits 16 n rounded up to a power of
two; 104 seeded requests with log-normal sizes (median 512 KiB, clipped to
1 KiB-256 MiB) plus the range's edges: 1 KiB, 128 MiB + 1 (two pieces) and
256 MiB (three pieces).  The server has the benchmark's configuration (1 GiB
HBM budget, 4 slots of 128 MiB).  Every request's match list, submitted all
at once and again at seeded Poisson arrivals, must equal the oracle's flow
(build, scan with feed-forward, pattern pass, verification) on that text
alone.  Needs an MI355X."""
import os
import sys
import threading
import time

import numpy as np
import pytest

from conftest import ROOT

import ffbf_amd as fa
from oracle import Oracle

sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the benchmark's workload definition)

pytestmark = pytest.mark.gpu

FORCED = {0: 1024, 1: (128 << 20) + 1, 2: 256 << 20}
N_REQ = 104


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def workload():
    t0 = time.perf_counter()
    d = bench.serve_inputs(fa, N_REQ, seed=51, forced=FORCED)
    sizes = [t.size for _, t in d["requests"]]
    say(f"serving workload: {N_REQ} requests, {sum(sizes) / 2**20:.0f} MiB, sizes {min(sizes)}..{max(sizes)}, "
        f"generated in {time.perf_counter() - t0:.1f} s")
    assert sizes[0] == 1024 and sizes[1] == (128 << 20) + 1 and sizes[2] == 256 << 20
    # the oracle's flow on every request's text, each filter built once
    t0 = time.perf_counter()
    threads = min(16, os.cpu_count() or 1)
    expect = []
    orcs = [(Oracle(d["W"], d["k"], log2, d["sbox"]), pb, po) for pb, po, log2, _ in d["sets"]]
    Fs = [o.build(pb, po) for o, pb, po in orcs]
    for f, t in d["requests"]:
        o, pb, po = orcs[f]
        R, cand = o.scan(t, Fs[f], threads=threads)
        surv = o.pattern_pass(pb, po, R)
        expect.append(o.verify(t, pb, po, cand, surv))
    say(f"oracle over every request in {time.perf_counter() - t0:.1f} s, "
        f"{sum(e[0].size for e in expect)} matches")
    ms = []
    for pb, po, log2, _ in d["sets"]:
        m = fa.Matcher(d["W"], d["k"], log2, d["sbox"])
        m.set_patterns(pb, po)
        m.build()
        ms.append(m)
    yield d, expect, ms
    for m in ms:
        m.close()
    d["requests"] = None
    d["pinned"].free()


def check(got, exp, i):
    ids, offs = exp
    assert np.array_equal(got["pattern_id"], ids), f"request {i}"
    assert np.array_equal(got["offset"], offs), f"request {i}"


def test_serving_workload_all_at_once(workload):
    d, expect, ms = workload
    with fa.Server(ms, hbm_budget_bytes=1 << 30, slots=4, max_batch_bytes=128 << 20) as srv:
        ids = srv.submit_many(d["requests"])
        for i, rid in enumerate(ids):
            got, lat = srv.wait(rid)
            check(got, expect[i], i)
        st = srv.stats()
        say(f"all at once: {st['batches']} batches for {st['completed']} requests")
        assert st["completed"] == N_REQ and st["rejected"] == 0 and st["in_flight_bytes"] == 0
    # the three forced sizes hold matches, and the big ones span piece borders
    assert all(expect[i][0].size > 0 for i in (1, 2))


def test_serving_workload_poisson(workload):
    d, expect, ms = workload
    total = sum(t.size for _, t in d["requests"])
    # arrivals at ~20 GB/s of offered text: the queue stays short but slots overlap
    lam = N_REQ / (total / 20e9)
    arrivals = np.cumsum(d["gaps"]) / lam
    with fa.Server(ms, hbm_budget_bytes=1 << 30, slots=4, max_batch_bytes=128 << 20) as srv:
        rids = [None] * N_REQ
        ready = [threading.Event() for _ in range(N_REQ)]

        def submitter():
            start = time.perf_counter()
            for i, (f, t) in enumerate(d["requests"]):
                dt = arrivals[i] - (time.perf_counter() - start)
                if dt > 0:
                    time.sleep(dt)
                rids[i] = srv.submit(f, t)
                ready[i].set()

        th = threading.Thread(target=submitter)
        th.start()
        lat = []
        for i in range(N_REQ):
            ready[i].wait()
            got, ms_ = srv.wait(rids[i])
            lat.append(ms_)
            check(got, expect[i], i)
        th.join()
        say(f"poisson at {lam:.0f} req/s: p50 {np.percentile(lat, 50):.2f} ms, p99 {np.percentile(lat, 99):.2f} ms")


def test_serving_cancel_in_the_mix(workload):
    d, expect, ms = workload
    with fa.Server(ms, hbm_budget_bytes=1 << 30, slots=4, max_batch_bytes=128 << 20) as srv:
        ids = srv.submit_many(d["requests"][:12])
        # the 256 MiB request (three pieces) cancelled while the rest run
        cancelled = srv.cancel(ids[2])
        for i, rid in enumerate(ids):
            if i == 2:
                if cancelled:
                    with pytest.raises(fa.FFBFError) as e:
                        srv.wait(rid)
                    assert e.value.status == fa.ECANCELED
                else:  # it had completed already: its result stands
                    got, _ = srv.wait(rid)
                    check(got, expect[i], i)
                continue
            got, _ = srv.wait(rid)
            check(got, expect[i], i)
        assert srv.stats()["in_flight_bytes"] == 0
        # the server goes on normally after the cancellation
        got, _ = srv.wait(srv.submit(*d["requests"][5]))
        check(got, expect[5], 5)
