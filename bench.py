"""Benchmark of the FFBF scan path on one MI355X.

A step is one pass of the whole path over one synthetic batch whose inputs are
already resident in HBM: build F from the pattern set, clear R and scan the
text with feed-forward, run the pattern pass, verify the survivors exactly,
and bring the sorted (pattern id, offset) match list back to the host.

Default workload: 1,000,000 DNA patterns (lengths 32-100, GC 40%), a 2 GiB
GC-40 DNA text with 50,000 planted occurrences, window 32, k=4, 2^28-bit
filter.  `value` is text GB/s over the whole step.  The roofline object
prices the scan kernel against HBM: algorithmic bytes = 1 byte read per text
byte, over its average launch time from HIP events on the library's stream.
The CPU baseline times the oracle (a C restatement of the reference
algorithm) on this machine's host cores over a bounded sample of the same
inputs.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
                      [--workload plain|ascii10k-64m|ascii100k-1536m|batch-4096texts-1set|
                                  batch-2048texts-64sets|serve-8sets|ingest|
                                  split-10k-64m-lines|split-1m-2g-lines]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
METRIC = "text GB/s (+% HBM peak) at 1M patterns, plain vs batched; p99 req latency"

WORKLOADS = {
    # name: alphabet, W, k, log2 bits, patterns, lmin, lmax, text bytes, plants, seed
    "plain": ("dna", 32, 4, 28, 1_000_000, 32, 100, 2 << 30, 50_000, 2),
    "ascii10k-64m": ("ascii", 19, 4, 24, 10_000, 20, 40, 64 << 20, 1_000, 1),
    "ascii100k-1536m": ("ascii", 19, 4, 26, 100_000, 20, 64, 1536 << 20, 1_536, 3),
}


# The CPU binary's split filter (ffbf_create_split) at its compile-time shape:
# W=19, k=5, s=2, cache part 2^24 bits, external part 2^28 bits; ASCII
# patterns and an ASCII text cut into lines (mean length `line` bytes).
# name: W, k, s, log2 cache, log2 ext, patterns, lmin, lmax, text bytes, line, plants, seed
SPLIT_WORKLOADS = {
    "split-10k-64m-lines": (19, 5, 2, 24, 28, 10_000, 20, 40, 64 << 20, 80, 1_000, 4),
    "split-1m-2g-lines": (19, 5, 2, 24, 28, 1_000_000, 20, 64, 2 << 30, 80, 50_000, 5),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def describe(name, wl):
    alpha, W, k, log2, n_pat, lmin, lmax, n_text, plants, seed = wl
    return {
        "workload": f"{name}: {n_pat} {alpha} patterns ({lmin}-{lmax} B), {n_text / 2**30:.3g} GiB {alpha} text, "
                    f"{plants} planted, W={W} k={k} M=2^{log2}",
        "patterns": n_pat,
        "text_bytes": n_text,
        "window": W,
        "k": k,
        "filter_bits": 1 << log2,
        "alphabet": "DNA GC40" if alpha == "dna" else "printable ASCII",
        "seed": seed,
    }


def make_inputs(fa, wl):
    alpha, W, k, log2, n_pat, lmin, lmax, n_text, plants, seed = wl
    a = fa.ALPHA_DNA_GC40 if alpha == "dna" else fa.ALPHA_ASCII
    sbox = fa.sbox_from_seed(seed)
    pb, po = fa.gen_patterns(a, seed * 10 + 1, n_pat, lmin, lmax)
    text = fa.gen_bytes(a, seed * 10 + 2, n_text)
    fa.plant(seed * 10 + 3, plants, text, pb, po)
    return sbox, pb, po, text


def next_pow2(x: int) -> int:
    p = 1
    while p < x:
        p <<= 1
    return p


def batch_inputs(fa, name, scale=1.0):
    """Synthetic batch of texts routed to resident pattern sets.

    batch-4096texts-1set: 4,096 texts, log-normal sizes (median 256 KiB, sigma 0.9, clipped
    to 4 KiB-64 MiB, ~1.5 GiB), printable ASCII, one 100,000-pattern set
    (lengths 20-64), W=19, k=4, 2^26-bit filter; one planted occurrence per
    started MiB of text.
    batch-2048texts-64sets: 64 pattern sets with n log-uniform in [1e3, 1e6] (lengths 24-80,
    DNA), 2,048 DNA texts (log-normal, median 1 MiB, sigma 1.0, clipped to
    4 KiB-64 MiB) with one tandem-repeat run (period 1-6, 205-614 bases, ~10%
    of bases) per 4 KiB, each text routed to a seeded set; W=24, k=3,
    filter bits = 16 n rounded up to a power of two; one plant per started MiB.
    """
    import numpy as np

    if name == "batch-4096texts-1set":
        W, k, seed = 19, 4, 31
        n_texts = max(8, int(4096 * scale))
        sizes = fa.lognormal_sizes(seed, n_texts, 256 * 1024.0, 0.9, 4096, 64 << 20)
        sets = [(100_000, 20, 64, 26, fa.ALPHA_ASCII)]
        route = np.zeros(n_texts, dtype=np.uint32)
        alpha_of = lambda f: fa.ALPHA_ASCII
        repeats = False
    elif name == "batch-2048texts-64sets":
        W, k, seed = 24, 3, 41
        n_texts = max(8, int(2048 * scale))
        sizes = fa.lognormal_sizes(seed, n_texts, 1024 * 1024.0, 1.0, 4096, 64 << 20)
        sets = []
        for f in range(64):
            u = (fa.gen_u64(seed, 21, f) >> 11) / float(1 << 53)
            n = int(round(1000.0 * 1000.0 ** u))
            sets.append((n, 24, 80, max(10, next_pow2(16 * n).bit_length() - 1), fa.ALPHA_DNA_GC40))
        route = np.array([fa.gen_u64(seed, 22, i) % 64 for i in range(n_texts)], dtype=np.uint32)
        alpha_of = lambda f: fa.ALPHA_DNA_GC40
        repeats = True
    else:
        raise ValueError(name)
    sbox = fa.sbox_from_seed(seed)
    psets = []
    for f, (n, lmin, lmax, log2, alpha) in enumerate(sets):
        pb, po = fa.gen_patterns(alpha, seed * 100 + f, n, lmin, lmax)
        psets.append((pb, po, log2))
    off = np.zeros(n_texts + 1, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    buf = np.empty(int(off[-1]), dtype=np.uint8)
    for i in range(n_texts):
        a, b = int(off[i]), int(off[i + 1])
        t = buf[a:b]
        t[:] = fa.gen_bytes(alpha_of(route[i]), seed * 1000 + 1, b - a, start=a)
        if repeats:
            fa.tandem_repeats(seed * 1000 + 2 + i, t, block=4096)
        pb, po, _ = psets[route[i]]
        fa.plant(seed * 1000 + 3 + i, (b - a + (1 << 20) - 1) >> 20, t, pb, po)
    return {"W": W, "k": k, "sbox": sbox, "sets": psets, "texts": buf, "offsets": off, "route": route,
            "n_texts": n_texts, "bytes": int(off[-1])}


def run_batched(fa, name, steps=10, per_text_reps=2, scale=1.0, cpu=False):
    """Batched (one persistent launch per batch, hipGraph replay, autotuned)
    vs the per-text launch path on the same texts; GB/s of text for each;
    with cpu, the oracle's batched flow on the host beside them."""
    import numpy as np

    t0 = time.perf_counter()
    d = batch_inputs(fa, name, scale)
    log(f"{name}: {d['n_texts']} texts, {d['bytes'] / 2**30:.2f} GiB, {len(d['sets'])} pattern sets, "
        f"generated in {time.perf_counter() - t0:.1f} s")
    dev = fa.DeviceBuffer.from_host(d["texts"])
    ms = []
    for pb, po, log2 in d["sets"]:
        m = fa.Matcher(d["W"], d["k"], log2, d["sbox"])
        m.set_patterns(pb, po)
        m.build()
        ms.append(m)
    bt = fa.Batch(ms)
    bt.set_texts(dev.ptr, d["offsets"], d["route"])
    t0 = time.perf_counter()
    best, best_ms = bt.autotune()
    tune_s = time.perf_counter() - t0
    bt.run()
    ref = bt.matches()
    # timed without the run's two timing events (~10 us), as the plain leg
    bt.set_timing(False)
    t0 = time.perf_counter()
    for _ in range(steps):
        bt.run()
        got = bt.matches()
    batched_s = (time.perf_counter() - t0) / steps
    bt.set_timing(True)
    bt.run()
    run_ms, _ = bt.timings()
    # scan kernel alone: one stage-by-stage run
    bt.set_tuning(best["chunk_iters"], best["probes"], best["block"], use_graph=False,
                  bytes_per_lane=best["bytes_per_lane"])
    bt.run()
    _, scan_ms = bt.timings()
    bt.set_tuning(best["chunk_iters"], best["probes"], best["block"], use_graph=True,
                  bytes_per_lane=best["bytes_per_lane"])
    bt.run_per_text()
    t0 = time.perf_counter()
    for _ in range(per_text_reps):
        bt.run_per_text()
        per = bt.matches()
    per_s = (time.perf_counter() - t0) / per_text_reps
    same = bool(np.array_equal(per, ref) and np.array_equal(got, ref))
    bt.close()
    for m in ms:
        m.close()
    dev.free()
    gbs = d["bytes"] / batched_s / 1e9
    per_gbs = d["bytes"] / per_s / 1e9
    return {
        "workload": name,
        "texts": d["n_texts"],
        "text_bytes": d["bytes"],
        "pattern_sets": len(d["sets"]),
        "value": round(gbs, 3),
        "unit": "GB/s",
        "ms_per_batch": round(batched_s * 1e3, 3),
        "graph_run_ms": round(run_ms, 3),
        "scan_kernel_ms": round(scan_ms, 3),
        "scan_kernel_gbs": round(d["bytes"] / (scan_ms * 1e-3) / 1e9, 3) if scan_ms > 0 else None,
        "per_text_gbs": round(per_gbs, 3),
        "per_text_ms": round(per_s * 1e3, 3),
        "speedup_vs_per_text": round(gbs / per_gbs, 2),
        "matches": int(ref.size),
        "same_matches_as_per_text": same,
        "autotuned": best,
        "autotune_s": round(tune_s, 2),
        "cpu_baseline": cpu_batched_baseline(d) if cpu else None,
    }


def serve_inputs(fa, n_requests, seed=51, forced=None, gap=0):
    """Serving workload: 8 resident pattern sets, n log-spaced 10k..2M
    (lengths 24-80; sets alternate DNA and printable ASCII), W=24, k=3,
    filter bits = 16 n rounded up to a power of two (the multi-filter rule);
    requests with log-normal sizes (median 512 KiB, sigma 1.5, clipped to
    1 KiB-256 MiB), each routed to a seeded set and drawn over that set's
    alphabet, one planted occurrence per started MiB.  forced: {request
    index: size} overriding drawn sizes (the range's edges)."""
    import numpy as np

    W, k = 24, 3
    sbox = fa.sbox_from_seed(seed)
    sets = []
    for f in range(8):
        n = int(round(10_000 * 200.0 ** (f / 7.0)))
        alpha = fa.ALPHA_DNA_GC40 if f % 2 == 0 else fa.ALPHA_ASCII
        pb, po = fa.gen_patterns(alpha, seed * 100 + f, n, 24, 80)
        sets.append((pb, po, next_pow2(16 * n).bit_length() - 1, alpha))
    sizes = fa.lognormal_sizes(seed, n_requests, 512 * 1024.0, 1.5, 1024, 256 << 20)
    for i, sz in (forced or {}).items():
        sizes[i] = sz
    route = [int(fa.gen_u64(seed, 23, i) % 8) for i in range(n_requests)]
    # request texts live in page-locked host memory, as a serving frontend's
    # receive buffers would
    # (gap > 0: that many bytes between consecutive requests, as requests in
    # separate buffers are)
    pinned = fa.PinnedBuffer(int(sizes.sum()) + gap * n_requests)
    reqs, pos = [], 0
    for i in range(n_requests):
        pb, po, _, alpha = sets[route[i]]
        t = pinned.array[pos:pos + int(sizes[i])]
        pos += int(sizes[i]) + gap
        t[:] = fa.gen_bytes(alpha, seed * 1000 + 1 + route[i], int(sizes[i]), start=int(fa.gen_u64(seed, 24, i) >> 24))
        fa.plant(seed * 1000 + 9 + i, (t.size + (1 << 20) - 1) >> 20, t, pb, po)
        reqs.append((route[i], t))
    # Poisson arrivals: exponential gaps from 53-bit uniforms (log of a
    # uniform through numpy; the gaps only pace submission)
    u = np.array([(fa.gen_u64(seed, 25, i) >> 11) / float(1 << 53) for i in range(n_requests)])
    gaps = -np.log1p(-u)  # mean 1, scaled by 1/lambda at run time
    return {"W": W, "k": k, "sbox": sbox, "sets": sets, "requests": reqs, "gaps": gaps, "pinned": pinned}


def percentile(xs, q):
    import numpy as np

    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None


UPLOADS = {"default": 0, "copy": 1, "kernel": 2}


def run_serving(fa, n_requests=600, load=0.7, budget=1 << 30, slots=4, max_batch_bytes=128 << 20, reps=5, upload=0,
                gap=0):
    """Continuous-batching server vs the plain per-request path (one
    ffbf_match_text per request, first come first served) on the same
    requests.  Saturation throughput of each (everything queued at once);
    then latency under seeded Poisson arrivals, each path driven for real by a
    submitting thread: at loads stated against the plain path's own
    saturation rate (where both are stable), and at loads against the
    server's (where the plain path is past its saturation, so it is reported
    as unstable and not driven).  Each load point is `reps` passes over the
    same arrivals."""
    import queue
    import threading

    import numpy as np

    t0 = time.perf_counter()
    d = serve_inputs(fa, n_requests, gap=gap)
    total = sum(t.size for _, t in d["requests"])
    log(f"serve: {n_requests} requests, {total / 2**30:.2f} GiB, 8 filters, generated in {time.perf_counter() - t0:.1f} s")
    ms = []
    for pb, po, log2, _ in d["sets"]:
        m = fa.Matcher(d["W"], d["k"], log2, d["sbox"])
        m.set_patterns(pb, po)
        m.build()
        ms.append(m)
    reqs = d["requests"]
    # plain per-request path: every request alone, back to back
    for f, t in reqs[:8]:
        ms[f].match(t)  # warm
    plain_res = []
    t1 = time.perf_counter()
    for f, t in reqs:
        plain_res.append(ms[f].match(t).matches)
    plain_s = time.perf_counter() - t1
    plain_rate = n_requests / plain_s
    t1 = time.perf_counter()
    srv = fa.Server(ms, hbm_budget_bytes=budget, slots=slots, max_batch_bytes=max_batch_bytes, upload=upload)
    create_s = time.perf_counter() - t1
    # the first pass over the new server, then one more before measuring
    t1 = time.perf_counter()
    for rid in srv.submit_many(reqs):
        srv.wait(rid)
    cold_s = time.perf_counter() - t1
    for rid in srv.submit_many(reqs):
        srv.wait(rid)
    # saturation: everything submitted at once; the timed region ends when
    # the last request's matches are in hand (compared with the plain path's
    # afterwards)
    sat_passes = []
    for r in range(reps):
        t1 = time.perf_counter()
        ids = srv.submit_many(reqs)
        got = [srv.wait(rid)[0] for rid in ids]
        sat_passes.append(time.perf_counter() - t1)
        if r == 0:
            got_all = got
    sat_s = float(np.median(sat_passes))
    srv_rate = n_requests / sat_s
    same = all(np.array_equal(g["pattern_id"], p["pattern_id"]) and np.array_equal(g["offset"], p["offset"])
               for g, p in zip(got_all, plain_res))
    got_all = got = None

    def arrivals_at(lam):
        return np.cumsum(d["gaps"]) / lam

    def drive_server(lam):
        arrivals = arrivals_at(lam)
        lat = [None] * n_requests
        rids = [None] * n_requests
        lag = [0.0]
        ready = [threading.Event() for _ in range(n_requests)]

        def submitter():
            start = time.perf_counter()
            for i, (f, t) in enumerate(reqs):
                dt = arrivals[i] - (time.perf_counter() - start)
                if dt > 0:
                    time.sleep(dt)
                elif -dt > lag[0]:
                    lag[0] = -dt  # submitted late: the arrivals behind it come as a burst
                rids[i] = srv.submit(f, t)
                ready[i].set()

        th = threading.Thread(target=submitter)
        t1 = time.perf_counter()
        th.start()
        for i in range(n_requests):
            ready[i].wait()
            _, lat[i] = srv.wait(rids[i])
        th.join()
        drive_server.lag_ms = lag[0] * 1e3
        return lat, time.perf_counter() - t1

    def drive_plain(lam):
        """The plain path at the same arrivals: a submitting thread queues
        each request at its arrival time; one worker serves them in order,
        one ffbf_match_text each; latency is arrival to matches in hand."""
        arrivals = arrivals_at(lam)
        q = queue.Queue()
        lat = [None] * n_requests

        def submitter():
            start = time.perf_counter()
            for i in range(n_requests):
                dt = arrivals[i] - (time.perf_counter() - start)
                if dt > 0:
                    time.sleep(dt)
                q.put((i, time.perf_counter()))

        th = threading.Thread(target=submitter)
        t1 = time.perf_counter()
        th.start()
        for _ in range(n_requests):
            i, t_arr = q.get()
            f, t = reqs[i]
            ms[f].match(t)
            lat[i] = (time.perf_counter() - t_arr) * 1e3
        th.join()
        return lat, time.perf_counter() - t1

    def pct(xs):
        return round(percentile(xs, 50), 3), round(percentile(xs, 99), 3)

    def passes(drive, lam):
        """`reps` passes over the same seeded arrivals: all their latencies
        (p50/p99 over reps x n_requests samples, so that one host stall does
        not make the p99), the mean pass time and each pass's own p99."""
        lat, el, p99s = [], 0.0, []
        passes.lags = []
        for _ in range(reps):
            l1, e1 = drive(lam)
            lat += l1
            el += e1
            p99s.append(round(percentile(l1, 99), 3))
            passes.lags.append(round(getattr(drive, "lag_ms", 0.0), 3))
        return lat, el / reps, p99s

    # both paths at the same arrival rates, below the plain path's saturation
    same_rate = []
    for ld in (0.5, 0.8):
        lam = ld * plain_rate
        s_lat, s_el, s_p99s = passes(drive_server, lam)
        s_lags = passes.lags
        p_lat, p_el, p_p99s = passes(drive_plain, lam)
        row = {"load_vs_plain_saturation": ld, "load_vs_server_saturation": round(lam / srv_rate, 4),
               "rate_per_s": round(lam, 1), "server_p50_ms": pct(s_lat)[0], "server_p99_ms": pct(s_lat)[1],
               "plain_p50_ms": pct(p_lat)[0], "plain_p99_ms": pct(p_lat)[1],
               "server_p99_ms_per_pass": s_p99s, "plain_p99_ms_per_pass": p_p99s, "server_submit_lag_ms_per_pass": s_lags,
               "server_gbs": round(total / s_el / 1e9, 3), "plain_gbs": round(total / p_el / 1e9, 3)}
        same_rate.append(row)
        log(f"serve: same arrivals at {lam:.0f} req/s ({ld} of plain saturation): server p99 {row['server_p99_ms']} ms, "
            f"plain p99 {row['plain_p99_ms']} ms")
    # the server at loads of its own saturation rate
    sweep = []
    lat_at_load = None
    for ld in (0.3, 0.5, 0.7, 0.9):
        lam = ld * srv_rate
        s_lat, s_el, s_p99s = passes(drive_server, lam)
        if ld == load:
            lat_at_load, poisson_s = s_lat, s_el
        row = {"load": ld, "rate_per_s": round(lam, 1), "p50_ms": pct(s_lat)[0], "p99_ms": pct(s_lat)[1],
               "p99_ms_per_pass": s_p99s, "submit_lag_ms_per_pass": passes.lags,
               "gbs": round(total / s_el / 1e9, 3), "load_vs_plain_saturation": round(lam / plain_rate, 3),
               "plain": "unstable: arrivals above its saturation rate" if lam >= plain_rate else "stable"}
        sweep.append(row)
        log(f"serve: load {ld}: {lam:.0f} req/s, p50 {row['p50_ms']} ms, p99 {row['p99_ms']} ms "
            f"({row['load_vs_plain_saturation']}x the plain path's saturation rate)")
    stats = srv.stats()
    srv.close()
    # The same requests through the server's upload kernel (text read from
    # pinned memory by a kernel instead of the DMA engine: no DMA-path
    # stalls, lower saturation): saturation, then two loads of its own rate
    kernel_leg = None
    if upload != UPLOADS["kernel"]:
        srv = fa.Server(ms, hbm_budget_bytes=budget, slots=slots, max_batch_bytes=max_batch_bytes,
                        upload=UPLOADS["kernel"])
        for _ in range(2):
            for rid in srv.submit_many(reqs):
                srv.wait(rid)
        k_sat = []
        for _ in range(reps):
            t1 = time.perf_counter()
            for rid in srv.submit_many(reqs):
                srv.wait(rid)
            k_sat.append(time.perf_counter() - t1)
        k_rate = n_requests / float(np.median(k_sat))
        k_loads = []
        for ld in (0.5, 0.7):
            lam = ld * k_rate
            s_lat, _, s_p99s = passes(drive_server, lam)
            k_loads.append({"load": ld, "rate_per_s": round(lam, 1), "p50_ms": pct(s_lat)[0], "p99_ms": pct(s_lat)[1],
                            "p99_ms_per_pass": s_p99s, "submit_lag_ms_per_pass": passes.lags})
            log(f"serve (upload kernel): load {ld}: {lam:.0f} req/s, p50 {pct(s_lat)[0]} ms, p99 {pct(s_lat)[1]} ms")
        kernel_leg = {"saturation_gbs": round(total / float(np.median(k_sat)) / 1e9, 3),
                      "saturation_gbs_per_pass": [round(total / x / 1e9, 3) for x in k_sat],
                      "saturation_rate_per_s": round(k_rate, 1), "load_sweep": k_loads,
                      "kernel_upload_bytes": int(srv.stats()["kernel_upload_bytes"])}
        srv.close()
    for m in ms:
        m.close()
    d["requests"] = None
    d["pinned"].free()
    return {
        "workload": "serve-8sets",
        "requests": n_requests,
        "text_bytes": int(total),
        "filters": 8,
        "saturation_gbs": round(total / sat_s / 1e9, 3),
        "saturation_gbs_per_pass": [round(total / x / 1e9, 3) for x in sat_passes],
        "saturation_rate_per_s": round(srv_rate, 1),
        "cold_start_gbs": round(total / cold_s / 1e9, 3),
        "server_create_s": round(create_s, 3),
        "plain_gbs": round(total / plain_s / 1e9, 3),
        "plain_saturation_rate_per_s": round(plain_rate, 1),
        "speedup_vs_plain": round(plain_s / sat_s, 2),
        "same_matches_as_plain": bool(same),
        "poisson_load": load,
        "poisson_rate_per_s": round(load * srv_rate, 1),
        "p50_ms": round(percentile(lat_at_load, 50), 3),
        "p99_ms": round(percentile(lat_at_load, 99), 3),
        "poisson_gbs": round(total / poisson_s / 1e9, 3),
        "same_arrivals": same_rate,
        "load_sweep": sweep,
        "batches": int(stats["batches"]),
        "upload": {0: "default", 1: "copy", 2: "kernel"}.get(upload, str(upload)),
        "request_layout": f"{gap} bytes between requests" if gap else "back to back in one pinned buffer",
        "kernel_upload_bytes": int(stats["kernel_upload_bytes"]),
        "upload_kernel_leg": kernel_leg,
        "hbm_budget_bytes": budget,
        "slots": slots,
        "max_batch_bytes": max_batch_bytes,
        "latency_drivers": "both paths driven by a submitting thread at seeded Poisson arrivals; latency from "
                           f"submission to matches in hand; percentiles over {reps} passes of the same arrivals",
        "passes_per_load": reps,
    }


def run_ingest(fa, reps=3):
    """End-to-end rate from host memory and from a file for the plain
    workload (1M DNA patterns, 2 GiB text): pipelined ingest (pinned staging,
    copies overlapped with per-piece scans) vs copying the whole pageable text
    to the device first and then scanning it.  Matches are identical."""
    import tempfile

    import numpy as np

    wl = WORKLOADS["plain"]
    alpha, W, k, log2, n_pat, lmin, lmax, n_text, plants, seed = wl
    sbox, pb, po, text = make_inputs(fa, wl)
    m = fa.Matcher(W, k, log2, sbox)
    m.set_patterns(pb, po)
    m.build()
    m.synchronize()

    def timed(fn):
        fn()  # warm (staging buffers, device text, candidate lists)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            r = fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), r

    def copy_then_scan():
        dev = fa.DeviceBuffer.from_host(text)
        nm = m.run_device(dev.ptr, n_text)
        got = m.matches()
        dev.free()
        return got

    t_seq, ref = timed(copy_then_scan)
    log(f"ingest: copy-then-scan {n_text / t_seq / 1e9:.2f} GB/s")
    t_host, got_h = timed(lambda: m.match(text).matches)
    log(f"ingest: pipelined host memory {n_text / t_host / 1e9:.2f} GB/s")
    tmpdir = os.environ.get("TMPDIR") or tempfile.gettempdir()
    fd, path = tempfile.mkstemp(dir=tmpdir, suffix=".text")
    os.close(fd)
    try:
        text.tofile(path)

        def from_file():
            m.run_file(path)
            return m.matches()

        t_file, got_f = timed(from_file)
    finally:
        os.unlink(path)
    log(f"ingest: pipelined file {n_text / t_file / 1e9:.2f} GB/s")
    same = bool(np.array_equal(got_h, ref) and np.array_equal(got_f, ref))
    m.close()
    return {
        "workload": "ingest (plain workload from host memory / a page-cached file)",
        "text_bytes": n_text,
        "copy_then_scan_gbs": round(n_text / t_seq / 1e9, 3),
        "host_pipelined_gbs": round(n_text / t_host / 1e9, 3),
        "file_pipelined_gbs": round(n_text / t_file / 1e9, 3),
        "speedup_host_vs_copy_then_scan": round(t_seq / t_host, 2),
        "same_matches": same,
        "staging": "4 x 64 MiB pinned, 8 host threads",
    }


def probe_ceiling(fa, m, k, n_text, W, t_scan_ms):
    """For the scans that probe a prefilter in global memory (an L2-resident
    fold of F, behind an LDS fold or not, or F itself): random 4-byte probes
    per second the scan made, beside the chip's random-gather rate into an
    array of the prefilter's size measured now.  Probes per window follow
    from the prefilter's fill f: bit 0 for every window, bit 1 where bit 0 is
    set, bits 2..k-1 together where both are (F's own bits after a full fold
    hit add ~f^k).  Behind an LDS fold of fill l only some windows probe: when
    its P up-front bits pass more than 5%, the share l^(P+1) passing those and
    one more LDS bit (they probe before queueing), else the share l^k passing
    all its k bits (they probe in the drain)."""
    import numpy as np

    t = m.timings()
    if t["scan_mode"] not in (0, 3, 4):
        return None
    F = m.filter(fa.FILTER_FORWARD)

    def fold(log2_bits):
        words = 1 << (log2_bits - 5)
        return np.bitwise_or.reduce(F.reshape(-1, words), axis=0) if words < F.size else F

    def fill(a):
        return float(np.bitwise_count(a).sum()) / (a.size * 32.0)

    pre_log2 = t["l2_fold_log2_bits"] if t["scan_mode"] in (3, 4) else (F.size * 32).bit_length() - 1
    pre = fold(pre_log2) if t["scan_mode"] in (3, 4) else F
    f = fill(pre)
    lds_pass = 1.0
    if t["scan_mode"] == 4:
        lf = fill(fold(t["fold_log2_bits"]))
        lds_pass = lf ** min(t["probes"] + 1, k) if lf ** t["probes"] > 0.05 else lf ** k
    per_window = lds_pass * (1.0 + (f if k > 1 else 0.0) + (max(0, k - 2) * f * f))
    windows = max(0, n_text - W + 1)
    rate = windows * per_window / (t_scan_ms * 1e-3)
    _, ceiling = fa.gather_bench(pre_log2 - 3, steps=16)
    return {
        "prefilter": {0: "F", 3: "L2-resident fold of F",
                      4: "L2-resident fold of F behind an LDS fold"}[t["scan_mode"]],
        "prefilter_bytes": pre.size * 4,
        "prefilter_fill": round(f, 5),
        "lds_pass": round(lds_pass, 5),
        "probes_per_window": round(per_window, 4),
        "scan_probes_per_s": round(rate, 1),
        "gather_probes_per_s": round(ceiling, 1),
        "frac": round(rate / ceiling, 4) if ceiling > 0 else None,
        # L2 counters of this scan from the committed rocprofv3 run: requests per
        # window and per L2 channel-cycle, beside the gather kernel's
        "counters": traffic_from_profiles("plain_l2"),
    }


def copy_kernel_rate(fa, nbytes=4 << 30, reps=5):
    """HBM read+write GB/s of the library's plain copy kernel (median of
    `reps` synchronous copies of `nbytes`): the practical streaming ceiling
    beside the 8 TB/s peak."""
    import numpy as np

    src, dst = fa.DeviceBuffer(nbytes), fa.DeviceBuffer(nbytes)
    dst.copy_from(src, nbytes)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        dst.copy_from(src, nbytes)
        ts.append(time.perf_counter() - t0)
    src.free()
    dst.free()
    return round(2 * nbytes / float(np.median(ts)) / 1e9, 1)


def cpu_threads():
    """Host threads of the threaded CPU baseline: the CPUs this process may
    run on, at most 16 -- one GPU's share of the host's cores on the MI355X
    boxes, whatever the machine's total count is."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def cpu_baseline(wl, sbox, pb, po, text, budget_s=12.0):
    """Oracle (C restatement of the reference path) on a bounded text sample."""
    from oracle import Oracle

    alpha, W, k, log2, n_pat, lmin, lmax, n_text, plants, seed = wl
    return _cpu_baseline(Oracle(W, k, log2, sbox), n_pat, pb, po, text, budget_s)


def _cpu_baseline(o, n_pat, pb, po, text, budget_s):
    """The oracle's whole flow over one text on 1 thread and on cpu_threads()
    threads, each stage timed on its own: build (all patterns), scan with
    feed-forward (the whole text when the budget allows, else its first
    bytes, the time scaled to the whole), pattern pass, verification (scaled
    like the scan).  The scan is the median of three after one warm run."""
    import numpy as np

    n_text = text.size
    threads = cpu_threads()
    t0 = time.perf_counter()
    F = o.build(pb, po)
    t_build = time.perf_counter() - t0
    # calibrate on 8 MiB with one thread, then size the samples to the budget
    probe = text[: 8 << 20]
    t0 = time.perf_counter()
    o.scan(probe, F)
    rate1 = probe.size / max(time.perf_counter() - t0, 1e-9)
    n1 = int(min(text.size, max(8 << 20, rate1 * budget_s * 0.3 / 4)))
    nT = int(min(text.size, max(8 << 20, rate1 * threads * budget_s * 0.5 / 4)))
    n1 = n1 if n1 == text.size else n1 & ~4095
    nT = nT if nT == text.size else nT & ~4095
    res = {}
    for label, n, th in (("1", n1, 1), ("T", nT, threads)):
        sample = np.ascontiguousarray(text[:n])
        o.scan(sample, F, threads=th)  # warm
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            R, cand = o.scan(sample, F, threads=th)
            runs.append(time.perf_counter() - t0)
        t_scan = float(np.median(runs)) * (n_text / n)
        t0 = time.perf_counter()
        surv = o.pattern_pass(pb, po, R)
        t_pass = time.perf_counter() - t0
        t0 = time.perf_counter()
        o.verify(sample, pb, po, cand, surv)
        t_ver = (time.perf_counter() - t0) * (n_text / n)
        total = t_build + t_scan + t_pass + t_ver
        res[label] = {"threads": th, "sample_bytes": n, "build_s": round(t_build, 4), "scan_s": round(t_scan, 4),
                      "pattern_pass_s": round(t_pass, 4), "verify_s": round(t_ver, 4), "total_s": round(total, 4),
                      "gbs": round(n_text / total / 1e9, 4), "scan_gbs": round(n_text / t_scan / 1e9, 4)}
    T = res["T"]
    return {
        "value": T["gbs"],
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle build of all {n_pat} patterns ({t_build:.2f} s) + scan with feed-forward of "
                   + ("the whole text" if T["sample_bytes"] == n_text else f"the first {T['sample_bytes'] >> 20} MiB")
                   + f" on {threads} threads ({T['scan_gbs']:.3f} GB/s, median of 3 after one warm run, scaled to "
                   f"{n_text >> 20} MiB) + pattern pass + verification; 1 thread: {res['1']['gbs']:.4f} GB/s on "
                   + ("the whole text" if res["1"]["sample_bytes"] == n_text else
                      f"the first {res['1']['sample_bytes'] >> 20} MiB")
                   + f"; host {cpu_model()}, {os.cpu_count()} logical CPUs, {threads} used"),
        "single_thread_value": res["1"]["gbs"],
        "stages": res,
    }


def cpu_batched_baseline(d, budget_s=8.0):
    """The oracle over a batch the way the batched path runs it: every filter
    built once; every text scanned with feed-forward into its filter's R; one
    pattern pass per filter; each text's candidates verified against its
    filter's survivors.  On 1 thread and on cpu_threads() threads (texts
    spread over a pool; the oracle's C calls run outside the interpreter
    lock).  Texts are taken in order until the scan budget is spent; scan and
    verification times are scaled to all the batch's bytes."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    from oracle import Oracle

    threads = cpu_threads()
    off, route, buf = d["offsets"], d["route"], d["texts"]
    n_texts, total = d["n_texts"], d["bytes"]
    orcs = [Oracle(d["W"], d["k"], log2, d["sbox"]) for _, _, log2 in d["sets"]]
    t0 = time.perf_counter()
    Fs = [o.build(pb, po) for o, (pb, po, _) in zip(orcs, d["sets"])]
    t_build = time.perf_counter() - t0
    active = sorted(set(int(f) for f in route))

    def text(i):
        return buf[int(off[i]):int(off[i + 1])]

    def scan_one(i):
        f = int(route[i])
        R, cand = orcs[f].scan(text(i), Fs[f])
        return i, f, R, cand

    res = {}
    for label, th in (("1", 1), ("T", threads)):
        R_union = {f: np.zeros_like(Fs[f]) for f in active}
        cands, sample = [], 0
        t0 = time.perf_counter()
        with ThreadPoolExecutor(th) as ex:
            i = 0
            while i < n_texts and (time.perf_counter() - t0 < budget_s * 0.4 or i < 8):
                for j, f, R, cand in ex.map(scan_one, range(i, min(n_texts, i + 4 * th))):
                    R_union[f] |= R
                    cands.append((j, f, cand))
                    sample += int(off[j + 1] - off[j])
                i = min(n_texts, i + 4 * th)
            t_scan = time.perf_counter() - t0
            t0 = time.perf_counter()
            surv = dict(zip(active, ex.map(lambda f: orcs[f].pattern_pass(d["sets"][f][0], d["sets"][f][1],
                                                                           R_union[f]), active)))
            t_pass = time.perf_counter() - t0
            t0 = time.perf_counter()
            list(ex.map(lambda c: orcs[c[1]].verify(text(c[0]), d["sets"][c[1]][0], d["sets"][c[1]][1], c[2],
                                                    surv[c[1]]), cands))
            t_ver = time.perf_counter() - t0
        scale = total / max(sample, 1)
        tot = t_build + t_scan * scale + t_pass + t_ver * scale
        res[label] = {"threads": th, "sample_bytes": sample, "sample_texts": len(cands), "build_s": round(t_build, 4),
                      "scan_s": round(t_scan * scale, 4), "pattern_pass_s": round(t_pass, 4),
                      "verify_s": round(t_ver * scale, 4), "total_s": round(tot, 4), "gbs": round(total / tot / 1e9, 4)}
    return {"value": res["T"]["gbs"], "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": (f"oracle: {len(d['sets'])} filters built, then the first {res['T']['sample_texts']} of "
                       f"{n_texts} texts ({res['T']['sample_bytes'] >> 20} MiB) scanned on {threads} threads, one "
                       f"pattern pass per filter, per-text verification; scan and verification scaled to "
                       f"{total >> 20} MiB; 1 thread: {res['1']['gbs']:.4f} GB/s on the first "
                       f"{res['1']['sample_texts']} texts"),
            "single_thread_value": res["1"]["gbs"], "stages": res}


def split_inputs(fa, wl):
    import numpy as np

    W, k, s, c, e, n_pat, lmin, lmax, n_text, line, plants, seed = wl
    sbox64 = np.random.default_rng(seed).integers(0, 2**64, size=256, dtype=np.uint64)
    pb, po = fa.gen_patterns(fa.ALPHA_ASCII, seed * 10 + 1, n_pat, lmin, lmax)
    text = fa.gen_bytes(fa.ALPHA_ASCII, seed * 10 + 2, n_text)
    rng = np.random.default_rng(seed * 10 + 4)
    text[rng.integers(0, n_text, size=n_text // line)] = 10
    fa.plant(seed * 10 + 3, plants, text, pb, po)
    return sbox64, pb, po, text


def run_split(fa, name, steps=10, warmup=2, cpu=True):
    """The split filter's whole path (build, scan, pattern pass, verify,
    sorted matches to the host) over a device-resident text, as the plain
    leg measures the pair method; CPU baseline: the oracle's split method."""
    import numpy as np

    wl = SPLIT_WORKLOADS[name]
    W, k, s, c, e, n_pat, lmin, lmax, n_text, line, plants, seed = wl
    sbox64, pb, po, text = split_inputs(fa, wl)
    dev = fa.DeviceBuffer.from_host(text)
    m = fa.Matcher.split(W, k, s, c, e, sbox64)
    m.set_patterns(pb, po)

    def step():
        m.build()
        m.run_device(dev.ptr, n_text)
        return m.matches()

    for _ in range(warmup):
        step()
    m.synchronize()
    # as in the plain leg: no stage events in the timed steps, the scan
    # kernel's duration from separate steps with them on
    m.set_timing(False)
    t0 = time.perf_counter()
    for _ in range(steps):
        res = step()
    m.synchronize()
    elapsed = time.perf_counter() - t0
    m.set_timing(True)
    scan_ms = []
    for _ in range(max(3, min(steps, 20))):
        res = step()
        scan_ms.append(m.timings()["scan_kernel_ms"])
    t = m.timings()
    t_scan = float(np.mean(scan_ms))
    achieved = n_text / (t_scan * 1e-3) / 1e9
    out = {
        "workload": (f"{name}: split filter W={W} k={k} s={s} cache 2^{c} + external 2^{e} bits, {n_pat} ASCII "
                     f"patterns ({lmin}-{lmax} B), {n_text / 2**20:.0f} MiB ASCII text in lines of ~{line} B, "
                     f"{plants} planted"),
        "value": round(n_text / (elapsed / steps) / 1e9, 3),
        "unit": "GB/s",
        "ms_per_step": round(elapsed * 1e3 / steps, 4),
        "steps": steps,
        "roofline": {"kernel": "k_scan (split)", "bound": "hbm", "achieved": round(achieved, 3),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "avg_launch_ms": round(t_scan, 4), "algorithmic_bytes": n_text},
        "scan_mode": {0: "F probed directly (cache part first)", 1: "LDS fold of the cache part"}.get(t["scan_mode"]),
        "probes": t["probes"],
        "matches": int(res.size),
        "candidates": int(m.n_candidates()),
    }
    m.close()
    dev.free()
    if cpu:
        from oracle import SplitOracle

        log(f"{name}: cpu baseline ...")
        out["cpu_baseline"] = _cpu_baseline(SplitOracle(W, k, s, c, e, sbox64), n_pat, pb, po, text, 12.0)
    return out


def traffic_from_profiles(name):
    p = os.path.join(ROOT, "profiles", "scan_traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get(name)
    except (OSError, ValueError):
        return None


PMC_COPY_BYTES = 1 << 30  # the copy kernel whose counters scale the scan's


def measure_traffic(name, timeout_s=150):
    """HBM bytes per k_scan launch of workload `name`, measured now: two
    rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE) over a child run of
    the same workload (one warmup and two timed launches) followed by three
    launches of the library's copy kernel over PMC_COPY_BYTES.  Each counter is
    scaled by what it reads for the copy, which reads and writes exactly
    PMC_COPY_BYTES.  None when rocprofv3 is missing, when this process already
    runs under a profiler (a nested one would profile itself), or when a pass
    fails; the caller then reports the committed profile's figure and says so."""
    import csv
    import shutil
    import signal
    import subprocess
    import tempfile

    if any(k.startswith("ROCPROF") or k == "ROCP_TOOL_LIBRARIES" for k in os.environ):
        return None
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    got = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="ffbf_pmc_")
        try:
            cmd = [prof, "--pmc", counter, "--kernel-include-regex", "k_scan<|k_copy", "-f", "csv", "-d", d, "-o", "p",
                   "--", sys.executable, os.path.abspath(__file__), "--workload", name, "--steps", "2", "--warmup", "1",
                   "--traffic-child", str(PMC_COPY_BYTES)]
            env = dict(os.environ, TMPDIR=tempfile.gettempdir())
            p = subprocess.Popen(cmd, cwd=d, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                                 start_new_session=True)
            try:
                p.wait(timeout=timeout_s)
            except subprocess.TimeoutExpired:
                os.killpg(p.pid, signal.SIGKILL)
                p.wait()
                log(f"traffic: {counter} pass timed out")
                return None
            # rocprofv3 can exit non-zero after a complete run; the CSV decides
            path = os.path.join(d, "p_counter_collection.csv")
            if not os.path.exists(path):
                log(f"traffic: {counter} pass wrote no counters (rc {p.returncode})")
                return None
            scan, copy = [], []
            with open(path) as f:
                for r in csv.DictReader(f):
                    if r.get("Counter_Name") != counter:
                        continue
                    if "k_scan<" in r["Kernel_Name"]:
                        scan.append(float(r["Counter_Value"]))
                    elif "k_copy" in r["Kernel_Name"]:
                        copy.append(float(r["Counter_Value"]))
            if len(scan) < 3 or len(copy) < 3 or min(copy) <= 0:
                log(f"traffic: {counter} pass incomplete ({len(scan)} scan, {len(copy)} copy dispatches)")
                return None
            scan.sort()
            copy.sort()
            got[counter] = (scan[len(scan) // 2], copy[len(copy) // 2])
        finally:
            shutil.rmtree(d, ignore_errors=True)
    f_scan, f_copy = got["FETCH_SIZE"]
    w_scan, w_copy = got["WRITE_SIZE"]
    f_scale = PMC_COPY_BYTES / (f_copy * 1024.0)
    w_scale = PMC_COPY_BYTES / (w_copy * 1024.0)
    return {
        "bytes": int(round(f_scan * 1024.0 * f_scale + w_scan * 1024.0 * w_scale)),
        "fetch_kib": round(f_scan, 1),
        "write_kib": round(w_scan, 1),
        "fetch_scale": round(f_scale, 4),
        "write_scale": round(w_scale, 4),
        "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE in this run (median of 3 scan launches), "
                  f"scaled by the same counters on a {PMC_COPY_BYTES >> 20} MiB copy kernel",
    }


def run_plain(fa, name, steps, warmup, mode=-1, copy_rate=True, traffic=None):
    """The single-text path over a device-resident text: each step builds F,
    scans with feed-forward, runs the pattern pass, verifies and brings the
    sorted matches to the host.  Returns the bench fields and the inputs."""
    import numpy as np

    wl = WORKLOADS[name]
    alpha, W, k, log2, n_pat, lmin, lmax, n_text, plants, seed = wl
    t0 = time.perf_counter()
    sbox, pb, po, text = make_inputs(fa, wl)
    log(f"{name}: inputs generated in {time.perf_counter() - t0:.1f} s")
    dev = fa.DeviceBuffer.from_host(text)
    m = fa.Matcher(W, k, log2, sbox)
    m.set_patterns(pb, po)
    m.set_scan_mode(mode)

    def step():
        m.build()
        m.run_device(dev.ptr, n_text)
        return m.matches()

    for _ in range(warmup):
        step()
    m.synchronize()
    # the timed steps record no stage events (each costs ~4-5 us of device
    # time between dependent kernels, ~10 per step); the stage times and the
    # scan kernel's duration come from separate steps with the events on
    m.set_timing(False)
    t0 = time.perf_counter()
    for _ in range(steps):
        res = step()
    m.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"{name}: timed {steps} steps in {elapsed:.3f} s")
    ms_per_step = elapsed * 1e3 / steps
    m.set_timing(True)
    stage_steps = max(3, min(steps, 20))
    scan_ms, stages = [], []
    for _ in range(stage_steps):
        res = step()
        t = m.timings()
        scan_ms.append(t["scan_kernel_ms"])
        stages.append((t["build_ms"], t["scan_ms"], t["pattern_pass_ms"], t["verify_ms"]))
    value = n_text / (ms_per_step * 1e-3) / 1e9
    t_scan = float(np.mean(scan_ms))
    achieved = n_text / (t_scan * 1e-3) / 1e9
    st = np.mean(np.array(stages), axis=0)
    tm = m.timings()
    ceiling = probe_ceiling(fa, m, k, n_text, W, t_scan)
    out = {
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded counter-based generator)",
        "config": describe(name, wl),
        "hbm_frac_of_step": round(value / HBM_PEAK_GBS, 5),
        "roofline": {
            "kernel": "k_scan",
            "bound": "hbm",
            # what holds the kernel below that roof, from its L2 and SQ counters
            "limiter": ("L2 random-request rate"
                        if tm["scan_mode"] in (0, 3) or (ceiling and ceiling["lds_pass"] > 0.04)
                        else "VALU/LDS instruction issue"),
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic["bytes"] if traffic else traffic_from_profiles(name),
            "traffic_source": traffic["source"] if traffic else "committed profile (profiles/scan_traffic.json)",
            "traffic_counters": ({k: traffic[k] for k in ("fetch_kib", "write_kib", "fetch_scale", "write_scale")}
                                 if traffic else None),
            "avg_launch_ms": round(t_scan, 4),
            "algorithmic_bytes": n_text,
            "copy_kernel_gbs": copy_kernel_rate(fa) if copy_rate else None,
        },
        "stages_ms": {"build": round(st[0], 4), "scan": round(st[1], 4), "pattern_pass": round(st[2], 4),
                      "verify": round(st[3], 4)},
        "stage_steps": stage_steps,  # steps with stage events, after the timed ones
        "scan_mode": {0: "F probed in L2/HBM", 1: "LDS fold prefilter", 2: "filter in LDS",
                      3: "L2-resident global fold", 4: "LDS fold, then L2-resident global fold"}.get(tm["scan_mode"]),
        "fold_log2_bits": tm["fold_log2_bits"],
        "l2_fold_log2_bits": tm["l2_fold_log2_bits"],
        "probes": tm["probes"],
        "l2_probe_ceiling": ceiling,
        "matches": int(res.size),
        "candidates": int(m.n_candidates()),
    }
    m.close()
    dev.free()
    return out, (sbox, pb, po, text)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="plain",
                    choices=sorted(WORKLOADS) + sorted(SPLIT_WORKLOADS) +
                    ["batch-4096texts-1set", "batch-2048texts-64sets", "serve-8sets", "ingest"])
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched, serving and ingest legs of the default run")
    ap.add_argument("--mode", type=int, default=-1, help="scan mode override (-1 auto)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES (HBM counter reference) and exit")
    ap.add_argument("--gather-sweep", action="store_true",
                    help="only run the random-gather microbenchmark (4-byte probes into 16 KiB..2 GiB) and exit")
    ap.add_argument("--gather-log2", type=int, default=0, metavar="L",
                    help="only run the random-gather microbenchmark into a 2^L-byte array (two launches) and exit")
    ap.add_argument("--probe-sweep", action="store_true",
                    help="only run the access-kind microbenchmark (random loads of four cache policies and "
                         "device-scope atomicOr into 1 MiB..256 MiB; streaming passes over 16 MiB..4 GiB; "
                         "pinned-host uploads by the upload kernel of 1 MiB..1 GiB) and exit")
    ap.add_argument("--serve-upload", choices=["default", "copy", "kernel"], default="default",
                    help="how the serving leg's request text reaches the device (server upload method)")
    ap.add_argument("--serve-gap", type=int, default=0, metavar="BYTES",
                    help="bytes left between consecutive serving requests in their pinned buffer (0: back to "
                         "back, as in a receive ring; >0: no two requests adjacent, as in separate buffers)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc passes that measure the scan's HBM traffic")
    ap.add_argument("--traffic-child", type=int, default=0, metavar="BYTES", help=argparse.SUPPRESS)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU sharding is design-only "
                                   "(text slices with a W-1 byte overlap, local filters, gathered match lists)",
                          "n_gpus": args.gpus}))
        return 2

    # the scan's HBM traffic, from counters of child runs, before this process
    # touches the GPU
    traffic = None
    if (not args.traffic_child and not args.no_traffic and args.workload in WORKLOADS and
            not (args.copy_calibration or args.gather_log2 or args.gather_sweep or args.probe_sweep)):
        log("traffic: two rocprofv3 --pmc passes ...")
        traffic = measure_traffic(args.workload)

    import numpy as np

    import ffbf_amd as fa

    if args.traffic_child:
        # a profiled child of measure_traffic: the workload's scans, then the
        # copy kernel whose counters scale theirs
        run_plain(fa, args.workload, args.steps, args.warmup, args.mode, copy_rate=False)
        n = args.traffic_child & ~15
        src, dst = fa.DeviceBuffer(n), fa.DeviceBuffer(n)
        for _ in range(3):
            dst.copy_from(src, n)
        src.free()
        dst.free()
        return 0

    if args.copy_calibration:
        n = args.copy_calibration & ~15
        src, dst = fa.DeviceBuffer(n), fa.DeviceBuffer(n)
        for _ in range(3):
            dst.copy_from(src, n)
        print(json.dumps({"copy_calibration_bytes": n, "launches": 3}), flush=True)
        return 0

    if args.gather_log2:
        ms, rate = fa.gather_bench(args.gather_log2, steps=16)
        print(json.dumps({"gather_array_bytes": 1 << args.gather_log2, "ms": round(ms, 4),
                          "probes_per_s": round(rate, 1)}), flush=True)
        return 0

    if args.probe_sweep:
        for kind in (fa.PROBE_LOAD, fa.PROBE_LOAD_NONTEMPORAL, fa.PROBE_LOAD_AGENT, fa.PROBE_LOAD_SYSTEM,
                     fa.PROBE_ATOMIC_OR):
            for l in (20, 22, 25, 28):
                ms, rate = fa.probe_bench(kind, l, steps=16)
                print(json.dumps({"kind": fa.PROBE_KIND_NAMES[kind], "array_bytes": 1 << l, "ms": round(ms, 4),
                                  "accesses_per_s": round(rate, 1)}), flush=True)
        for l in range(24, 33):
            ms, rate = fa.probe_bench(fa.PROBE_STREAM, l, steps=max(4, 1 << (33 - l)))
            print(json.dumps({"kind": fa.PROBE_KIND_NAMES[fa.PROBE_STREAM], "array_bytes": 1 << l,
                              "ms": round(ms, 4), "bytes_per_s": round(rate, 1)}), flush=True)
        for l in (20, 24, 27, 30):
            ms, rate = fa.probe_bench(fa.PROBE_UPLOAD, l, steps=max(4, 1 << (31 - l)))
            print(json.dumps({"kind": fa.PROBE_KIND_NAMES[fa.PROBE_UPLOAD], "buffer_bytes": 1 << l,
                              "ms": round(ms, 4), "bytes_per_s": round(rate, 1)}), flush=True)
        return 0

    if args.gather_sweep:
        for l in range(14, 32):
            ms, rate = fa.gather_bench(l, steps=16 if l < 28 else 8)
            print(json.dumps({"gather_array_bytes": 1 << l, "ms": round(ms, 4), "probes_per_s": round(rate, 1)}),
                  flush=True)
        return 0

    if args.workload == "ingest":
        r = run_ingest(fa, reps=max(1, min(args.steps, 5)))
        print(json.dumps({"metric": METRIC, "value": r["host_pipelined_gbs"], "unit": "GB/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "u32",
                          "data": "synthetic (seeded counter-based generator)",
                          "config": {"workload": "ingest", "text_bytes": r["text_bytes"]}, "ingest": r}), flush=True)
        return 0

    if args.workload in SPLIT_WORKLOADS:
        r = run_split(fa, args.workload, steps=args.steps, warmup=args.warmup, cpu=not args.no_cpu)
        print(json.dumps({"metric": METRIC, "value": r["value"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "u64",
                          "data": "synthetic (seeded counter-based generator)",
                          "config": {"workload": r["workload"]}, "roofline": r["roofline"],
                          "cpu_baseline": r.get("cpu_baseline"), "split": r}), flush=True)
        return 0

    if args.workload == "serve-8sets":
        r = run_serving(fa, n_requests=max(64, args.steps * 100), upload=UPLOADS[args.serve_upload],
                        gap=args.serve_gap)
        print(json.dumps({"metric": METRIC, "value": r["saturation_gbs"], "unit": "GB/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "u32",
                          "data": "synthetic (seeded counter-based generator)",
                          "config": {"workload": "serve-8sets", "requests": r["requests"], "text_bytes": r["text_bytes"]},
                          "serving": r}), flush=True)
        return 0
    if args.workload in ("batch-4096texts-1set", "batch-2048texts-64sets"):
        r = run_batched(fa, args.workload, steps=args.steps)
        print(json.dumps({"metric": METRIC, "value": r["value"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "ms_per_step": r["ms_per_batch"], "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "u32",
                          "data": "synthetic (seeded counter-based generator)",
                          "config": {"workload": r["workload"], "texts": r["texts"], "text_bytes": r["text_bytes"],
                                     "pattern_sets": r["pattern_sets"]},
                          "batched": r}), flush=True)
        return 0

    out, inputs = run_plain(fa, args.workload, args.steps, args.warmup, args.mode, traffic=traffic)
    out = {"metric": METRIC, **out}
    if args.workload == "plain" and not args.no_batched:
        log("config 1 leg (ascii10k-64m) ...")
        # a 0.2 ms step, bound by launch and completion latency: 200 timed
        # steps (~45 ms).  Right after the 2 GiB leg the same steps take ~30%
        # longer for about two seconds (the 2 GiB steps themselves do not
        # change), so the leg is timed then and again after a 2 s pause; the
        # settled figure is the leg's value, the other is reported beside it
        c1_hot, _ = run_plain(fa, "ascii10k-64m", 200, 20, -1, copy_rate=False)
        time.sleep(2.0)
        c1, c1_in = run_plain(fa, "ascii10k-64m", 200, 20, -1, copy_rate=False)
        c1["settle_s"] = 2.0
        c1["right_after_2gib_leg"] = {"value": c1_hot["value"], "ms_per_step": c1_hot["ms_per_step"]}
        if not args.no_cpu:
            c1["cpu_baseline"] = cpu_baseline(WORKLOADS["ascii10k-64m"], *c1_in, budget_s=6.0)
        out["config1"] = c1
        log("batched leg (batch-4096texts-1set) ...")
        out["batched"] = run_batched(fa, "batch-4096texts-1set", steps=10, cpu=not args.no_cpu)
        log("multi-filter batched leg (batch-2048texts-64sets) ...")
        out["batched_multi_filter"] = run_batched(fa, "batch-2048texts-64sets", steps=5, per_text_reps=1,
                                                  cpu=not args.no_cpu)
        log("serving leg (serve-8sets) ...")
        out["serving"] = run_serving(fa, n_requests=600, upload=UPLOADS[args.serve_upload], gap=args.serve_gap)
        log("ingest leg (host memory / file) ...")
        out["ingest"] = run_ingest(fa, reps=2)
        log("split filter leg (split-1m-2g-lines) ...")
        out["split_method"] = run_split(fa, "split-1m-2g-lines", steps=10, cpu=not args.no_cpu)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(WORKLOADS[args.workload], *inputs)
        if args.workload == "plain" and not args.no_batched:
            # the CPU stages of configs 1-4 side by side (1 thread / cpu_threads())
            out["cpu_table"] = {
                "config1 ascii10k-64m": out["config1"]["cpu_baseline"]["stages"],
                "config2 plain": out["cpu_baseline"]["stages"],
                "config3 batch-4096texts-1set": out["batched"]["cpu_baseline"]["stages"],
                "config4 batch-2048texts-64sets": out["batched_multi_filter"]["cpu_baseline"]["stages"],
            }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
