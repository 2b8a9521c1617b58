"""Full-size parity: the HIP path against the oracle run over the WHOLE text
(on the host's cores) at the sizes the benchmark measures, not on a sample.

* 1,000,000 DNA patterns (32..100 bytes, GC 40%), a 2 GiB text with 50,000
  planted occurrences, W=32, k=4, a 2^28-bit filter: forward filter, every
  candidate offset, the reverse filter, the survivor bitmap and the match
  list are compared bit for bit, in the automatic mode (an L2-resident fold),
  with F probed directly, and with the LDS fold (which lets ~93% of windows
  through, so F's first bit is tested before queueing).
* 100,000 ASCII patterns over a 1.5 GiB text (W=19, k=4, 2^26 bits) in the
  LDS-fold, L2-fold and direct modes.
* The split filter (W=19, k=5, s=2, 2^24 + 2^28 bits), 1,000,000 ASCII
  patterns, a 2 GiB text cut into lines.
* A 4.2 GiB ASCII text (10,000 patterns, W=19, k=4, 2^24 bits) with patterns
  planted after 2^32, across the 2^32 boundary, at 2^31 and at the very end: offsets
  beyond 32 bits through every stage.

Needs an MI355X (-m gpu)."""
import os
import sys
import time

import numpy as np
import pytest

import ffbf_amd as fa
from oracle import Oracle

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_all(o, text, pb, po):
    t0 = time.perf_counter()
    r = o.run(text, pb, po, threads=THREADS)
    say(f"  oracle over {text.size >> 20} MiB on {THREADS} threads: {time.perf_counter() - t0:.1f} s, "
        f"{r['cand'].size} candidates, {r['ids'].size} matches")
    return r


def gpu_staged(m, dev, n):
    m.scan_device(dev.ptr, n)
    out = {"cand": m.candidates(), "R": m.filter(fa.FILTER_REVERSE)}
    out["n_surv"] = m.pattern_pass()
    out["surv"] = m.survivors()
    out["n_match"] = m.verify()
    out["matches"] = m.matches()
    return out


def assert_same(g, r):
    assert np.array_equal(g["cand"], r["cand"]), f"candidates differ: {g['cand'].size} vs {r['cand'].size}"
    assert np.array_equal(g["R"], r["R"]), "reverse filter differs"
    assert np.array_equal(g["surv"], r["surv"]), "survivor bitmap differs"
    assert g["n_surv"] == int(np.unpackbits(r["surv"].view(np.uint8)).sum())
    assert np.array_equal(g["matches"]["pattern_id"], r["ids"]), "match ids differ"
    assert np.array_equal(g["matches"]["offset"], r["offs"]), "match offsets differ"


def intact_plants(text, pb, po, ids, offs):
    """(id, offset) of the planted copies that later plants did not overwrite."""
    keep = []
    for i, q in zip(ids.tolist(), offs.tolist()):
        a, b = int(po[i]), int(po[i + 1])
        if np.array_equal(text[q:q + b - a], pb[a:b]):
            keep.append((i, q))
    return keep


def test_full_size_1m_dna_patterns_2gib_text():
    W, K, LOG2, N_TEXT = 32, 4, 28, 2 << 30
    say("full size: 1M DNA patterns, 2 GiB text")
    t0 = time.perf_counter()
    sbox = fa.sbox_from_seed(2)
    pb, po = fa.gen_patterns(fa.ALPHA_DNA_GC40, 21, 1_000_000, 32, 100)
    text = fa.gen_bytes(fa.ALPHA_DNA_GC40, 22, N_TEXT)
    pid, poff = fa.plant(23, 50_000, text, pb, po)
    say(f"  inputs in {time.perf_counter() - t0:.1f} s")
    r = oracle_all(Oracle(W, K, LOG2, sbox), text, pb, po)
    planted = intact_plants(text, pb, po, pid, poff)
    assert len(planted) > 45_000
    found = set(zip(r["ids"].tolist(), r["offs"].tolist()))
    assert all(p in found for p in planted), "a planted occurrence is missing from the oracle's list"
    dev = fa.DeviceBuffer.from_host(text)
    with fa.Matcher(W, K, LOG2, sbox) as m:
        m.set_patterns(pb, po)
        m.build()
        assert np.array_equal(m.filter(fa.FILTER_FORWARD), r["F"]), "forward filter differs"
        # the LDS fold passes ~93% of windows: F (mode 1) or the L2 fold (mode 4) tested before queueing
        for mode in (fa.MODE_AUTO, fa.MODE_GLOBAL, fa.MODE_FOLD, fa.MODE_GFOLD, fa.MODE_HYBRID):
            m.set_scan_mode(mode)
            say(f"  mode request {mode}: scan mode {m.timings()['scan_mode']}")
            assert_same(gpu_staged(m, dev, N_TEXT), r)
        # fused path (one sync), automatic mode
        m.set_scan_mode(fa.MODE_AUTO)
        assert m.run_device(dev.ptr, N_TEXT) == r["ids"].size
        mm = m.matches()
        assert np.array_equal(mm["pattern_id"], r["ids"]) and np.array_equal(mm["offset"], r["offs"])
    dev.free()


def test_text_beyond_4gib():
    W, K, LOG2 = 19, 4, 24
    n = (4 << 30) + (200 << 20) + 13
    say(f"text of {n} bytes")
    t0 = time.perf_counter()
    sbox = fa.sbox_from_seed(31)
    pb, po = fa.gen_patterns(fa.ALPHA_ASCII, 32, 10_000, 20, 40)
    text = fa.gen_bytes(fa.ALPHA_ASCII, 33, n)
    fa.plant(34, 2_000, text, pb, po)
    hi = 1 << 32
    tail = text[hi:]  # a view: plants land after 2^32
    fa.plant(35, 500, tail, pb, po)
    # one pattern across the 2^32 boundary, one starting exactly at 2^31, one
    # ending at the last byte
    edges = ((7, hi - 9), (8, 1 << 31), (9, n - int(po[10] - po[9])))
    for pid, q in edges:
        a, b = int(po[pid]), int(po[pid + 1])
        text[q:q + b - a] = pb[a:b]
    say(f"  inputs in {time.perf_counter() - t0:.1f} s")
    r = oracle_all(Oracle(W, K, LOG2, sbox), text, pb, po)
    found = set(zip(r["ids"].tolist(), r["offs"].tolist()))
    for p in edges:
        assert p in found, p
    assert int((r["offs"] >= hi).sum()) > 400
    dev = fa.DeviceBuffer.from_host(text)
    with fa.Matcher(W, K, LOG2, sbox) as m:
        m.set_patterns(pb, po)
        m.build()
        assert np.array_equal(m.filter(fa.FILTER_FORWARD), r["F"]), "forward filter differs"
        for mode in (fa.MODE_AUTO, fa.MODE_GLOBAL):
            m.set_scan_mode(mode)
            say(f"  mode request {mode}: scan mode {m.timings()['scan_mode']}")
            assert_same(gpu_staged(m, dev, n), r)
        m.set_scan_mode(fa.MODE_AUTO)
        assert m.run_device(dev.ptr, n) == r["ids"].size
        mm = m.matches()
        assert np.array_equal(mm["pattern_id"], r["ids"]) and np.array_equal(mm["offset"], r["offs"])
    dev.free()


def test_full_size_100k_ascii_patterns_1536mib_text():
    """100,000 ASCII patterns (20..64 bytes), W=19, k=4, 2^26 bits, a 1.5 GiB
    text with 1,536 plants: the LDS-fold scan, the L2 fold and F directly."""
    W, K, LOG2, N_TEXT = 19, 4, 26, 1536 << 20
    say("full size: 100k ASCII patterns, 1.5 GiB text")
    sbox = fa.sbox_from_seed(3)
    pb, po = fa.gen_patterns(fa.ALPHA_ASCII, 31, 100_000, 20, 64)
    text = fa.gen_bytes(fa.ALPHA_ASCII, 32, N_TEXT)
    fa.plant(33, 1_536, text, pb, po)
    r = oracle_all(Oracle(W, K, LOG2, sbox), text, pb, po)
    assert r["ids"].size >= 1_500
    dev = fa.DeviceBuffer.from_host(text)
    with fa.Matcher(W, K, LOG2, sbox) as m:
        m.set_patterns(pb, po)
        m.build()
        assert np.array_equal(m.filter(fa.FILTER_FORWARD), r["F"]), "forward filter differs"
        for mode in (fa.MODE_AUTO, fa.MODE_GFOLD, fa.MODE_HYBRID, fa.MODE_GLOBAL):
            m.set_scan_mode(mode)
            say(f"  mode request {mode}: scan mode {m.timings()['scan_mode']}")
            assert_same(gpu_staged(m, dev, N_TEXT), r)
    dev.free()


def test_full_size_split_1m_patterns_2gib_lines():
    """The CPU binary's split filter at its compile-time shape (W=19, k=5,
    s=2, 2^24-bit cache part, 2^28-bit external part), 1,000,000 ASCII
    patterns and a 2 GiB text cut into lines of mean length 80 with 50,000
    plants: the automatic choice, the cache-part fold in LDS and F probed
    directly."""
    from oracle import SplitOracle

    W, K, S, C, E, N_TEXT = 19, 5, 2, 24, 28, 2 << 30
    say("full size: split filter, 1M patterns, 2 GiB of lines")
    sbox64 = np.random.default_rng(5).integers(0, 2**64, size=256, dtype=np.uint64)
    pb, po = fa.gen_patterns(fa.ALPHA_ASCII, 51, 1_000_000, 20, 64)
    text = fa.gen_bytes(fa.ALPHA_ASCII, 52, N_TEXT)
    rng = np.random.default_rng(54)
    text[rng.integers(0, N_TEXT, size=N_TEXT // 80)] = 10
    fa.plant(53, 50_000, text, pb, po)
    r = oracle_all(SplitOracle(W, K, S, C, E, sbox64), text, pb, po)
    assert r["ids"].size > 30_000
    dev = fa.DeviceBuffer.from_host(text)
    with fa.Matcher.split(W, K, S, C, E, sbox64) as m:
        m.set_patterns(pb, po)
        m.build()
        assert np.array_equal(m.filter(fa.FILTER_FORWARD), r["F"]), "forward filter differs"
        for mode in (fa.MODE_AUTO, fa.MODE_FOLD, fa.MODE_GLOBAL):
            m.set_scan_mode(mode)
            say(f"  mode request {mode}: scan mode {m.timings()['scan_mode']}")
            assert_same(gpu_staged(m, dev, N_TEXT), r)
    dev.free()


@pytest.mark.parametrize("name", ["batch-4096texts-1set", "batch-2048texts-64sets"])
def test_full_size_batches(name):
    """The two benchmarked batches in full (4,096 ASCII texts, 1.5 GiB, one
    100k-pattern set; 2,048 DNA texts with tandem repeats, 3.3 GiB, 64 sets
    of 1k..1M patterns): every text's matches against the oracle's flow
    (scan, pattern pass, verification) on that text alone, and each filter's
    R against the union of its texts' reverse filters."""
    from concurrent.futures import ThreadPoolExecutor

    from bench import batch_inputs

    t0 = time.perf_counter()
    d = batch_inputs(fa, name)
    say(f"{name}: {d['n_texts']} texts, {d['bytes'] >> 20} MiB, generated in {time.perf_counter() - t0:.1f} s")
    W, K, sbox, texts, off, route = d["W"], d["k"], d["sbox"], d["texts"], d["offsets"], d["route"]
    t0 = time.perf_counter()
    orc = [Oracle(W, K, log2, sbox) for _, _, log2 in d["sets"]]
    Fs = [o.build(pb, po) for o, (pb, po, _) in zip(orc, d["sets"])]

    def one(i):
        f = int(route[i])
        pb, po, _ = d["sets"][f]
        t = texts[int(off[i]):int(off[i + 1])]
        R, cand = orc[f].scan(t, Fs[f])
        ids, offs = orc[f].verify(t, pb, po, cand, orc[f].pattern_pass(pb, po, R))
        return f, R, ids, offs

    rows, R_union = [], {}
    with ThreadPoolExecutor(THREADS) as ex:
        for lo in range(0, d["n_texts"], 256):
            for i, (f, R, ids, offs) in zip(range(lo, d["n_texts"]), ex.map(one, range(lo, min(lo + 256, d["n_texts"])))):
                rows.append(np.rec.fromarrays([np.full(ids.size, i, np.uint32), ids, offs]))
                R_union[f] = R if f not in R_union else (R_union[f] | R)
    expect = np.concatenate(rows)
    expect.sort(order=["f0", "f1", "f2"])
    say(f"  oracle on {THREADS} threads: {time.perf_counter() - t0:.1f} s, {expect.size} matches")
    assert expect.size >= d["n_texts"]
    dev = fa.DeviceBuffer.from_host(texts)
    ms = []
    for pb, po, log2 in d["sets"]:
        m = fa.Matcher(W, K, log2, sbox)
        m.set_patterns(pb, po)
        m.build()
        ms.append(m)
    with fa.Batch(ms) as bt:
        bt.set_texts(dev.ptr, off, route)
        for use_graph in (True, False):
            bt.set_tuning(use_graph=use_graph)
            n = bt.run()
            got = bt.matches()
            say(f"  graph={use_graph}: {n} matches")
            assert n == expect.size
            assert np.array_equal(got["text_id"], expect["f0"]), "text ids differ"
            assert np.array_equal(got["pattern_id"], expect["f1"]), "pattern ids differ"
            assert np.array_equal(got["offset"], expect["f2"]), "offsets differ"
            for f, m in enumerate(ms):
                R = m.filter(fa.FILTER_REVERSE)
                assert np.array_equal(R, R_union[f]) if f in R_union else not R.any(), f"R of filter {f}"
    for m in ms:
        m.close()
    dev.free()
