"""Several host threads, each driving its own context (and one a batch), on
one GPU at the same time: every result equals the oracle's.  Contexts own
their streams and device memory, so calls on different contexts may run
concurrently.  Needs an MI355X (-m gpu)."""
import sys
import threading

import numpy as np
import pytest

import ffbf_amd as fa
from oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# alphabet, W, k, log2 bits, patterns, lmin, lmax, scan mode
SHAPES = [
    (fa.ALPHA_ASCII, 19, 4, 16, 2000, 19, 40, fa.MODE_LDS),
    (fa.ALPHA_ASCII, 19, 4, 22, 20000, 20, 60, fa.MODE_FOLD),
    (fa.ALPHA_DNA_GC40, 32, 4, 22, 30000, 32, 80, fa.MODE_GFOLD),
    (fa.ALPHA_DNA_GC40, 24, 3, 18, 5000, 24, 60, fa.MODE_GLOBAL),
    (fa.ALPHA_ASCII, 13, 2, 20, 8000, 13, 30, fa.MODE_AUTO),
]


def test_contexts_on_concurrent_threads():
    rounds = 4
    cases = []
    for i, (alpha, W, k, log2, n, lmin, lmax, mode) in enumerate(SHAPES):
        sbox = fa.sbox_from_seed(300 + i)
        b, off = fa.gen_patterns(alpha, 310 + i, n, lmin, lmax)
        o = Oracle(W, k, log2, sbox)
        texts, refs = [], []
        for r in range(rounds):
            t = fa.gen_bytes(alpha, 320 + 10 * i + r, 300_000 + 77_777 * r)
            fa.plant(330 + 10 * i + r, 60, t, b, off)
            texts.append(t)
            refs.append(o.run(t, b, off, threads=4))
        cases.append((sbox, b, off, W, k, log2, mode, texts, refs))
    errors = []
    start = threading.Barrier(len(cases))

    def worker(ci):
        sbox, b, off, W, k, log2, mode, texts, refs = cases[ci]
        try:
            with fa.Matcher(W, k, log2, sbox) as m:
                m.set_patterns(b, off)
                m.build()
                m.set_scan_mode(mode)
                start.wait(timeout=60)
                for rep in range(3):
                    for t, r in zip(texts, refs):
                        if rep % 2 == 0:
                            mm = m.match(t).matches
                            ids, offs = mm["pattern_id"], mm["offset"]
                        else:
                            buf = fa.DeviceBuffer.from_host(t)
                            m.run_device(buf.ptr, t.size)
                            mm = m.matches()
                            buf.free()
                            ids, offs = mm["pattern_id"], mm["offset"]
                        if not (np.array_equal(ids, r["ids"]) and np.array_equal(offs, r["offs"])):
                            errors.append(f"context {ci} rep {rep}: matches differ")
                        if not np.array_equal(m.filter(fa.FILTER_REVERSE), r["R"]):
                            errors.append(f"context {ci} rep {rep}: reverse filter differs")
        except Exception as e:  # reported below, on the main thread
            start.abort()
            errors.append(f"context {ci}: {e!r}")

    threads = [threading.Thread(target=worker, args=(ci,), daemon=True) for ci in range(len(cases))]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads), "a worker did not finish"
    assert not errors, errors[:5]
    say(f"{len(cases)} threads x {3 * rounds} scans agree with the oracle")


def test_batch_beside_plain_contexts():
    """A batch runs on one thread while two plain contexts scan on two others."""
    sbox = fa.sbox_from_seed(340)
    W, k = 24, 3
    sets = [fa.gen_patterns(fa.ALPHA_DNA_GC40, 341 + f, 3000 + 2000 * f, 24, 60) for f in range(3)]
    o = [Oracle(W, k, 18, sbox) for _ in sets]
    rng = np.random.default_rng(342)
    lens = rng.integers(1000, 200_000, size=40)
    route = rng.integers(0, len(sets), size=40).astype(np.uint32)
    texts = []
    for i, (n, f) in enumerate(zip(lens.tolist(), route.tolist())):
        t = fa.gen_bytes(fa.ALPHA_DNA_GC40, 350 + i, n)
        fa.plant(400 + i, 5, t, *sets[f])
        texts.append(t)
    off = np.zeros(len(texts) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(lens)
    expect = []
    for i, t in enumerate(texts):
        f = int(route[i])
        r = o[f].run(t, *sets[f])
        expect += [(i, int(p), int(q)) for p, q in zip(r["ids"], r["offs"])]
    expect.sort()
    plain_texts = [fa.gen_bytes(fa.ALPHA_DNA_GC40, 460 + j, 500_000) for j in range(2)]
    for j, t in enumerate(plain_texts):
        fa.plant(470 + j, 80, t, *sets[j])
    plain_refs = [o[j].run(plain_texts[j], *sets[j]) for j in range(2)]
    errors = []
    dev = fa.DeviceBuffer.from_host(np.concatenate(texts))
    ms = []
    for pb, po in sets:
        m = fa.Matcher(W, k, 18, sbox)
        m.set_patterns(pb, po)
        m.build()
        ms.append(m)
    start = threading.Barrier(3)

    def batch_worker():
        try:
            with fa.Batch(ms) as bt:
                bt.set_texts(dev.ptr, off, route)
                start.wait(timeout=60)
                for rep in range(6):
                    bt.run()
                    got = bt.matches()
                    rows = list(zip(got["text_id"].tolist(), got["pattern_id"].tolist(), got["offset"].tolist()))
                    if rows != expect:
                        errors.append(f"batch rep {rep}: matches differ")
        except Exception as e:
            start.abort()
            errors.append(f"batch: {e!r}")

    def plain_worker(j):
        try:
            with fa.Matcher(W, k, 18, sbox) as m:
                m.set_patterns(*sets[j])
                m.build()
                start.wait(timeout=60)
                for rep in range(6):
                    mm = m.match(plain_texts[j]).matches
                    if not (np.array_equal(mm["pattern_id"], plain_refs[j]["ids"])
                            and np.array_equal(mm["offset"], plain_refs[j]["offs"])):
                        errors.append(f"plain {j} rep {rep}: matches differ")
        except Exception as e:
            start.abort()
            errors.append(f"plain {j}: {e!r}")

    threads = [threading.Thread(target=batch_worker, daemon=True)] + [threading.Thread(target=plain_worker, args=(j,), daemon=True)
                                                          for j in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads), "a worker did not finish"
    for m in ms:
        m.close()
    dev.free()
    assert not errors, errors[:5]
