"""Batched scan (many texts x many resident filters, one persistent launch,
hipGraph replay) against the oracle text by text, against the per-text
launch path, and across every tuning.  Needs an MI355X (-m gpu)."""
import sys

import numpy as np
import pytest

from conftest import bits_of

import ffbf_amd as fa
from oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


W, K = 24, 3


def make_filters(sbox, specs, seed):
    """specs: list of (n_patterns, log2_bits, alphabet, scan mode); returns matchers and pattern sets."""
    out = []
    for i, (n, log2, alpha, mode) in enumerate(specs):
        b, off = fa.gen_patterns(alpha, seed + i, n, W, 60)
        m = fa.Matcher(W, K, log2, sbox)
        m.set_patterns(b, off)
        m.build()
        m.set_scan_mode(mode)
        assert m.timings()["scan_mode"] == mode
        out.append((m, b, off, log2))
    return out


def make_texts(filters, n_texts, seed, max_len=200_000):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len, size=n_texts)
    lens[:4] = [0, 5, W - 1, W]  # empty and too-short texts
    route = rng.integers(0, len(filters), size=n_texts).astype(np.uint32)
    texts = []
    for i, (n, f) in enumerate(zip(lens.tolist(), route.tolist())):
        alpha = fa.ALPHA_DNA_GC40 if f % 2 == 0 else fa.ALPHA_ASCII
        t = fa.gen_bytes(alpha, seed * 1000 + i, n)
        _, b, off, _ = filters[f]
        if n > 100:
            fa.plant(seed * 7 + i, 1 + n // 20000, t, b, off)
        texts.append(t)
    off = np.zeros(n_texts + 1, dtype=np.uint64)
    off[1:] = np.cumsum(lens)
    return texts, off, route


def oracle_batch(sbox, filters, texts, route):
    """Sorted (text, pattern, offset) rows and, per filter, the OR of the
    reverse filters of the texts routed to it."""
    rows, R = [], {}
    for i, t in enumerate(texts):
        f = int(route[i])
        _, b, off, log2 = filters[f]
        r = Oracle(W, K, log2, sbox).run(t, b, off)
        rows += [(i, int(p), int(q)) for p, q in zip(r["ids"], r["offs"])]
        R[f] = r["R"] if f not in R else (R[f] | r["R"])
    return sorted(rows), R


def as_rows(m):
    return list(zip(m["text_id"].tolist(), m["pattern_id"].tolist(), m["offset"].tolist()))


SPECS = [
    (500, 14, fa.ALPHA_DNA_GC40, fa.MODE_LDS),       # whole filter in LDS
    (3000, 22, fa.ALPHA_ASCII, fa.MODE_FOLD),        # LDS fold
    (20000, 18, fa.ALPHA_DNA_GC40, fa.MODE_GLOBAL),  # F probed in global memory
    (60000, 26, fa.ALPHA_ASCII, fa.MODE_GFOLD),      # L2-sized global fold
    (200000, 27, fa.ALPHA_DNA_GC40, fa.MODE_HYBRID),  # LDS fold, then the L2-sized fold
]


def test_batch_equals_oracle_and_per_text():
    say("batch vs oracle")
    sbox = fa.sbox_from_seed(77)
    filters = make_filters(sbox, SPECS, 300)
    modes = [f[0].timings()["scan_mode"] for f in filters]
    say(f"  filter modes {modes}")
    texts, off, route = make_texts(filters, 60, 5)
    dev = fa.DeviceBuffer.from_host(np.concatenate(texts) if texts else np.zeros(1, np.uint8))
    expect, R = oracle_batch(sbox, filters, texts, route)
    assert len(expect) > 20
    with fa.Batch([f[0] for f in filters]) as bt:
        bt.set_texts(dev.ptr, off, route)
        n = bt.run()
        got = as_rows(bt.matches())
        assert n == len(expect)
        assert got == expect
        # a filter's R after the batch is the union over its texts
        for f, (m, _, _, _) in enumerate(filters):
            if f in R:
                assert np.array_equal(m.filter(fa.FILTER_REVERSE), R[f]), f"R of filter {f}"
        n2 = bt.run_per_text()
        assert as_rows(bt.matches()) == expect and n2 == n
        # a second replay of the same graph gives the same answer
        bt.run()
        assert as_rows(bt.matches()) == expect
    dev.free()
    for f in filters:
        f[0].close()


@pytest.mark.parametrize("tuning", [
    dict(chunk_iters=1, probes=1, block=1024, use_graph=True, bytes_per_lane=32),
    dict(chunk_iters=7, probes=2, block=512, use_graph=False, bytes_per_lane=16),
    dict(chunk_iters=64, probes=3, block=1024, use_graph=True, bytes_per_lane=16),
    dict(chunk_iters=1000, probes=0, block=0, use_graph=False),
    dict(chunk_iters=3, probes=3, block=512, use_graph=True, bytes_per_lane=32),
])
def test_batch_tunings_agree(tuning):
    say(f"batch tuning {tuning}")
    sbox = fa.sbox_from_seed(78)
    filters = make_filters(sbox, SPECS[:3], 400)
    texts, off, route = make_texts(filters, 40, 6, max_len=120_000)
    dev = fa.DeviceBuffer.from_host(np.concatenate(texts))
    with fa.Batch([f[0] for f in filters]) as bt:
        bt.set_texts(dev.ptr, off, route)
        bt.run()
        base = as_rows(bt.matches())
        bt.set_tuning(**tuning)
        bt.run()
        assert as_rows(bt.matches()) == base
        t = bt.tuning()
        assert t["block"] in (512, 1024)
        # W=24, pair method: 32 bytes per lane unless 16 was asked for
        assert t["bytes_per_lane"] == (16 if tuning.get("bytes_per_lane") == 16 else 32)
        with pytest.raises(fa.FFBFError) as e:
            bt.set_tuning(bytes_per_lane=8)
        assert e.value.status == fa.EINVAL
    dev.free()
    for f in filters:
        f[0].close()


def test_batch_autotune_and_new_texts():
    say("batch autotune")
    sbox = fa.sbox_from_seed(79)
    filters = make_filters(sbox, SPECS[:2], 500)
    texts, off, route = make_texts(filters, 30, 7)
    dev = fa.DeviceBuffer.from_host(np.concatenate(texts))
    expect, _ = oracle_batch(sbox, filters, texts, route)
    with fa.Batch([f[0] for f in filters]) as bt:
        bt.set_texts(dev.ptr, off, route)
        # the autotuner times its candidates with the run events off too,
        # and leaves them off
        bt.set_timing(False)
        best, ms = bt.autotune()
        assert ms > 0 and best["chunk_iters"] in (32, 64, 128, 256)
        bt.run()
        assert as_rows(bt.matches()) == expect
        assert bt.timings()[0] == 0
        bt.set_timing(True)
        bt.run()
        assert as_rows(bt.matches()) == expect
        assert bt.timings()[0] > 0
        # a different batch on the same object (graph reused)
        texts2, off2, route2 = make_texts(filters, 25, 8)
        dev2 = fa.DeviceBuffer.from_host(np.concatenate(texts2))
        expect2, _ = oracle_batch(sbox, filters, texts2, route2)
        bt.set_texts(dev2.ptr, off2, route2)
        bt.run()
        assert as_rows(bt.matches()) == expect2
        dev2.free()
    dev.free()
    for f in filters:
        f[0].close()


def test_batch_candidate_overflow():
    say("batch overflow")
    sbox = fa.sbox_from_seed(80)
    # a saturated 1 Kibit filter: almost every window is a candidate
    filters = make_filters(sbox, [(2000, 10, fa.ALPHA_DNA_GC40, fa.MODE_LDS), (100, 14, fa.ALPHA_ASCII, fa.MODE_LDS)], 600)
    texts, off, route = make_texts(filters, 12, 9, max_len=150_000)
    dev = fa.DeviceBuffer.from_host(np.concatenate(texts))
    expect, _ = oracle_batch(sbox, filters, texts, route)
    with fa.Batch([f[0] for f in filters]) as bt:
        bt.set_texts(dev.ptr, off, route)
        bt.run()
        assert as_rows(bt.matches()) == expect
    dev.free()
    for f in filters:
        f[0].close()


def test_batch_regrows_descriptor_buffers_between_batches():
    # chunk counts that cross the descriptor buffers' capacity: every new
    # batch must be scanned from its own descriptors (the replayed graph is
    # rebuilt whenever a buffer moves)
    say("batch descriptor regrowth")
    sbox = fa.sbox_from_seed(81)
    filters = make_filters(sbox, SPECS[:2], 700)
    with fa.Batch([f[0] for f in filters]) as bt:
        bt.set_tuning(chunk_iters=32)
        for round_, n_chunks in enumerate([10, 230, 250, 270, 300, 330, 500, 20]):
            # texts of 4..12 chunks of 32 KiB each, up to n_chunks in total
            rng = np.random.default_rng(round_)
            lens, left = [], n_chunks
            while left > 0:
                c = int(min(left, rng.integers(4, 13)))
                lens.append(c * 32 * 1024 - int(rng.integers(0, 100)))
                left -= c
            route = (np.arange(len(lens)) % len(filters)).astype(np.uint32)
            texts = []
            for i, n in enumerate(lens):
                f = int(route[i])
                t = fa.gen_bytes(fa.ALPHA_DNA_GC40 if f % 2 == 0 else fa.ALPHA_ASCII, 50 * round_ + i, n)
                fa.plant(70 * round_ + i, 3, t, filters[f][1], filters[f][2])
                texts.append(t)
            off = np.zeros(len(lens) + 1, dtype=np.uint64)
            off[1:] = np.cumsum(lens)
            dev = fa.DeviceBuffer.from_host(np.concatenate(texts))
            bt.set_texts(dev.ptr, off, route)
            bt.run()
            got = as_rows(bt.matches())
            bt.run_per_text()
            assert got == as_rows(bt.matches()), f"round {round_} ({n_chunks} chunks)"
            assert len(got) >= len(lens)
            dev.free()
    for f in filters:
        f[0].close()


def test_batch_matches_at_unaligned_text_starts():
    # every text starts at a different residue mod 16 in the batch buffer and
    # holds a pattern at its very first offsets, inside the 16-byte unit that
    # straddles the text's start: those candidates must keep their text id
    say("batch matches at unaligned text starts")
    sbox = fa.sbox_from_seed(82)
    filters = make_filters(sbox, SPECS[:3], 800)
    rng = np.random.default_rng(11)
    texts, lens = [], []
    route = (np.arange(48) % len(filters)).astype(np.uint32)
    for i in range(48):
        f = int(route[i])
        _, b, off, _ = filters[f]
        n = 1001 + 16 * int(rng.integers(0, 40))  # n = 9 mod 16: starts walk all residues
        t = fa.gen_bytes(fa.ALPHA_DNA_GC40 if f % 2 == 0 else fa.ALPHA_ASCII, 900 + i, n)
        pid = int(rng.integers(0, off.size - 1))
        p = b[int(off[pid]):int(off[pid + 1])]
        at = i % 12
        t[at:at + p.size] = p
        texts.append(t)
        lens.append(n)
    off = np.zeros(len(lens) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(lens)
    assert len({int(o) % 16 for o in off[:-1]}) == 16
    dev = fa.DeviceBuffer.from_host(np.concatenate(texts))
    expect, _ = oracle_batch(sbox, filters, texts, route)
    assert sum(1 for r in expect if r[2] < 12) >= 48
    with fa.Batch([f[0] for f in filters]) as bt:
        bt.set_texts(dev.ptr, off, route)
        bt.run()
        assert as_rows(bt.matches()) == expect
        bt.run_per_text()
        assert as_rows(bt.matches()) == expect
    dev.free()
    for f in filters:
        f[0].close()


def test_batch_run_time_window_length():
    # a window length without a compiled fast path (the kernels take W at run
    # time): batch equals the oracle text by text
    say("batch with W=40")
    w, k = 40, 3
    sbox = fa.sbox_from_seed(83)
    filters = []
    for i, (n, log2, alpha, mode) in enumerate([(400, 14, fa.ALPHA_DNA_GC40, fa.MODE_LDS),
                                                (4000, 22, fa.ALPHA_ASCII, fa.MODE_FOLD),
                                                (30000, 24, fa.ALPHA_DNA_GC40, fa.MODE_GFOLD)]):
        b, off = fa.gen_patterns(alpha, 850 + i, n, w, 90)
        m = fa.Matcher(w, k, log2, sbox)
        m.set_patterns(b, off)
        m.build()
        m.set_scan_mode(mode)
        filters.append((m, b, off, log2))
    rng = np.random.default_rng(13)
    texts, route = [], []
    for i in range(24):
        f = i % 3
        n = int(rng.integers(0, 150_000))
        t = fa.gen_bytes(fa.ALPHA_DNA_GC40 if f != 1 else fa.ALPHA_ASCII, 870 + i, n)
        if n > 200:
            fa.plant(880 + i, 1 + n // 20000, t, filters[f][1], filters[f][2])
        texts.append(t)
        route.append(f)
    route = np.array(route, dtype=np.uint32)
    off = np.zeros(len(texts) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([t.size for t in texts])
    rows = []
    for i, t in enumerate(texts):
        m, b, po, log2 = filters[int(route[i])]
        r = Oracle(w, k, log2, sbox).run(t, b, po)
        rows += [(i, int(p), int(q)) for p, q in zip(r["ids"], r["offs"])]
    expect = sorted(rows)
    assert len(expect) > 10
    dev = fa.DeviceBuffer.from_host(np.concatenate(texts))
    with fa.Batch([f[0] for f in filters]) as bt:
        bt.set_texts(dev.ptr, off, route)
        bt.run()
        assert as_rows(bt.matches()) == expect
    dev.free()
    for f in filters:
        f[0].close()


def test_batch_many_matches_sorted_on_device():
    # more than a few thousand matches: the batch sorts them on the device
    say("batch with many matches")
    sbox = fa.sbox_from_seed(81)
    specs = [(800, 18, fa.ALPHA_DNA_GC40, fa.MODE_LDS), (1500, 20, fa.ALPHA_ASCII, fa.MODE_FOLD)]
    filters = make_filters(sbox, specs, 500)
    rng = np.random.default_rng(81)
    n_texts = 40
    route = rng.integers(0, len(filters), size=n_texts).astype(np.uint32)
    texts = []
    for i in range(n_texts):
        _, b, off, _ = filters[int(route[i])]
        alpha = fa.ALPHA_DNA_GC40 if int(route[i]) % 2 == 0 else fa.ALPHA_ASCII
        t = fa.gen_bytes(alpha, 8100 + i, 200_000)
        fa.plant(8200 + i, 300, t, b, off)
        texts.append(t)
    expect, _ = oracle_batch(sbox, filters, texts, route)
    assert len(expect) > 8000
    off = np.zeros(n_texts + 1, dtype=np.uint64)
    off[1:] = np.cumsum([t.size for t in texts])
    dev = fa.DeviceBuffer.from_host(np.concatenate(texts))
    with fa.Batch([f[0] for f in filters]) as bt:
        bt.set_texts(dev.ptr, off, route)
        assert bt.run() == len(expect)
        assert as_rows(bt.matches()) == expect
    dev.free()
    for f in filters:
        f[0].close()


def test_batch_follows_changes_to_its_filters():
    """A batch borrows its filters: a new scan mode (which moves a filter to
    the other prefilter class), a new pattern set with a rebuild, or a set
    not yet built must be seen by the next run."""
    say("batch after its filters change")
    sbox = fa.sbox_from_seed(81)
    filters = make_filters(sbox, SPECS[:3], 600)
    texts, off, route = make_texts(filters, 30, 9, max_len=120_000)
    dev = fa.DeviceBuffer.from_host(np.concatenate(texts))
    expect, _ = oracle_batch(sbox, filters, texts, route)
    with fa.Batch([f[0] for f in filters]) as bt:
        bt.set_texts(dev.ptr, off, route)
        bt.run()
        assert as_rows(bt.matches()) == expect
        # LDS filter -> global, global filter -> LDS fold (graph replayed before)
        filters[0][0].set_scan_mode(fa.MODE_GLOBAL)
        filters[2][0].set_scan_mode(fa.MODE_FOLD)
        say("  modes swapped")
        bt.run()
        assert as_rows(bt.matches()) == expect
        bt.run_per_text()
        assert as_rows(bt.matches()) == expect
        # new patterns for filter 1, not yet built: refused
        m1, _, _, log2 = filters[1]
        b2, o2 = fa.gen_patterns(fa.ALPHA_ASCII, 601, 4000, W, 60)
        m1.set_patterns(b2, o2)
        with pytest.raises(fa.FFBFError) as e:
            bt.run()
        assert e.value.status == fa.ESTATE
        m1.build()
        say("  filter 1 rebuilt with a new set")
        filters[1] = (m1, b2, o2, log2)
        expect2, _ = oracle_batch(sbox, filters, texts, route)
        bt.run()
        assert as_rows(bt.matches()) == expect2
    dev.free()
    for f in filters:
        f[0].close()
