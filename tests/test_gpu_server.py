"""Serving loop: every request's matches in the mix (continuous batching,
two slots overlapping on the device) equal its matches scanned alone, and
admission by HBM footprint rejects what can never fit.  Needs an MI355X."""
import sys
import threading

import numpy as np
import pytest

import ffbf_amd as fa
from oracle import Oracle

pytestmark = pytest.mark.gpu

W, K = 24, 3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_filters(sbox):
    # whole filter in LDS, LDS fold, F in L2, L2-sized global fold: each mode
    # is copied into the server's replicas differently
    specs = [(800, 14, fa.ALPHA_DNA_GC40, fa.MODE_LDS), (5000, 22, fa.ALPHA_ASCII, fa.MODE_FOLD),
             (40000, 20, fa.ALPHA_DNA_GC40, fa.MODE_GLOBAL), (200000, 22, fa.ALPHA_DNA_GC40, fa.MODE_AUTO),
             (1_000_000, 26, fa.ALPHA_ASCII, fa.MODE_GFOLD)]
    out = []
    for i, (n, log2, alpha, mode) in enumerate(specs):
        b, off = fa.gen_patterns(alpha, 900 + i, n, W, 70)
        m = fa.Matcher(W, K, log2, sbox)
        m.set_patterns(b, off)
        m.build()
        m.set_scan_mode(mode)
        assert mode == fa.MODE_AUTO or m.timings()["scan_mode"] == mode
        out.append((m, b, off, log2, alpha))
    return out


def make_requests(filters, n, seed, max_len):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        f = int(rng.integers(0, len(filters)))
        size = int(rng.integers(0, max_len)) if i % 7 else int(rng.integers(0, 64))
        _, b, off, _, alpha = filters[f]
        t = fa.gen_bytes(alpha, seed * 100 + i, size)
        if size > 200:
            fa.plant(seed * 31 + i, 1 + size // 30000, t, b, off)
        reqs.append((f, t))
    return reqs


def alone(filters, f, t):
    return filters[f][0].match(t).matches


def test_server_results_equal_alone():
    say("server mix vs alone")
    sbox = fa.sbox_from_seed(91)
    filters = make_filters(sbox)
    say(f"  filter modes {[f[0].timings()['scan_mode'] for f in filters]}")
    reqs = make_requests(filters, 80, 3, 300_000)
    expect = [alone(filters, f, t) for f, t in reqs]
    # spot-check the reference answer itself against the oracle
    for i in (1, 2, 3):
        f, t = reqs[i]
        _, b, off, log2, _ = filters[f]
        r = Oracle(W, K, log2, sbox).run(t, b, off)
        assert np.array_equal(expect[i]["pattern_id"], r["ids"]) and np.array_equal(expect[i]["offset"], r["offs"])
    assert sum(e.size for e in expect) > 20
    with fa.Server([f[0] for f in filters], hbm_budget_bytes=64 << 20, slots=2) as srv:
        # all at once
        ids = [srv.submit(f, t) for f, t in reqs]
        for i, rid in enumerate(ids):
            got, lat = srv.wait(rid)
            assert lat >= 0
            assert np.array_equal(got["pattern_id"], expect[i]["pattern_id"]), i
            assert np.array_equal(got["offset"], expect[i]["offset"]), i
        # staggered, from two submitting threads, waits in another order
        ids = [None] * len(reqs)

        def submit_range(lo, hi):
            for j in range(lo, hi):
                ids[j] = srv.submit(*reqs[j])

        th = [threading.Thread(target=submit_range, args=(0, 40)), threading.Thread(target=submit_range, args=(40, 80))]
        for x in th:
            x.start()
        for x in th:
            x.join()
        for i in reversed(range(len(reqs))):
            got, _ = srv.wait(ids[i])
            assert np.array_equal(got["pattern_id"], expect[i]["pattern_id"]), i
            assert np.array_equal(got["offset"], expect[i]["offset"]), i
        st = srv.stats()
        assert st["completed"] == 2 * len(reqs) and st["pending"] == 0 and st["in_flight_bytes"] == 0
        assert st["batches"] < st["completed"]  # requests were batched together
    for f in filters:
        f[0].close()


def test_server_pinned_requests():
    say("server pinned requests")
    sbox = fa.sbox_from_seed(93)
    filters = make_filters(sbox)
    reqs = make_requests(filters, 40, 4, 400_000)
    pins = []
    for f, t in reqs:
        p = fa.PinnedBuffer(t.size)
        p.array[:] = t
        pins.append((f, p))
    expect = [alone(filters, f, t) for f, t in reqs]
    with fa.Server([f[0] for f in filters], hbm_budget_bytes=32 << 20, slots=2) as srv:
        ids = [srv.submit(f, p.array) for f, p in pins]
        for i, rid in enumerate(ids):
            got, _ = srv.wait(rid)
            assert np.array_equal(got["pattern_id"], expect[i]["pattern_id"]), i
            assert np.array_equal(got["offset"], expect[i]["offset"]), i
    for _, p in pins:
        p.free()
    for f in filters:
        f[0].close()


def dense_text(b, off, seed, size):
    """Patterns of the set back to back: every byte lies in a planted
    occurrence, so one wrong byte anywhere changes the match list."""
    rng = np.random.default_rng(seed)
    lens = np.diff(off)
    parts, total = [], 0
    while total < size:
        j = int(rng.integers(0, lens.size))
        parts.append(b[int(off[j]):int(off[j + 1])])
        total += int(lens[j])
    return np.concatenate(parts)[:size] if parts else np.zeros(0, dtype=np.uint8)


def test_server_upload_kernel():
    """Text in pinned host memory uploaded by the server's copy kernel: every
    source alignment mod 16, lengths around the kernel's 16-byte chunks and
    63-chunk segments, a request cut into pieces, batches packing pieces at
    odd device offsets.  Matches equal each request scanned alone; pageable
    text still goes by the DMA engine."""
    say("server upload kernel")
    sbox = fa.sbox_from_seed(94)
    filters = make_filters(sbox)[:2]
    sizes = [1, 2, 15, 16, 17, 31, 33, 1007, 1008, 1009, 1024, 4099, 65543, 300_001, 3 << 20]
    sizes += [int(x) for x in np.random.default_rng(5).integers(1, 200_000, size=33)]
    pin = fa.PinnedBuffer(sum(sizes) + 32 * len(sizes) + 64)
    reqs, at = [], 0
    for i, n in enumerate(sizes):
        at = (at + 15) // 16 * 16 + i % 16  # source addresses at every alignment mod 16
        f = i % 2
        _, b, off, _, _ = filters[f]
        t = dense_text(b, off, 40 + i, n)
        pin.array[at:at + n] = t
        reqs.append((f, at, n))
        at += n
    expect = [alone(filters, f, pin.array[a:a + n].copy()) for f, a, n in reqs]
    assert sum(e.size for e in expect) > 10_000
    with fa.Server([f[0] for f in filters], hbm_budget_bytes=16 << 20, slots=2, max_batch_bytes=2 << 20,
                   upload=fa.UPLOAD_KERNEL) as srv:
        for rep in range(2):
            if rep == 1:
                # the host rewrites the same pinned bytes with other text: the
                # second uploads must read the new bytes, not lines the first
                # left in a cache
                for i, (f, a, n) in enumerate(reqs):
                    _, b, off, _, _ = filters[f]
                    pin.array[a:a + n] = dense_text(b, off, 600 + i, n)
                expect = [alone(filters, f, pin.array[a:a + n].copy()) for f, a, n in reqs]
            ids = [srv.submit(f, pin.array[a:a + n]) for f, a, n in reqs]
            for i, rid in enumerate(ids):
                got, _ = srv.wait(rid)
                assert np.array_equal(got["pattern_id"], expect[i]["pattern_id"]), (rep, i, reqs[i])
                assert np.array_equal(got["offset"], expect[i]["offset"]), (rep, i, reqs[i])
        up = srv.stats()["kernel_upload_bytes"]
        assert up >= 2 * sum(sizes)  # every pinned byte (pieces of the long one overlap)
        # pageable copies of the same requests: the DMA engine, same matches
        ids = [srv.submit(f, pin.array[a:a + n].copy()) for f, a, n in reqs[:12]]
        for i, rid in enumerate(ids):
            got, _ = srv.wait(rid)
            assert np.array_equal(got["offset"], expect[i]["offset"]), i
        assert srv.stats()["kernel_upload_bytes"] == up
    # 300 small pinned requests queued behind eight pageable 2 MiB ones on a
    # one-slot server: they wait together and go in a few batches of more
    # than 64 pieces each, so several upload launches per batch
    small = []
    tiny = fa.PinnedBuffer(300 * 4200)
    at = 0
    for i in range(300):
        n = 1 + (i * 37) % 4000
        at = (at + 15) // 16 * 16 + (i * 7) % 16
        f = i % 2
        _, b, off, _, _ = filters[f]
        tiny.array[at:at + n] = dense_text(b, off, 900 + i, n)
        small.append((f, at, n))
        at += n
    want = [alone(filters, f, tiny.array[a:a + n].copy()) for f, a, n in small]
    blockers = [(i % 2, fa.gen_bytes(fa.ALPHA_ASCII, 950 + i, 2 << 20)) for i in range(8)]
    with fa.Server([f[0] for f in filters], hbm_budget_bytes=16 << 20, slots=1, max_batch_bytes=2 << 20,
                   upload=fa.UPLOAD_KERNEL) as srv:
        first = srv.submit_many(blockers)
        ids = srv.submit_many([(f, tiny.array[a:a + n]) for f, a, n in small])
        for rid in first:
            srv.wait(rid)
        b0 = srv.stats()["batches"]
        for i, rid in enumerate(ids):
            got, _ = srv.wait(rid)
            assert np.array_equal(got["pattern_id"], want[i]["pattern_id"]), (i, small[i])
            assert np.array_equal(got["offset"], want[i]["offset"]), (i, small[i])
        tiny_batches = srv.stats()["batches"] - b0
        say(f"  300 small requests in {tiny_batches} batches")
        assert tiny_batches <= 4  # so some batch held at least 75 pieces
        assert srv.stats()["kernel_upload_bytes"] == sum(n for _, _, n in small)
    tiny.free()
    with fa.Server([f[0] for f in filters], hbm_budget_bytes=16 << 20, slots=1, max_batch_bytes=2 << 20) as srv:
        got, _ = srv.wait(srv.submit(reqs[13][0], pin.array[reqs[13][1]:reqs[13][1] + reqs[13][2]]))
        assert np.array_equal(got["offset"], expect[13]["offset"])
        assert srv.stats()["kernel_upload_bytes"] == 0  # the default copies
    with pytest.raises(fa.FFBFError) as e:
        fa.Server([filters[0][0]], upload=7)
    assert e.value.status == fa.EINVAL
    pin.free()
    for f in filters:
        f[0].close()


def test_server_adjacent_requests_merged_copies():
    """Requests back to back in one buffer (pinned and pageable), as a
    receive ring holds them: the DMA copies of adjacent pieces in a batch are
    merged.  A request longer than a slot sits among them, its overlapping
    pieces never adjacent to each other.  Matches equal each alone."""
    say("server adjacent requests")
    sbox = fa.sbox_from_seed(95)
    filters = make_filters(sbox)[:2]
    rng = np.random.default_rng(95)
    sizes = [int(x) for x in rng.integers(1, 150_000, size=60)]
    sizes[30] = 3 << 20  # pieces of a 2 MiB slot
    total = sum(sizes)
    pin = fa.PinnedBuffer(total)
    reqs, at = [], 0
    for i, n in enumerate(sizes):
        f = i % 2
        _, b, off, _, _ = filters[f]
        pin.array[at:at + n] = dense_text(b, off, 300 + i, n)
        reqs.append((f, at, n))
        at += n
    page = pin.array.copy()  # the same ring in pageable memory
    expect = [alone(filters, f, pin.array[a:a + n].copy()) for f, a, n in reqs]
    assert sum(e.size for e in expect) > 10_000
    for buf in (pin.array, page):
        with fa.Server([f[0] for f in filters], hbm_budget_bytes=16 << 20, slots=2, max_batch_bytes=2 << 20) as srv:
            ids = srv.submit_many([(f, buf[a:a + n]) for f, a, n in reqs])
            for i, rid in enumerate(ids):
                got, _ = srv.wait(rid)
                assert np.array_equal(got["pattern_id"], expect[i]["pattern_id"]), (i, reqs[i])
                assert np.array_equal(got["offset"], expect[i]["offset"]), (i, reqs[i])
            assert srv.stats()["batches"] < len(reqs)
    pin.free()
    for f in filters:
        f[0].close()


def test_server_admission_limits():
    say("server admission")
    sbox = fa.sbox_from_seed(92)
    filters = make_filters(sbox)[:2]
    # a slot's buffer (with its candidate reserve) larger than the budget:
    # requests needing a whole slot can never be admitted
    with fa.Server([f[0] for f in filters], hbm_budget_bytes=1 << 20, slots=1, max_batch_bytes=2 << 20) as tight:
        with pytest.raises(fa.FFBFError) as e:
            tight.submit(0, np.zeros(3 << 20, dtype=np.uint8))
        assert e.value.status == fa.ECAPACITY
        assert tight.stats()["rejected"] == 1
        small = make_requests(filters, 1, 6, 100_000)[0]
        got, _ = tight.wait(tight.submit(*small))
        assert np.array_equal(got["offset"], alone(filters, *small)["offset"])
    with fa.Server([f[0] for f in filters], hbm_budget_bytes=4 << 20, slots=2, max_batch_requests=3) as srv:
        # larger than one slot (2 MiB): scanned in two overlapping pieces
        big = fa.gen_bytes(filters[0][4], 77, 3 << 20)
        fa.plant(78, 300, big, filters[0][1], filters[0][2])
        # plants around the piece border: the second piece owns the window
        # starts from 2 MiB - (longest pattern - 1) on; one plant starts just
        # below that and ends at the first piece's last byte
        b, off = filters[0][1], filters[0][2]
        lens = np.diff(off)
        border = (2 << 20) - (int(lens.max()) - 1)
        at_list = [border - 150, border - 75, border - 1, border + 72, border + 150]
        longest = int(np.argmax(lens))
        for j, at in zip([3, 4, longest, 5, 6], at_list):
            p = b[int(off[j]):int(off[j + 1])]
            big[at:at + p.size] = p
        expect_big = alone(filters, 0, big)
        got, _ = srv.wait(srv.submit(0, big))
        for at in at_list:
            assert at in set(expect_big["offset"].tolist()), at
        assert np.array_equal(got["pattern_id"], expect_big["pattern_id"])
        assert np.array_equal(got["offset"], expect_big["offset"])
        with pytest.raises(fa.FFBFError):
            srv.submit(5, b"x" * 100)  # no such filter
        # many requests that only fit a few at a time: all complete, in budget
        reqs = make_requests(filters, 30, 5, 900_000)
        expect = [alone(filters, f, t) for f, t in reqs]
        ids = [srv.submit(f, t) for f, t in reqs]
        for i, rid in enumerate(ids):
            got, _ = srv.wait(rid)
            assert np.array_equal(got["offset"], expect[i]["offset"]), i
        assert srv.stats()["rejected"] == 0
        # two waits on one request: exactly one gets the matches, the other
        # an error (already waited for, or already released)
        f0, t0 = reqs[0]
        rid = srv.submit(f0, t0)
        res = []

        def waiter():
            try:
                res.append(srv.wait(rid)[0])
            except fa.FFBFError as err:
                res.append(err.status)

        ths = [threading.Thread(target=waiter) for _ in range(2)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        errs = [x for x in res if isinstance(x, int)]
        oks = [x for x in res if not isinstance(x, int)]
        assert errs == [fa.EINVAL] and len(oks) == 1
        got = oks[0]
        assert np.array_equal(got["offset"], expect[0]["offset"])
    for f in filters:
        f[0].close()


def test_server_rejects_unbuilt_filter():
    # the slots copy each filter's F as built: a source with patterns but no
    # built F is refused at creation
    say("server unbuilt filter")
    sbox = fa.sbox_from_seed(94)
    b, off = fa.gen_patterns(fa.ALPHA_DNA_GC40, 95, 200, W, 40)
    m = fa.Matcher(W, K, 16, sbox)
    m.set_patterns(b, off)
    with pytest.raises(fa.FFBFError) as e:
        fa.Server([m], hbm_budget_bytes=8 << 20, slots=1)
    assert e.value.status == fa.ESTATE
    m.build()
    with fa.Server([m], hbm_budget_bytes=8 << 20, slots=1) as srv:
        t = fa.gen_bytes(fa.ALPHA_DNA_GC40, 96, 100_000)
        fa.plant(97, 20, t, b, off)
        got, _ = srv.wait(srv.submit(0, t))
        want = m.match(t).matches
        assert np.array_equal(got["pattern_id"], want["pattern_id"])
        assert np.array_equal(got["offset"], want["offset"])
    m.close()
