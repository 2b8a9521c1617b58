"""HIP path vs the CPU oracle, through the C ABI.  Bit-exact: forward and
reverse filter words, candidate offsets, survivor bitmap and the sorted
(pattern id, offset) match list.  Needs an MI355X (-m gpu)."""
import sys

import numpy as np
import pytest

from conftest import bits_of, load_golden

import ffbf_amd as fa
from oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_all(m, text=None, dev=None):
    if dev is not None:
        m.scan_device(*dev)
    else:
        m.scan(text)
    out = {"cand": m.candidates(), "R": m.filter(fa.FILTER_REVERSE)}
    out["n_surv"] = m.pattern_pass()
    out["surv"] = m.survivors()
    out["n_match"] = m.verify()
    out["matches"] = m.matches()
    return out


def assert_same(g, r, F_gpu=None):
    if F_gpu is not None:
        assert np.array_equal(F_gpu, r["F"]), "forward filter differs"
    assert np.array_equal(g["cand"], r["cand"]), f"candidates differ: {g['cand'].size} vs {r['cand'].size}"
    assert np.array_equal(g["R"], r["R"]), "reverse filter differs"
    assert np.array_equal(g["surv"], r["surv"]), "survivor bitmap differs"
    assert g["n_surv"] == sum(bin(int(x)).count("1") for x in r["surv"])
    assert np.array_equal(g["matches"]["pattern_id"], r["ids"]), "match ids differ"
    assert np.array_equal(g["matches"]["offset"], r["offs"]), "match offsets differ"


def _case_arrays(case):
    pats = [bytes.fromhex(p) for p in case["patterns"]]
    b, off = fa.patterns_from_list(pats)
    text = np.frombuffer(bytes.fromhex(case["text"]), dtype=np.uint8).copy()
    return pats, b, off, text


@pytest.mark.parametrize("mode", [fa.MODE_GLOBAL, fa.MODE_FOLD, fa.MODE_LDS, fa.MODE_GFOLD, fa.MODE_AUTO])
@pytest.mark.parametrize("ci", [0, 1, 2])
def test_golden_cases(ref_sbox, ci, mode):
    case = load_golden("e2e_small.json")["cases"][ci]
    say(f"golden case {ci} mode {mode}")
    pats, b, off, text = _case_arrays(case)
    with fa.Matcher(case["window"], case["k"], case["log2_bits"], ref_sbox) as m:
        m.set_patterns(b, off)
        m.build()
        fold = {fa.MODE_FOLD: case["log2_bits"] - 2, fa.MODE_GFOLD: max(10, case["log2_bits"] - 2)}.get(mode, 0)
        m.set_scan_mode(mode, fold)
        assert bits_of(m.filter(fa.FILTER_FORWARD)) == case["forward_bits"]
        g = gpu_all(m, text)
        assert g["cand"].tolist() == case["candidates"]
        assert bits_of(g["R"]) == case["reverse_bits"]
        assert [i for i in range(len(pats)) if (g["surv"][i >> 5] >> (i & 31)) & 1] == case["survivors"]
        assert [tuple(x) for x in zip(g["matches"]["pattern_id"].tolist(), g["matches"]["offset"].tolist())] == \
            [tuple(x) for x in case["matches"]]


def make_case(alpha, W, n_pat, lmin, lmax, n_text, n_plants, seed):
    b, off = fa.gen_patterns(alpha, seed, n_pat, lmin, lmax)
    text = fa.gen_bytes(alpha, seed + 1, n_text)
    fa.plant(seed + 2, n_plants, text, b, off)
    return b, off, text


RANDOM_CASES = [
    # alpha, W, k, log2, n_pat, lmin, lmax, n_text, plants, mode
    (fa.ALPHA_DNA_GC40, 32, 4, 20, 3000, 32, 100, 1 << 20, 300, fa.MODE_GLOBAL),
    (fa.ALPHA_DNA_GC40, 32, 4, 20, 3000, 32, 100, 1 << 20, 300, fa.MODE_FOLD),
    (fa.ALPHA_DNA_GC40, 32, 4, 18, 3000, 32, 100, 1 << 20, 300, fa.MODE_LDS),
    (fa.ALPHA_ASCII, 19, 4, 16, 2000, 20, 40, 777_777, 200, fa.MODE_AUTO),
    (fa.ALPHA_ASCII, 19, 4, 22, 2000, 20, 40, 777_777, 200, fa.MODE_FOLD),
    (fa.ALPHA_DNA_GC40, 24, 3, 15, 1500, 24, 80, 500_001, 150, fa.MODE_AUTO),
    (fa.ALPHA_DNA_GC40, 24, 3, 15, 1500, 24, 80, 500_001, 150, fa.MODE_GLOBAL),
    (fa.ALPHA_ASCII, 7, 2, 12, 800, 3, 30, 300_003, 100, fa.MODE_GLOBAL),     # generic W, heavy FP
    (fa.ALPHA_ASCII, 7, 2, 12, 800, 3, 30, 300_003, 100, fa.MODE_LDS),
    (fa.ALPHA_DNA_GC40, 33, 5, 20, 1000, 30, 70, 400_000, 100, fa.MODE_FOLD),  # generic W, k=5
    (fa.ALPHA_DNA_GC40, 64, 1, 16, 1000, 60, 90, 400_000, 100, fa.MODE_GLOBAL),  # W=64, k=1
    (fa.ALPHA_ASCII, 1, 4, 10, 20, 1, 3, 50_000, 10, fa.MODE_AUTO),            # W=1
    (fa.ALPHA_DNA_GC40, 19, 5, 26, 5000, 19, 64, 2_000_000, 500, fa.MODE_AUTO),
    (fa.ALPHA_DNA_GC40, 32, 4, 22, 20000, 32, 100, 3_000_000, 500, fa.MODE_GFOLD),
    (fa.ALPHA_ASCII, 24, 3, 20, 5000, 24, 80, 1_000_000, 300, fa.MODE_GFOLD),
    (fa.ALPHA_ASCII, 13, 4, 18, 5000, 13, 40, 1_000_000, 300, fa.MODE_GFOLD),     # generic W
    # LDS fold, then the L2 fold: F beyond L2 (2^26 bits and up)
    (fa.ALPHA_DNA_GC40, 32, 4, 26, 400_000, 32, 100, 2_000_000, 500, fa.MODE_HYBRID),
    (fa.ALPHA_ASCII, 19, 5, 27, 50_000, 19, 64, 1_000_000, 300, fa.MODE_HYBRID),   # k=5
    (fa.ALPHA_ASCII, 13, 3, 26, 20_000, 13, 40, 1_000_000, 300, fa.MODE_HYBRID),   # generic W, k=3
    (fa.ALPHA_DNA_GC40, 24, 2, 26, 300_000, 24, 60, 1_000_000, 300, fa.MODE_HYBRID),  # k=2
]


@pytest.mark.parametrize("ci", range(len(RANDOM_CASES)))
def test_random_vs_oracle(ci):
    alpha, W, k, log2, n_pat, lmin, lmax, n_text, plants, mode = RANDOM_CASES[ci]
    say(f"random case {ci}: W={W} k={k} log2={log2} n_pat={n_pat} n_text={n_text} mode={mode}")
    sbox = fa.sbox_from_seed(1000 + ci)
    b, off, text = make_case(alpha, W, n_pat, lmin, lmax, n_text, plants, 77 + ci)
    o = Oracle(W, k, log2, sbox)
    r = o.run(text, b, off, threads=8)
    with fa.Matcher(W, k, log2, sbox) as m:
        m.set_patterns(b, off)
        m.build()
        if mode == fa.MODE_FOLD:
            m.set_scan_mode(mode, max(7, log2 - 3))
        elif mode == fa.MODE_GFOLD:
            m.set_scan_mode(mode, max(10, log2 - 4))
        else:
            m.set_scan_mode(mode)
        g = gpu_all(m, text)
        assert_same(g, r, m.filter(fa.FILTER_FORWARD))
    assert r["ids"].size >= min(plants, 1)


@pytest.mark.parametrize("shift", [0, 1, 3, 7, 8, 13, 15])
def test_unaligned_device_text(shift):
    say(f"unaligned shift {shift}")
    sbox = fa.sbox_from_seed(5)
    b, off, text = make_case(fa.ALPHA_DNA_GC40, 32, 500, 32, 60, 100_003, 60, 9)
    o = Oracle(32, 4, 18, sbox)
    r = o.run(text, b, off)
    buf = fa.DeviceBuffer(text.size + 32)
    buf.upload(text, offset=shift)
    with fa.Matcher(32, 4, 18, sbox) as m:
        m.set_patterns(b, off)
        m.build()
        for mode in (fa.MODE_GLOBAL, fa.MODE_LDS):
            m.set_scan_mode(mode)
            g = gpu_all(m, dev=(buf.ptr + shift, text.size))
            assert_same(g, r)
    buf.free()


def test_timing_off_same_results():
    """Stage events off: the same F, candidates, R, survivors and matches;
    the stage times read 0 until they are switched back on."""
    sbox = fa.sbox_from_seed(7)
    W = 24
    b, off = fa.gen_patterns(fa.ALPHA_DNA_GC40, 8, 3000, 24, 60)
    text = fa.gen_bytes(fa.ALPHA_DNA_GC40, 9, 3 << 20)
    fa.plant(10, 40, text, b, off)
    r = Oracle(W, 4, 20, sbox).run(text, b, off)
    with fa.Matcher(W, 4, 20, sbox) as m:
        m.set_timing(False)
        m.set_patterns(b, off)
        m.build()
        assert_same(gpu_all(m, text), r, m.filter(fa.FILTER_FORWARD))
        t = m.timings()
        assert t["build_ms"] == 0 and t["scan_kernel_ms"] == 0 and t["verify_ms"] == 0
        assert t["scan_mode"] in (fa.MODE_LDS, fa.MODE_FOLD, fa.MODE_GLOBAL)
        m.set_timing(True)
        m.build()
        assert_same(gpu_all(m, text), r, m.filter(fa.FILTER_FORWARD))
        t = m.timings()
        assert t["build_ms"] > 0 and t["scan_kernel_ms"] > 0 and t["verify_ms"] > 0


def test_edge_sizes():
    sbox = fa.sbox_from_seed(6)
    W = 19
    b, off = fa.gen_patterns(fa.ALPHA_ASCII, 3, 50, 19, 30)
    o = Oracle(W, 4, 14, sbox)
    with fa.Matcher(W, 4, 14, sbox) as m:
        m.set_patterns(b, off)
        m.build()
        for n in (0, 1, W - 1, W, W + 1, 16, 17, 31, 32, 33, 1023, 1024, 1025, 4097):
            say(f"edge size {n}")
            text = fa.gen_bytes(fa.ALPHA_ASCII, 4, n)
            if n >= 30:
                text[n - 30:n - 30 + int(off[1] - off[0])] = b[int(off[0]):int(off[1])]
            r = o.run(text, b, off)
            g = gpu_all(m, text)
            assert_same(g, r)


def test_no_patterns_and_short_patterns():
    sbox = fa.sbox_from_seed(7)
    text = fa.gen_bytes(fa.ALPHA_DNA_GC40, 1, 10000)
    with fa.Matcher(24, 4, 16, sbox) as m:
        say("empty pattern set")
        m.set_pattern_list([])
        m.build()
        assert not m.filter().any()
        g = gpu_all(m, text)
        assert g["cand"].size == 0 and g["n_match"] == 0 and g["n_surv"] == 0
        say("only short patterns")
        m.set_pattern_list([b"ACGT", b"", b"A" * 23])
        m.build()
        assert not m.filter().any()
        g = gpu_all(m, text)
        assert g["cand"].size == 0 and g["n_match"] == 0


def test_candidate_overflow_grows_and_matches():
    # a 1 Kibit filter from 3000 patterns is saturated: nearly every offset is
    # a candidate, far beyond the initial candidate capacity
    say("candidate overflow")
    sbox = fa.sbox_from_seed(8)
    b, off, text = make_case(fa.ALPHA_DNA_GC40, 12, 3000, 12, 20, 600_000, 400, 21)
    o = Oracle(12, 2, 10, sbox)
    r = o.run(text, b, off, threads=8)
    assert r["cand"].size > 200_000
    with fa.Matcher(12, 2, 10, sbox) as m:
        m.set_patterns(b, off)
        m.build()
        for mode in (fa.MODE_GLOBAL, fa.MODE_LDS):
            m.set_scan_mode(mode)
            assert_same(gpu_all(m, text), r)
            # fused path: one sync, same answer
            buf = fa.DeviceBuffer.from_host(text)
            assert m.run_device(buf.ptr, text.size) == r["ids"].size
            mm = m.matches()
            assert np.array_equal(mm["pattern_id"], r["ids"]) and np.array_equal(mm["offset"], r["offs"])
            buf.free()


def test_match_text_and_lines_interface(ref_sbox):
    say("lines interface")
    pats = [b"GATTACAGATTACAGATTACA", b"short", b"TTTTTTTTTTTTTTTTTTTTTTTT", b"GATTACAGATTACAGATTACAGG"]
    text = b"xx" + pats[0] + b"GG" + b"TTTTTTTTTTTTTTTTTTTTTTTTTT" + b"yy"
    with fa.Matcher(19, 4, 20, ref_sbox) as m:
        m.set_patterns_lines(b"\n".join(pats) + b"\n")
        m.build()
        res = m.match(text)
        expect = sorted([(0, 2), (3, 2), (2, 25), (2, 26), (2, 27)])
        assert res.pairs == expect


def test_filter_persistence_round_trip():
    say("filter save/load")
    from oracle import filter_save_deltas

    sbox = fa.sbox_from_seed(9)
    b, off, text = make_case(fa.ALPHA_DNA_GC40, 24, 2000, 24, 60, 300_000, 100, 31)
    o = Oracle(24, 3, 18, sbox)
    r = o.run(text, b, off)
    with fa.Matcher(24, 3, 18, sbox) as m:
        m.set_patterns(b, off)
        m.build()
        d = m.filter_save()
        assert np.array_equal(d, filter_save_deltas(r["F"], 18))
    with fa.Matcher(24, 3, 18, sbox) as m2:
        m2.set_patterns(b, off)
        m2.filter_load(d)
        assert_same(gpu_all(m2, text), r, m2.filter())


SCAN_CONFIGS = [
    # mode, fold log2 bits, probes, block -> effective block
    (fa.MODE_GLOBAL, 0, 0, 0, 256), (fa.MODE_GLOBAL, 0, 0, 512, 256),
    (fa.MODE_FOLD, 14, 0, 0, 1024), (fa.MODE_FOLD, 17, 1, 512, 512), (fa.MODE_FOLD, 17, 2, 1024, 1024),
    (fa.MODE_FOLD, 20, 4, 512, 512), (fa.MODE_FOLD, 18, 3, 0, 1024),
    (fa.MODE_LDS, 0, 1, 512, 512), (fa.MODE_LDS, 0, 3, 1024, 1024), (fa.MODE_LDS, 0, 4, 0, 1024),
    (fa.MODE_LDS, 0, 2, 0, 1024),
    (fa.MODE_GFOLD, 12, 0, 0, 256), (fa.MODE_GFOLD, 16, 0, 512, 256), (fa.MODE_GFOLD, 20, 0, 0, 256),
]


@pytest.mark.parametrize("W", [19, 24, 32, 21])
def test_scan_configs_agree_on_one_input(W):
    say(f"config agreement W={W}")
    sbox = fa.sbox_from_seed(10 + W)
    b, off, text = make_case(fa.ALPHA_ASCII, W, 4000, W, 64, 3_000_011, 300, 41)
    o = Oracle(W, 4, 20, sbox)
    r = o.run(text, b, off, threads=8)
    with fa.Matcher(W, 4, 20, sbox) as m:
        m.set_patterns(b, off)
        m.build()
        for mode, fold, probes, block, eff in SCAN_CONFIGS:
            say(f"  mode={mode} fold={fold} probes={probes} block={block}")
            m.set_scan_mode(mode, fold)
            m.set_scan_tuning(probes, block)
            assert_same(gpu_all(m, text), r)
            t = m.timings()
            assert t["scan_mode"] == mode
            assert t["block"] == eff


def test_gather_bench_runs():
    ms, rate = fa.gather_bench(20, steps=4)
    assert ms > 0 and rate > 1e9
    for kind in range(fa.PROBE_UPLOAD + 1):
        ms, rate = fa.probe_bench(kind, 20, steps=4)
        assert ms > 0 and rate > 1e9, fa.PROBE_KIND_NAMES[kind]
    for bad in ((fa.PROBE_UPLOAD, 15), (fa.PROBE_UPLOAD, 32), (fa.PROBE_UPLOAD + 1, 20)):
        with pytest.raises(fa.FFBFError) as err:
            fa.probe_bench(bad[0], bad[1], steps=1)
        assert err.value.status == fa.EINVAL


@pytest.mark.parametrize("mode", [fa.MODE_LDS, fa.MODE_GLOBAL])
def test_periodic_text_dense_identical_candidates(mode):
    # tandem repeats: runs of identical windows that are all candidates,
    # drained side by side (the feed-forward skips a bit that a nearby lane
    # of the same drain sets; R must still hold every bit)
    say(f"periodic text, mode {mode}")
    sbox = fa.sbox_from_seed(61)
    W = 19
    rng = np.random.default_rng(61)
    units = [b"A", b"AC", b"ACG", b"ACGT", b"AACGT", b"AACCGT"]
    parts, pats = [], []
    for i in range(300):
        u = units[i % len(units)]
        run = (u * (600 // len(u) + 1))[: int(rng.integers(100, 600))]
        parts.append(run)
        parts.append(fa.gen_bytes(fa.ALPHA_DNA_GC40, 6100 + i, int(rng.integers(50, 400))).tobytes())
        if i < 40:
            off0 = int(rng.integers(0, len(u)))
            pats.append((u * 40)[off0:off0 + W + int(rng.integers(0, 20))])
    text = np.frombuffer(b"".join(parts), dtype=np.uint8).copy()
    b0, o0 = fa.gen_patterns(fa.ALPHA_DNA_GC40, 6200, 3000, W, 50)
    pats += [bytes(b0[int(o0[i]):int(o0[i + 1])]) for i in range(o0.size - 1)]
    b, off = fa.patterns_from_list(pats)
    for log2 in (14, 20):  # dense false positives, then mostly true repeats
        r = Oracle(W, 4, log2, sbox).run(text, b, off)
        with fa.Matcher(W, 4, log2, sbox) as m:
            m.set_patterns(b, off)
            m.build()
            m.set_scan_mode(mode if (mode != fa.MODE_LDS or log2 <= 20) else fa.MODE_AUTO)
            assert_same(gpu_all(m, text), r, m.filter(fa.FILTER_FORWARD))
        assert r["cand"].size > 5000


def test_new_pattern_set_needs_a_build():
    say("pattern set replaced after a build")
    sbox = fa.sbox_from_seed(62)
    b1, o1, text = make_case(fa.ALPHA_ASCII, 19, 500, 19, 40, 200_000, 50, 63)
    b2, o2 = fa.gen_patterns(fa.ALPHA_ASCII, 64, 700, 19, 40)
    fa.plant(65, 40, text, b2, o2)
    with fa.Matcher(19, 4, 18, sbox) as m:
        m.set_patterns(b1, o1)
        m.build()
        m.scan(text)
        m.set_patterns(b2, o2)
        with pytest.raises(fa.FFBFError) as e:
            m.scan(text)
        assert e.value.status == fa.ESTATE
        m.build()
        r = Oracle(19, 4, 18, sbox).run(text, b2, o2)
        assert_same(gpu_all(m, text), r, m.filter())


@pytest.mark.parametrize("W", [19, 21])
def test_high_pass_fold_and_small_lds_filter(W):
    """A fold so small that nearly every window passes its up-front bits
    (F's first bit is then tested before queueing), and a 2^13-bit filter
    held whole in LDS with dense candidates (R gathered in LDS): both equal
    the oracle, alone and in one batch."""
    say(f"high-pass fold / small LDS filter W={W}")
    sbox = fa.sbox_from_seed(90 + W)
    b1, o1, t1 = make_case(fa.ALPHA_ASCII, W, 3000, W, 50, 400_000, 200, 91)
    b2, o2, t2 = make_case(fa.ALPHA_DNA_GC40, W, 600, W, 40, 400_000, 200, 92)
    r1 = Oracle(W, 3, 20, sbox).run(t1, b1, o1, threads=8)
    r2 = Oracle(W, 3, 13, sbox).run(t2, b2, o2, threads=8)
    assert r2["cand"].size > 1000  # dense candidates on the small filter
    m1 = fa.Matcher(W, 3, 20, sbox)
    m1.set_patterns(b1, o1)
    m1.build()
    m1.set_scan_mode(fa.MODE_FOLD, 10)  # 1 Kibit fold: nearly full
    m2 = fa.Matcher(W, 3, 13, sbox)
    m2.set_patterns(b2, o2)
    m2.build()
    m2.set_scan_mode(fa.MODE_LDS)
    for probes in (1, 2, 3):
        m1.set_scan_tuning(probes, 0)
        assert_same(gpu_all(m1, t1), r1)
    assert_same(gpu_all(m2, t2), r2)
    texts = [t1, t2]
    off = np.zeros(3, dtype=np.uint64)
    off[1] = t1.size
    off[2] = t1.size + t2.size
    dev = fa.DeviceBuffer.from_host(np.concatenate(texts))
    expect = sorted([(0, int(p), int(q)) for p, q in zip(r1["ids"], r1["offs"])] +
                    [(1, int(p), int(q)) for p, q in zip(r2["ids"], r2["offs"])])
    with fa.Batch([m1, m2]) as bt:
        bt.set_texts(dev.ptr, off, np.array([0, 1], dtype=np.uint32))
        for tuning in (dict(), dict(probes=1, bytes_per_lane=16), dict(probes=3, block=512)):
            bt.set_tuning(**tuning)
            bt.run()
            got = bt.matches()
            rows = list(zip(got["text_id"].tolist(), got["pattern_id"].tolist(), got["offset"].tolist()))
            assert rows == expect, tuning
            assert np.array_equal(m1.filter(fa.FILTER_REVERSE), r1["R"])
            assert np.array_equal(m2.filter(fa.FILTER_REVERSE), r2["R"])
    dev.free()
    m1.close()
    m2.close()
