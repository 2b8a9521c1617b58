"""The CPU binary's split filter (ffbf_create_split) on the GPU against the
oracle's restatement of src/rabin-karp/ffbf.cpp, through the C ABI.
Bit-exact: forward and reverse filter words (2^c + 2^e bits), candidate
offsets (windows without a newline), survivor bitmap and the sorted
(pattern id, offset) match list.  Needs an MI355X (-m gpu)."""
import sys

import numpy as np
import pytest

from conftest import bits_of, load_golden

import ffbf_amd as fa
from oracle import SplitOracle, candidate_lines, filter_save_deltas_bits, survivor_files

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
    assert np.array_equal(g["matches"]["pattern_id"], r["ids"]), "match ids differ"
    assert np.array_equal(g["matches"]["offset"], r["offs"]), "match offsets differ"


def sbox64(seed):
    return np.random.default_rng(seed).integers(0, 2**64, size=256, dtype=np.uint64)


def lines_text(seed, n, b, off, line_mean=60, high_frac=0.01, plants=0):
    """ASCII text cut into lines of random length, with some bytes >= 0x80,
    and pattern copies planted over it (a plant may span a line break)."""
    rng = np.random.default_rng(seed)
    t = fa.gen_bytes(fa.ALPHA_ASCII, seed, n)
    if n:
        t[rng.integers(0, n, size=max(1, int(n * high_frac)))] = rng.integers(0x80, 0x100, size=max(1, int(n * high_frac)))
        t[rng.integers(0, n, size=max(1, n // line_mean))] = 10
    if plants and n > 200:
        fa.plant(seed + 3, plants, t, b, off)
    return t


def make_split(W, k, s, c, e, sb, b, off):
    m = fa.Matcher.split(W, k, s, c, e, sb)
    m.set_patterns(b, off)
    m.build()
    return m


def _case_arrays(case):
    pats = [bytes.fromhex(p) for p in case["patterns"]]
    b, off = fa.patterns_from_list(pats)
    text = np.frombuffer(bytes.fromhex(case["text"]), dtype=np.uint8).copy()
    return pats, b, off, text


@pytest.mark.parametrize("mode", [fa.MODE_AUTO, fa.MODE_GLOBAL, fa.MODE_FOLD])
@pytest.mark.parametrize("ci", [0, 1, 2])
def test_split_golden_cases(ref_sbox64, ci, mode):
    case = load_golden("split_e2e_small.json")["cases"][ci]
    say(f"split golden case {ci} mode {mode}")
    pats, b, off, text = _case_arrays(case)
    with make_split(case["window"], case["k"], case["s"], case["log2_cache_bits"], case["log2_ext_bits"],
                    ref_sbox64, b, off) as m:
        if mode == fa.MODE_FOLD and case["s"] == 0:
            with pytest.raises(fa.FFBFError) as e:
                m.set_scan_mode(mode)
            assert e.value.status == fa.EINVAL
            return
        m.set_scan_mode(mode, max(7, case["log2_cache_bits"] - 2) if mode == fa.MODE_FOLD else 0)
        assert bits_of(m.filter(fa.FILTER_FORWARD)) == case["forward_bits"]
        g = gpu_all(m, text)
        assert g["cand"].tolist() == case["candidates"]
        assert bits_of(g["R"]) == case["reverse_bits"]
        assert [i for i in range(len(pats)) if (g["surv"][i >> 5] >> (i & 31)) & 1] == case["survivors"]
        assert [tuple(x) for x in zip(g["matches"]["pattern_id"].tolist(), g["matches"]["offset"].tolist())] == \
            [tuple(x) for x in case["matches"]]


# W, k, s, log2 cache, log2 ext, patterns, text bytes, plants, mode, fold log2, probes
RANDOM_CASES = [
    (19, 5, 2, 24, 28, 3000, 4_000_037, 300, fa.MODE_AUTO, 0, 0),   # the CPU binary's compile-time shape
    (19, 5, 2, 24, 28, 3000, 4_000_037, 300, fa.MODE_GLOBAL, 0, 0),
    (19, 11, 3, 20, 22, 5000, 3_000_001, 200, fa.MODE_FOLD, 18, 3),
    (19, 4, 2, 18, 20, 20000, 2_000_003, 200, fa.MODE_FOLD, 17, 1),  # dense cache part: many queued windows
    (23, 4, 1, 16, 20, 2000, 2_000_000, 100, fa.MODE_FOLD, 15, 1),   # run-time window length
    (12, 3, 0, 10, 14, 3000, 600_011, 100, fa.MODE_GLOBAL, 0, 0),    # no cache part, many candidates
    (64, 5, 2, 20, 24, 1000, 1_000_000, 50, fa.MODE_AUTO, 0, 0),
]


@pytest.mark.parametrize("ci", range(len(RANDOM_CASES)))
def test_split_random_vs_oracle(ci):
    W, k, s, c, e, n_pat, n_text, plants, mode, fold, probes = RANDOM_CASES[ci]
    say(f"split random case {ci}: W={W} k={k} s={s} c={c} e={e} n_pat={n_pat} n={n_text} mode={mode}")
    sb = sbox64(500 + ci)
    b, off = fa.gen_patterns(fa.ALPHA_ASCII, 600 + ci, n_pat, W, W + 40)
    text = lines_text(700 + ci, n_text, b, off, plants=plants)
    o = SplitOracle(W, k, s, c, e, sb)
    r = o.run(text, b, off, threads=8)
    with make_split(W, k, s, c, e, sb, b, off) as m:
        m.set_scan_mode(mode, fold)
        if probes:
            m.set_scan_tuning(probes, 0)
        g = gpu_all(m, text)
        assert_same(g, r, m.filter(fa.FILTER_FORWARD))
        t = m.timings()
        if mode != fa.MODE_AUTO:
            assert t["scan_mode"] == mode
    assert r["ids"].size >= plants // 2


@pytest.mark.parametrize("shift", [0, 1, 5, 8, 15])
def test_split_unaligned_device_text(shift):
    say(f"split unaligned shift {shift}")
    sb = sbox64(9)
    b, off = fa.gen_patterns(fa.ALPHA_ASCII, 10, 800, 19, 50)
    text = lines_text(11, 200_003, b, off, line_mean=25, plants=80)
    o = SplitOracle(19, 5, 2, 16, 20, sb)
    r = o.run(text, b, off)
    buf = fa.DeviceBuffer(text.size + 32)
    buf.upload(text, offset=shift)
    with make_split(19, 5, 2, 16, 20, sb, b, off) as m:
        for mode in (fa.MODE_GLOBAL, fa.MODE_FOLD):
            m.set_scan_mode(mode)
            assert_same(gpu_all(m, dev=(buf.ptr + shift, text.size)), r)
    buf.free()


def test_split_line_edges():
    """Lines of exactly W-1 bytes hold no window; lines of exactly W bytes
    hold one; windows at the very start and end of the text and of lines."""
    say("split line edges")
    sb = sbox64(12)
    W = 19
    pats = [bytes([65 + (i % 26)]) * 3 + b"%016d" % i for i in range(64)]  # 19 bytes each
    b, off = fa.patterns_from_list(pats)
    o = SplitOracle(W, 5, 2, 14, 16, sb)
    with make_split(W, 5, 2, 14, 16, sb, b, off) as m:
        texts = [
            b"",
            b"\n" * 1000,
            b"\n".join(p[:W - 1] for p in pats),
            b"\n".join(pats),
            b"\n".join(pats) + b"\n",
            pats[0] + b"x" * 3 + pats[1],
            b"\n" + pats[2] + b"\n" + pats[3],
            bytes(range(256)) * 40 + pats[5],
        ]
        for i, t in enumerate(texts):
            say(f"  text {i} ({len(t)} bytes)")
            arr = np.frombuffer(t, dtype=np.uint8).copy()
            r = o.run(arr, b, off)
            for mode in (fa.MODE_GLOBAL, fa.MODE_FOLD):
                m.set_scan_mode(mode)
                assert_same(gpu_all(m, arr), r)
        full = np.frombuffer(b"\n".join(pats), dtype=np.uint8).copy()
        assert m.match(full).pairs == [(i, i * (W + 1)) for i in range(64)]
        short = np.frombuffer(b"\n".join(p[:W - 1] for p in pats), dtype=np.uint8).copy()
        m.scan(short)
        assert m.candidates().size == 0


def test_split_ingest_outputs_and_persistence(tmp_path):
    say("split ingest, outputs, persistence")
    sb = sbox64(13)
    W, k, s, c, e = 19, 5, 2, 24, 28
    b, off = fa.gen_patterns(fa.ALPHA_ASCII, 14, 2000, 19, 60)
    text = lines_text(15, (3 << 20) + 77, b, off, plants=150)
    o = SplitOracle(W, k, s, c, e, sb)
    r = o.run(text, b, off, threads=8)
    with make_split(W, k, s, c, e, sb, b, off) as m:
        m.set_ingest(1 << 20, 3, 4)
        assert_same(gpu_all(m, text), r, m.filter())  # host text in 1 MiB pieces
        path = tmp_path / "t.bin"
        text.tofile(path)
        m.scan_file(path)
        m.pattern_pass()
        pats, idx = tmp_path / "p.out", tmp_path / "__index_p.out"
        m.write_survivors(pats, idx)
        want_p, want_i = survivor_files(b, off, r["surv"])
        assert pats.read_bytes() == want_p and idx.read_bytes() == want_i
        corpus = tmp_path / "corpus.out"
        m.write_candidate_lines(corpus, text)
        assert corpus.read_bytes() == candidate_lines(text, r["cand"])
        d = m.filter_save()
        assert np.array_equal(d, filter_save_deltas_bits(r["F"], o.n_bits))
    with fa.Matcher.split(W, k, s, c, e, sb) as m2:
        m2.set_patterns(b, off)
        m2.filter_load(d)
        assert_same(gpu_all(m2, text), r, m2.filter())


def test_split_rejected_where_unsupported():
    say("split: mode and mixing errors")
    sb = sbox64(16)
    b, off = fa.gen_patterns(fa.ALPHA_ASCII, 17, 100, 19, 30)
    with make_split(19, 5, 2, 16, 18, sb, b, off) as m:
        for mode in (fa.MODE_LDS, fa.MODE_GFOLD, fa.MODE_HYBRID):
            with pytest.raises(fa.FFBFError) as e:
                m.set_scan_mode(mode)
            assert e.value.status == fa.EINVAL
        m.set_scan_mode(fa.MODE_AUTO)
        # one hash method per batch
        with fa.Matcher(19, 5, 20, fa.sbox_from_seed(3)) as pair:
            pair.set_patterns(b, off)
            pair.build()
            with pytest.raises(fa.FFBFError) as e:
                fa.Batch([m, pair])
            assert e.value.status == fa.EINVAL
    # a cache part of one word cannot be folded (auto picks F probed directly)
    with make_split(19, 3, 3, 5, 10, sb, b, off) as m:
        with pytest.raises(fa.FFBFError) as e:
            m.set_scan_mode(fa.MODE_FOLD)
        assert e.value.status == fa.EINVAL
        m.set_scan_mode(fa.MODE_AUTO)
        assert m.timings()["scan_mode"] == fa.MODE_GLOBAL
    for bad in [(19, 12, 2, 24, 28), (19, 5, 6, 24, 28), (19, 5, 2, 4, 28), (19, 5, 2, 24, 32), (0, 5, 2, 24, 28)]:
        with pytest.raises(fa.FFBFError) as e:
            fa.Matcher.split(*bad, sb)
        assert e.value.status == fa.EINVAL


def split_batch_filters(W, k, sb, shapes, seed):
    out = []
    for i, (s, c, e, mode, n_pat) in enumerate(shapes):
        b, off = fa.gen_patterns(fa.ALPHA_ASCII, seed + i, n_pat, W, W + 40)
        m = make_split(W, k, s, c, e, sb, b, off)
        m.set_scan_mode(mode)
        out.append((m, b, off, (s, c, e)))
    return out


@pytest.mark.parametrize("W", [19, 23])
def test_split_batch_and_server(W):
    say(f"split batch and server W={W}")
    sb = sbox64(40 + W)
    k = 5
    shapes = [(2, 20, 24, fa.MODE_FOLD, 3000), (2, 24, 28, fa.MODE_GLOBAL, 20000), (1, 16, 22, fa.MODE_AUTO, 500),
              (0, 10, 18, fa.MODE_GLOBAL, 2000), (3, 18, 20, fa.MODE_FOLD, 8000)]
    filters = split_batch_filters(W, k, sb, shapes, 700)
    rng = np.random.default_rng(W)
    n_texts = 70
    route = rng.integers(0, len(filters), size=n_texts).astype(np.uint32)
    texts = []
    for i in range(n_texts):
        _, b, off, _ = filters[int(route[i])]
        n = int(rng.integers(0, 150_000)) if i > 3 else [0, 3, W - 1, W][i]
        texts.append(lines_text(900 + i, n, b, off, line_mean=int(rng.integers(20, 200)), plants=1 + n // 20000))
    rows = []
    for i, t in enumerate(texts):
        _, b, off, (s, c, e) = filters[int(route[i])]
        r = SplitOracle(W, k, s, c, e, sb).run(t, b, off)
        rows += [(i, int(p), int(q)) for p, q in zip(r["ids"], r["offs"])]
    rows.sort()
    assert len(rows) > 30
    offs = np.zeros(n_texts + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([t.size for t in texts])
    dev = fa.DeviceBuffer.from_host(np.concatenate(texts))
    with fa.Batch([f[0] for f in filters]) as bt:
        for tuning in (dict(), dict(chunk_iters=3, probes=3, use_graph=False), dict(chunk_iters=200, probes=1)):
            bt.set_tuning(**tuning)
            bt.set_texts(dev.ptr, offs, route)
            assert bt.run() == len(rows)
            m = bt.matches()
            assert list(zip(m["text_id"].tolist(), m["pattern_id"].tolist(), m["offset"].tolist())) == rows
        assert bt.run_per_text() == len(rows)
    dev.free()
    # the serving loop over the same split filters: every request's matches
    # equal its rows above
    with fa.Server([f[0] for f in filters], hbm_budget_bytes=16 << 20, slots=3) as srv:
        ids = [srv.submit(int(route[i]), texts[i]) for i in range(n_texts)]
        for i, rid in enumerate(ids):
            got, _ = srv.wait(rid)
            want = [(p, q) for (t, p, q) in rows if t == i]
            assert list(zip(got["pattern_id"].tolist(), got["offset"].tolist())) == want, i
    for f in filters:
        f[0].close()
