"""Seeded random configurations of both hash methods against the oracle:
window length, k, filter size (and for the split filter s and the two part
sizes), scan mode, up-front probes and block size, text length and
alphabet (including newline-rich texts with high bytes), pattern counts
with short and duplicate patterns, device-pointer alignment.  Every case
compares F, the candidates, R, the survivors and the sorted match list bit
for bit.  Needs an MI355X (-m gpu)."""
import sys

import numpy as np
import pytest

import ffbf_amd as fa
from oracle import Oracle, SplitOracle

pytestmark = pytest.mark.gpu

N_CASES = 24


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_gpu(m, text, shift):
    buf = fa.DeviceBuffer(text.size + 32)
    if text.size:
        buf.upload(text, offset=shift)
    m.scan_device(buf.ptr + shift, text.size)
    out = {"cand": m.candidates(), "R": m.filter(fa.FILTER_REVERSE)}
    m.pattern_pass()
    out["surv"] = m.survivors()
    m.verify()
    out["matches"] = m.matches()
    buf.free()
    return out


def check(g, r, F):
    assert np.array_equal(F, r["F"]), "forward filter"
    assert np.array_equal(g["cand"], r["cand"]), f"candidates {g['cand'].size} vs {r['cand'].size}"
    assert np.array_equal(g["R"], r["R"]), "reverse filter"
    assert np.array_equal(g["surv"], r["surv"]), "survivors"
    assert np.array_equal(g["matches"]["pattern_id"], r["ids"]), "match ids"
    assert np.array_equal(g["matches"]["offset"], r["offs"]), "match offsets"


def random_inputs(rng, W, lines):
    alpha = fa.ALPHA_ASCII if rng.random() < 0.6 else fa.ALPHA_DNA_GC40
    n_pat = int(rng.integers(0, 3000))
    lmin = max(1, W - int(rng.integers(0, 6)))  # some patterns shorter than W
    lmax = lmin + int(rng.integers(0, 60))
    seed = int(rng.integers(1, 1 << 30))
    b, off = fa.gen_patterns(alpha, seed, n_pat, lmin, lmax)
    if n_pat > 10:  # duplicates
        lst = [bytes(b[int(off[i]):int(off[i + 1])]) for i in range(n_pat)]
        lst += [lst[int(rng.integers(0, n_pat))] for _ in range(5)]
        b, off = fa.patterns_from_list(lst)
    n = int(rng.choice([0, 1, W - 1, W, int(rng.integers(W, 5000)), int(rng.integers(5000, 400_000))],
                       p=[0.04, 0.04, 0.04, 0.04, 0.34, 0.5]))
    n = max(n, 0)
    text = fa.gen_bytes(alpha, seed + 1, n)
    if n and lines:
        text[rng.integers(0, n, size=max(1, n // int(rng.integers(8, 200))))] = 10
        text[rng.integers(0, n, size=max(1, n // 300))] = rng.integers(0x80, 0x100, size=max(1, n // 300))
    if n > 200 and off.size > 1:
        fa.plant(seed + 2, 1 + n // 5000, text, b, off)
    return b, off, text


@pytest.mark.parametrize("case", range(N_CASES))
def test_pair_method_random(case):
    rng = np.random.default_rng(9000 + case)
    W = int(rng.choice([1, 5, 12, 19, 24, 31, 32, 33, 47, 64]))
    k = int(rng.integers(1, 6))
    log2 = int(rng.integers(10, 23))
    mode = int(rng.choice([fa.MODE_AUTO, fa.MODE_GLOBAL, fa.MODE_FOLD, fa.MODE_LDS, fa.MODE_GFOLD]))
    if mode == fa.MODE_LDS and log2 > 20:
        mode = fa.MODE_AUTO
    fold = {fa.MODE_FOLD: int(rng.integers(7, min(log2, 20) + 1)),
            fa.MODE_GFOLD: int(rng.integers(10, log2 + 1))}.get(mode, 0)
    probes = int(rng.integers(0, k + 1))
    block = int(rng.choice([0, 256, 512, 1024]))
    shift = int(rng.integers(0, 16))
    sbox = fa.sbox_from_seed(int(rng.integers(1, 1 << 30)))
    b, off, text = random_inputs(rng, W, lines=False)
    build = int(rng.choice([fa.BUILD_AUTO, fa.BUILD_ATOMIC, fa.BUILD_BINNED]))
    say(f"pair case {case}: W={W} k={k} log2={log2} mode={mode} fold={fold} probes={probes} block={block} "
        f"n_pat={off.size - 1} n={text.size} shift={shift} build={build}")
    r = Oracle(W, k, log2, sbox).run(text, b, off)
    with fa.Matcher(W, k, log2, sbox) as m:
        m.set_build_method(build)
        m.set_patterns(b, off)
        m.build()
        m.set_scan_mode(mode, fold)
        m.set_scan_tuning(probes, block)
        check(run_gpu(m, text, shift), r, m.filter(fa.FILTER_FORWARD))


@pytest.mark.parametrize("case", range(N_CASES // 2))
def test_pair_method_random_beyond_l2(case):
    """Filters of 2^26-2^27 bits (beyond one XCD's L2) in the modes that
    probe the L2-resident fold: mode 4 with LDS folds of 2^7-2^20 bits (the
    small ones pass most windows, so the L2 fold's first bit is tested before
    queueing; the large ones pass few, and the drain probes it), mode 3, the
    LDS fold alone and the automatic choice."""
    rng = np.random.default_rng(9300 + case)
    W = int(rng.choice([5, 19, 24, 32, 33]))
    k = int(rng.integers(1, 6))
    log2 = int(rng.integers(26, 28))
    mode = int(rng.choice([fa.MODE_AUTO, fa.MODE_FOLD, fa.MODE_GFOLD, fa.MODE_HYBRID, fa.MODE_HYBRID]))
    fold = {fa.MODE_FOLD: int(rng.integers(7, 21)), fa.MODE_HYBRID: int(rng.integers(7, 21)),
            fa.MODE_GFOLD: int(rng.integers(20, 26))}.get(mode, 0)
    probes = int(rng.integers(0, k + 1))
    block = int(rng.choice([0, 512, 1024]))
    shift = int(rng.integers(0, 16))
    sbox = fa.sbox_from_seed(int(rng.integers(1, 1 << 30)))
    b, off, text = random_inputs(rng, W, lines=False)
    build = int(rng.choice([fa.BUILD_AUTO, fa.BUILD_ATOMIC, fa.BUILD_BINNED]))
    say(f"pair case beyond L2 {case}: W={W} k={k} log2={log2} mode={mode} fold={fold} probes={probes} "
        f"block={block} n_pat={off.size - 1} n={text.size} shift={shift} build={build}")
    r = Oracle(W, k, log2, sbox).run(text, b, off)
    with fa.Matcher(W, k, log2, sbox) as m:
        m.set_build_method(build)
        m.set_patterns(b, off)
        m.build()
        m.set_scan_mode(mode, fold)
        m.set_scan_tuning(probes, block)
        assert m.timings()["scan_mode"] == mode or mode == fa.MODE_AUTO
        check(run_gpu(m, text, shift), r, m.filter(fa.FILTER_FORWARD))


@pytest.mark.parametrize("case", range(N_CASES // 2))
def test_split_method_random(case):
    rng = np.random.default_rng(9500 + case)
    W = int(rng.choice([1, 7, 19, 19, 24, 33, 64]))
    k = int(rng.integers(1, 12))
    s = int(rng.integers(0, k + 1))
    c = int(rng.integers(5, 22))
    e = int(rng.integers(5, 24))
    mode = int(rng.choice([fa.MODE_AUTO, fa.MODE_GLOBAL, fa.MODE_FOLD]))
    if s == 0:
        mode = fa.MODE_GLOBAL
    fold = int(rng.integers(7, min(c, 20) + 1)) if mode == fa.MODE_FOLD and c >= 7 else 0
    if mode == fa.MODE_FOLD and c < 7:
        mode = fa.MODE_GLOBAL
    probes = int(rng.integers(0, k + 1))
    shift = int(rng.integers(0, 16))
    sb = rng.integers(0, 2**64, size=256, dtype=np.uint64)
    b, off, text = random_inputs(rng, W, lines=True)
    build = int(rng.choice([fa.BUILD_AUTO, fa.BUILD_ATOMIC, fa.BUILD_BINNED]))
    say(f"split case {case}: W={W} k={k} s={s} c={c} e={e} mode={mode} fold={fold} probes={probes} "
        f"n_pat={off.size - 1} n={text.size} shift={shift} build={build}")
    r = SplitOracle(W, k, s, c, e, sb).run(text, b, off)
    with fa.Matcher.split(W, k, s, c, e, sb) as m:
        m.set_build_method(build)
        m.set_patterns(b, off)
        m.build()
        m.set_scan_mode(mode, fold)
        m.set_scan_tuning(probes, 0)
        check(run_gpu(m, text, shift), r, m.filter(fa.FILTER_FORWARD))


def _batch_case(rng):
    """Random resident filters (mode, size, k fixed per batch) and texts."""
    W = int(rng.choice([12, 19, 24, 32, 40]))
    k = int(rng.integers(1, 6))
    sbox = fa.sbox_from_seed(int(rng.integers(1, 1 << 30)))
    nf = int(rng.integers(1, 9))
    filters = []
    for f in range(nf):
        log2 = int(rng.integers(10, 24))
        mode = int(rng.choice([fa.MODE_AUTO, fa.MODE_GLOBAL, fa.MODE_FOLD, fa.MODE_LDS, fa.MODE_GFOLD]))
        if mode == fa.MODE_LDS and log2 > 20:
            mode = fa.MODE_FOLD
        fold = {fa.MODE_FOLD: int(rng.integers(7, min(log2, 20) + 1)),
                fa.MODE_GFOLD: int(rng.integers(10, log2 + 1))}.get(mode, 0)
        alpha = fa.ALPHA_ASCII if rng.random() < 0.5 else fa.ALPHA_DNA_GC40
        b, off = fa.gen_patterns(alpha, int(rng.integers(1, 1 << 30)), int(rng.integers(0, 4000)),
                                 max(1, W - 3), W + 40)
        m = fa.Matcher(W, k, log2, sbox)
        m.set_patterns(b, off)
        m.build()
        m.set_scan_mode(mode, fold)
        filters.append((m, b, off, log2, alpha))
    n_texts = int(rng.integers(1, 120))
    lens = rng.integers(0, int(rng.choice([100, 5000, 300_000])), size=n_texts)
    route = rng.integers(0, nf, size=n_texts).astype(np.uint32)
    texts = []
    for i in range(n_texts):
        _, b, off, _, alpha = filters[int(route[i])]
        t = fa.gen_bytes(alpha, int(rng.integers(1, 1 << 30)), int(lens[i]))
        if t.size > 200 and off.size > 1:
            fa.plant(int(rng.integers(1, 1 << 30)), 1 + t.size // 10000, t, b, off)
        texts.append(t)
    return W, k, sbox, filters, texts, route


@pytest.mark.parametrize("case", range(10))
def test_batch_random(case):
    rng = np.random.default_rng(9800 + case)
    W, k, sbox, filters, texts, route = _batch_case(rng)
    tuning = dict(chunk_iters=int(rng.choice([0, 1, 5, 32, 300])), probes=int(rng.integers(0, min(3, k) + 1)),
                  block=int(rng.choice([0, 512, 1024])), use_graph=bool(rng.random() < 0.7),
                  bytes_per_lane=int(np.random.default_rng(9900 + case).choice([0, 16, 32])))
    say(f"batch case {case}: W={W} k={k} filters={len(filters)} texts={len(texts)} "
        f"bytes={sum(t.size for t in texts)} tuning={tuning}")
    rows = []
    for i, t in enumerate(texts):
        _, b, off, log2, _ = filters[int(route[i])]
        r = Oracle(W, k, log2, sbox).run(t, b, off)
        rows += [(i, int(p), int(q)) for p, q in zip(r["ids"], r["offs"])]
    rows.sort()
    lens = np.array([t.size for t in texts], dtype=np.uint64)
    offs = np.zeros(len(texts) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(lens)
    dev = fa.DeviceBuffer.from_host(np.concatenate(texts) if offs[-1] else np.zeros(1, np.uint8))
    with fa.Batch([f[0] for f in filters]) as bt:
        bt.set_tuning(**tuning)
        bt.set_texts(dev.ptr, offs, route)
        n = bt.run()
        m = bt.matches()
        got = list(zip(m["text_id"].tolist(), m["pattern_id"].tolist(), m["offset"].tolist()))
        assert n == len(rows) and got == rows
    dev.free()
    for f in filters:
        f[0].close()


@pytest.mark.parametrize("case", range(6))
def test_host_ingest_random(case):
    rng = np.random.default_rng(9900 + case)
    W = int(rng.choice([7, 19, 32]))
    sbox = fa.sbox_from_seed(int(rng.integers(1, 1 << 30)))
    b, off = fa.gen_patterns(fa.ALPHA_ASCII, int(rng.integers(1, 1 << 30)), 2000, W, W + 30)
    chunk = int(rng.choice([1, 2, 3])) << 20
    depth, threads = int(rng.integers(2, 5)), int(rng.integers(1, 9))
    n = int(rng.choice([chunk - 1, chunk, chunk + W // 2, 3 * chunk + int(rng.integers(0, 1 << 20)),
                        int(rng.integers(1, 6 << 20))]))
    text = fa.gen_bytes(fa.ALPHA_ASCII, int(rng.integers(1, 1 << 30)), n)
    fa.plant(int(rng.integers(1, 1 << 30)), 1 + n // 50000, text, b, off)
    say(f"ingest case {case}: W={W} chunk={chunk} depth={depth} threads={threads} n={n}")
    r = Oracle(W, 4, 20, sbox).run(text, b, off, threads=8)
    with fa.Matcher(W, 4, 20, sbox) as m:
        m.set_patterns(b, off)
        m.build()
        m.set_ingest(chunk, depth, threads)
        m.scan(text)
        assert np.array_equal(m.candidates(), r["cand"])
        assert np.array_equal(m.filter(fa.FILTER_REVERSE), r["R"])
        m.pattern_pass()
        m.verify()
        got = m.matches()
        assert np.array_equal(got["pattern_id"], r["ids"]) and np.array_equal(got["offset"], r["offs"])
