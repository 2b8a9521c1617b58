"""The scan's prefilter folds, word for word, after every build.

An OR-fold of F onto w words (a power of two) has word j = OR of F's words
congruent to j modulo w.  A small atomic build sets the folds' bits itself,
beside F's (bit x of F is bit x mod 32w of the fold); larger atomic builds
and the binned build fold the finished F in a separate pass; a scan-mode
change after the build refolds F.  Each way must give numpy's fold of the F
the build produced, and F must be the oracle's.  Pair method: the LDS fold
(mode 1, automatic and forced sizes), the L2-resident fold alone (mode 3) and
both (mode 4, two folds set by one build); split method: the fold of F's
cache part.  Needs an MI355X."""
import ctypes
import sys

import numpy as np
import pytest

import ffbf_amd as fa
from oracle import Oracle, SplitOracle
from test_gpu_split import sbox64

pytestmark = pytest.mark.gpu

METHODS = (fa.BUILD_ATOMIC, fa.BUILD_BINNED, fa.BUILD_AUTO)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fold_of(F, words):
    return np.bitwise_or.reduce(F.reshape(-1, words), axis=0)


def check_folds(m, F, src_words=None):
    """The folds the mode in force holds, against numpy's folds of F (of its
    first src_words words when given)."""
    t = m.timings()
    src = F if src_words is None else F[:src_words]
    checked = 0
    if t["scan_mode"] in (fa.MODE_FOLD, fa.MODE_HYBRID):
        fold = m.filter(fa.FILTER_FOLD)
        assert fold.size == (1 << t["fold_log2_bits"]) // 32
        assert np.array_equal(fold, fold_of(src, fold.size)), "LDS fold"
        checked += 1
    if t["scan_mode"] in (fa.MODE_GFOLD, fa.MODE_HYBRID):
        g = m.filter(fa.FILTER_L2_FOLD)
        assert g.size == (1 << t["l2_fold_log2_bits"]) // 32
        assert np.array_equal(g, fold_of(src, g.size)), "L2 fold"
        checked += 1
    return checked


# (W, k, log2 bits, patterns, mode, fold log2 bits, folds the mode holds)
PAIR_CASES = [
    (19, 4, 24, 10_000, fa.MODE_AUTO, 0, 1),  # the 10k-pattern ASCII shape: a 2^20-bit fold, set by the atomics
    (19, 4, 24, 10_000, fa.MODE_FOLD, 17, 1),  # a smaller fold
    (19, 4, 26, 10_000, fa.MODE_GFOLD, 0, 1),  # the L2-resident fold alone
    (19, 4, 26, 8_000, fa.MODE_HYBRID, 0, 2),  # both folds from one atomic build
    (24, 3, 24, 60_000, fa.MODE_FOLD, 0, 1),  # 180k fold atomics: too many, folded in a separate pass
]


@pytest.mark.parametrize("ci", range(len(PAIR_CASES)))
def test_pair_folds(ci):
    W, k, log2, n, mode, fold_log2, n_folds = PAIR_CASES[ci]
    say(f"pair folds: W={W} k={k} 2^{log2} bits, {n} patterns, mode {mode}, fold 2^{fold_log2}")
    sbox = fa.sbox_from_seed(800 + ci)
    b, off = fa.gen_patterns(fa.ALPHA_ASCII, 810 + ci, n, W, W + 20)
    F_ref = Oracle(W, k, log2, sbox).build(b, off)
    for method in METHODS:
        with fa.Matcher(W, k, log2, sbox) as m:
            m.set_scan_mode(mode, fold_log2)
            m.set_build_method(method)
            m.set_patterns(b, off)
            m.build()
            F = m.filter()
            assert np.array_equal(F, F_ref), method
            assert check_folds(m, F) == n_folds, method
            # a second build over the same buffers: the clear must remove
            # nothing the rebuild sets and leave nothing else
            m.build()
            assert np.array_equal(m.filter(), F_ref), method
            assert check_folds(m, F_ref) == n_folds, method


def test_mode_change_after_build_refolds():
    W, k, log2 = 19, 4, 26
    sbox = fa.sbox_from_seed(830)
    b, off = fa.gen_patterns(fa.ALPHA_ASCII, 831, 9_000, 19, 40)
    with fa.Matcher(W, k, log2, sbox) as m:
        m.set_build_method(fa.BUILD_ATOMIC)
        m.set_patterns(b, off)
        m.build()
        F = m.filter()
        for mode, fl in ((fa.MODE_FOLD, 18), (fa.MODE_HYBRID, 0), (fa.MODE_GFOLD, 0), (fa.MODE_FOLD, 20)):
            m.set_scan_mode(mode, fl)
            assert check_folds(m, F) >= 1, (mode, fl)
        m.set_scan_mode(fa.MODE_GLOBAL)
        with pytest.raises(fa.FFBFError) as err:
            m.filter(fa.FILTER_FOLD)
        assert err.value.status == fa.ESTATE


@pytest.mark.parametrize("n", [20_000, 40_000])  # 100k fold atomics (set by the build) and 200k (separate pass)
def test_split_cache_fold(n):
    W, k, s, c, e = 19, 5, 2, 24, 28
    say(f"split fold: {n} patterns")
    sb = sbox64(840)
    b, off = fa.gen_patterns(fa.ALPHA_ASCII, 841, n, 20, 60)
    F_ref = SplitOracle(W, k, s, c, e, sb).build(b, off)
    for method in METHODS:
        with fa.Matcher.split(W, k, s, c, e, sb) as m:
            m.set_scan_mode(fa.MODE_FOLD)
            m.set_build_method(method)
            m.set_patterns(b, off)
            m.build()
            F = m.filter()
            assert np.array_equal(F, F_ref), method
            assert check_folds(m, F, src_words=(1 << c) // 32) == 1, method


def noalloc(m, dev, n, out, cap):
    nm = ctypes.c_uint64(0)
    st = fa.lib().ffbf_run_device_noalloc(m._h, dev.ptr, n, out.ptr, cap, ctypes.byref(nm))
    return st, nm.value


def run_and_check(m, dev, text, r):
    assert m.run_device(dev.ptr, text.size) == r["ids"].size
    mm = m.matches()
    assert np.array_equal(mm["pattern_id"], r["ids"]) and np.array_equal(mm["offset"], r["offs"])
    assert np.array_equal(m.candidates(), r["cand"])
    assert np.array_equal(m.filter(fa.FILTER_REVERSE), r["R"])


def test_scan_after_build_clear():
    """The atomic build clears the scan state (R, region counts, counters,
    survivor index) with F, and the scan right after it then skips its own
    clear.  Every order must still scan from zero: a scan after a build, a
    scan after a scan, two builds, an allocation-free run or a batch between
    the build and the scan (both write R and the survivor index)."""
    W, k, log2 = 19, 4, 24
    sbox = fa.sbox_from_seed(850)
    b, off = fa.gen_patterns(fa.ALPHA_ASCII, 851, 10_000, 20, 40)
    texts = []
    for s in range(2):
        t = fa.gen_bytes(fa.ALPHA_ASCII, 852 + s, 3 << 20)
        fa.plant(860 + s, 300, t, b, off)
        texts.append(t)
    o = Oracle(W, k, log2, sbox)
    refs = [o.run(t, b, off) for t in texts]
    assert all(r["ids"].size >= 300 for r in refs)
    devs = [fa.DeviceBuffer.from_host(t) for t in texts]
    out = fa.DeviceBuffer(4096 * fa.MATCH_DTYPE.itemsize)
    try:
        with fa.Matcher(W, k, log2, sbox) as m:
            m.set_build_method(fa.BUILD_ATOMIC)
            m.set_patterns(b, off)
            say("build, scan, scan")
            m.build()
            run_and_check(m, devs[0], texts[0], refs[0])
            run_and_check(m, devs[1], texts[1], refs[1])
            say("build, build, scan")
            m.build()
            m.build()
            run_and_check(m, devs[0], texts[0], refs[0])
            say("build, allocation-free run, scan")
            m.build()
            assert noalloc(m, devs[1], texts[1].size, out, 4096) == (0, refs[1]["ids"].size)
            run_and_check(m, devs[0], texts[0], refs[0])
            say("build, batch, scan")
            m.build()
            with fa.Batch([m]) as bt:
                bt.set_texts(devs[1].ptr, np.array([0, texts[1].size], dtype=np.uint64),
                             np.zeros(1, dtype=np.uint32))
                assert bt.run() == refs[1]["ids"].size
            run_and_check(m, devs[0], texts[0], refs[0])
            say("build, scan after the build's clear, binned build, scan")
            m.build()
            run_and_check(m, devs[1], texts[1], refs[1])
            m.set_build_method(fa.BUILD_BINNED)
            m.build()
            run_and_check(m, devs[0], texts[0], refs[0])
    finally:
        out.free()
        for d in devs:
            d.free()
