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
