"""Both filter builds give the oracle's F, word for word: the atomic build
(F cleared, one atomicOr per bit) and the binned build (bit indices sorted
by 2^20-bit slice, each slice set in LDS and written whole).  Pair and split
methods, filters from part of one slice to hundreds of slices, k = 1..5
(pair) and up to 11 (split), pattern sets with patterns shorter than the window (they set no bits), an
empty set, and one context rebuilt with the other method and another set
(no bits of the first build may survive).  Needs an MI355X."""
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


# (W, k, log2 bits, patterns, min len, max len, alphabet)
PAIR_CASES = [
    (16, 1, 16, 3_000, 8, 40, fa.ALPHA_DNA_GC40),  # part of one slice; some patterns shorter than W
    (19, 4, 20, 20_000, 12, 40, fa.ALPHA_ASCII),  # exactly one slice
    (24, 3, 21, 50_000, 16, 60, fa.ALPHA_DNA_GC40),  # two slices
    (32, 5, 24, 100_000, 24, 64, fa.ALPHA_ASCII),  # 16 slices
    (19, 2, 26, 300_000, 19, 50, fa.ALPHA_ASCII),  # 64 slices, 600k bit sets
]


def build_with(m, method, b, off):
    m.set_build_method(method)
    m.set_patterns(b, off)
    m.build()
    return m.filter()


@pytest.mark.parametrize("ci", range(len(PAIR_CASES)))
def test_pair_builds_equal_oracle(ci):
    W, k, log2, n, lmin, lmax, alpha = PAIR_CASES[ci]
    say(f"pair build methods: W={W} k={k} 2^{log2} bits, {n} patterns of {lmin}-{lmax} B")
    sbox = fa.sbox_from_seed(700 + ci)
    b, off = fa.gen_patterns(alpha, 710 + ci, n, lmin, lmax)
    F = Oracle(W, k, log2, sbox).build(b, off)
    assert F.any()
    for method in METHODS:
        with fa.Matcher(W, k, log2, sbox) as m:
            assert np.array_equal(build_with(m, method, b, off), F), method


def test_rebuild_with_other_method_and_set():
    """The binned build writes every word of F without clearing it first:
    bits of an earlier build with another set must not survive, in either
    order of the two methods."""
    W, k, log2 = 24, 4, 22
    say("rebuild with the other method and another set")
    sbox = fa.sbox_from_seed(720)
    b1, o1 = fa.gen_patterns(fa.ALPHA_DNA_GC40, 721, 80_000, 24, 60)
    b2, o2 = fa.gen_patterns(fa.ALPHA_DNA_GC40, 722, 5_000, 20, 60)
    o = Oracle(W, k, log2, sbox)
    F1, F2 = o.build(b1, o1), o.build(b2, o2)
    for first, second in ((fa.BUILD_ATOMIC, fa.BUILD_BINNED), (fa.BUILD_BINNED, fa.BUILD_ATOMIC),
                          (fa.BUILD_BINNED, fa.BUILD_BINNED)):
        with fa.Matcher(W, k, log2, sbox) as m:
            assert np.array_equal(build_with(m, first, b1, o1), F1), (first, second)
            assert np.array_equal(build_with(m, second, b2, o2), F2), (first, second)


def test_empty_and_all_short_sets():
    W, k, log2 = 32, 4, 22
    sbox = fa.sbox_from_seed(730)
    short_b, short_o = fa.gen_patterns(fa.ALPHA_ASCII, 731, 1_000, 4, 31)  # all shorter than W
    for method in METHODS:
        with fa.Matcher(W, k, log2, sbox) as m:
            # a first build with bits, so that the empty builds must clear them
            b, o = fa.gen_patterns(fa.ALPHA_ASCII, 732, 2_000, 32, 40)
            assert build_with(m, method, b, o).any()
            assert not build_with(m, method, short_b, short_o).any(), method
            assert not build_with(m, method, np.zeros(0, dtype=np.uint8), np.zeros(1, dtype=np.uint64)).any(), method


@pytest.mark.parametrize("shape", [(19, 5, 2, 16, 20, 5_000), (19, 5, 2, 24, 28, 200_000), (24, 3, 3, 20, 24, 60_000),
                                   (19, 11, 3, 20, 22, 30_000)])
def test_split_builds_equal_oracle(shape):
    W, k, s, c, e, n = shape
    say(f"split build methods: W={W} k={k} s={s} cache 2^{c} + external 2^{e} bits, {n} patterns")
    sb = sbox64(740 + c)
    b, off = fa.gen_patterns(fa.ALPHA_ASCII, 741 + c, n, max(8, W - 6), 64)
    F = SplitOracle(W, k, s, c, e, sb).build(b, off)
    assert F.any()
    for method in METHODS:
        with fa.Matcher.split(W, k, s, c, e, sb) as m:
            assert np.array_equal(build_with(m, method, b, off), F), method


def test_build_method_is_checked():
    with fa.Matcher(19, 4, 16, fa.sbox_from_seed(750)) as m:
        for bad in (-1, 3, 7):
            with pytest.raises(fa.FFBFError) as err:
                m.set_build_method(bad)
            assert err.value.status == fa.EINVAL
