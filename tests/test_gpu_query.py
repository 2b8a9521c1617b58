"""Line queries (ffbf_query_line / ffbf_query_lines, Filter::operator[] of
ffbf.cpp:285-308) on the GPU against the oracle, for both hash methods.
Needs an MI355X."""
import numpy as np
import pytest

import ffbf_amd as fa
from oracle import Oracle, SplitOracle

pytestmark = pytest.mark.gpu


def _lines(seed, alpha, n, lmax, pb, po, nl=False):
    rng = np.random.default_rng(seed)
    lines = []
    for i in range(n):
        L = int(rng.integers(0, lmax)) if i % 50 else int(rng.integers(0, 20000))
        t = fa.gen_bytes(alpha, seed * 1000 + i, L)
        if nl and L:
            t[rng.integers(0, L, size=max(1, L // 90))] = 10
            t[rng.integers(0, L, size=max(1, L // 120))] = 0xC3
        if i % 3 == 0 and L > 100:
            j = int(rng.integers(0, po.size - 1))
            p = pb[int(po[j]):int(po[j + 1])]
            at = int(rng.integers(0, L - p.size + 1)) if L >= p.size else 0
            t[at:at + p.size] = p[:L - at]
        lines.append(t)
    off = np.zeros(n + 1, dtype=np.uint64)
    off[1:] = np.cumsum([l.size for l in lines])
    return np.concatenate(lines), off


@pytest.mark.parametrize("W,k,log2,alpha", [(19, 4, 16, fa.ALPHA_ASCII), (32, 4, 22, fa.ALPHA_DNA_GC40),
                                            (7, 2, 12, fa.ALPHA_DNA_GC40), (64, 5, 20, fa.ALPHA_ASCII)])
def test_query_lines_pair(W, k, log2, alpha):
    sbox = fa.sbox_from_seed(300 + W)
    pb, po = fa.gen_patterns(alpha, 301 + W, 3000, W, W + 40)
    b, off = _lines(302 + W, alpha, 2000, 600, pb, po)
    o = Oracle(W, k, log2, sbox)
    F = o.build(pb, po)
    want = o.query_lines(b, off, F)
    with fa.Matcher(W, k, log2, sbox) as m:
        m.set_patterns(pb, po)
        with pytest.raises(fa.FFBFError) as e:
            m.query_line(b"x" * 100)
        assert e.value.status == fa.ESTATE
        m.build()
        R0 = m.filter(fa.FILTER_REVERSE)
        got = m.query_lines(b, off)
        assert np.array_equal(got, want)
        assert 100 < int(want.sum()) < 2000
        for i in (0, 3, 6, 50, 51):
            assert m.query_line(b[int(off[i]):int(off[i + 1])]) == bool(want[i])
        assert np.array_equal(m.filter(fa.FILTER_REVERSE), R0)  # nothing fed forward
        assert m.query_lines(b, np.zeros(1, dtype=np.uint64)).size == 0
        with pytest.raises(fa.FFBFError):
            m.query_lines(b, np.array([0, 5, 3], dtype=np.uint64))


def test_query_lines_split_method():
    sb64 = np.random.default_rng(310).integers(0, 2**64, size=256, dtype=np.uint64)
    pb, po = fa.gen_patterns(fa.ALPHA_ASCII, 311, 4000, 19, 50)
    # lines with newlines and high bytes inside: operator[] hashes through both
    b, off = _lines(312, fa.ALPHA_ASCII, 1500, 500, pb, po, nl=True)
    for params in ((19, 5, 2, 16, 22), (19, 11, 0, 10, 14), (23, 3, 3, 12, 12)):
        o = SplitOracle(*params, sb64)
        F = o.build(pb, po)
        want = o.query_lines(b, off, F)
        with fa.Matcher.split(*params, sb64) as m:
            m.set_patterns(pb, po)
            m.build()
            got = m.query_lines(b, off)
            assert np.array_equal(got, want), params
            assert int(want.sum()) > 100
