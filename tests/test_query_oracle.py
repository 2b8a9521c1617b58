"""The oracle's line query (Filter::operator[], ffbf.cpp:285-308), pinned on
the CPU.  Its answer for a line must equal "a scan of the line as a text has
a candidate" (scan is pinned by its own tests), lines shorter than W are
false, and -- unlike the text scan of the split method, which restarts at
every newline (ffbf.cpp:342-358) -- a line is hashed straight through a '\\n'
inside it."""
import numpy as np

import ffbf_amd as fa
from oracle import Oracle, SplitOracle


def _lines(rng, alpha, n, lmax, pats=None):
    out = []
    for i in range(n):
        L = int(rng.integers(0, lmax))
        t = fa.gen_bytes(alpha, int(rng.integers(1, 1 << 30)), L)
        if pats is not None and i % 3 == 0 and L > 80:
            b, off = pats
            j = int(rng.integers(0, off.size - 1))
            p = b[int(off[j]):int(off[j + 1])]
            at = int(rng.integers(0, L - p.size + 1)) if L >= p.size else 0
            t[at:at + p.size] = p[:L - at]
        out.append(t)
    return out


def _pack(lines):
    off = np.zeros(len(lines) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([l.size for l in lines])
    b = np.concatenate(lines) if lines else np.zeros(0, dtype=np.uint8)
    return np.ascontiguousarray(b, dtype=np.uint8), off


def test_pair_query_equals_scan_has_candidate():
    rng = np.random.default_rng(5)
    W, k, log2 = 19, 4, 14  # a small filter: false-positive windows occur too
    sbox = fa.sbox_from_seed(5)
    pb, po = fa.gen_patterns(fa.ALPHA_ASCII, 51, 600, 19, 40)
    o = Oracle(W, k, log2, sbox)
    F = o.build(pb, po)
    lines = _lines(rng, fa.ALPHA_ASCII, 300, 400, (pb, po))
    b, off = _pack(lines)
    hits = o.query_lines(b, off, F)
    want = []
    for l in lines:
        _, cand = o.scan(l, F)
        want.append(1 if cand.size else 0)
    assert hits.tolist() == want
    assert 20 < sum(want) < 300
    # shorter than W: false even when F is all ones
    full = np.full_like(F, 0xFFFFFFFF)
    sb, soff = _pack([np.frombuffer(b"x" * n, dtype=np.uint8).copy() for n in (0, 1, 18, 19, 20)])
    assert o.query_lines(sb, soff, full).tolist() == [0, 0, 0, 1, 1]
    # a line holding a pattern is true
    p = pb[int(po[7]):int(po[8])]
    pl = np.concatenate([np.frombuffer(b"zz", dtype=np.uint8), p, np.frombuffer(b"z", dtype=np.uint8)])
    qb, qoff = _pack([pl])
    assert o.query_lines(qb, qoff, F).tolist() == [1]


def test_split_query_equals_scan_on_newline_free_lines_and_hashes_through_newlines():
    rng = np.random.default_rng(6)
    sb64 = np.random.default_rng(61).integers(0, 2**64, size=256, dtype=np.uint64)
    W, k, s, c, e = 19, 5, 2, 12, 16
    pb, po = fa.gen_patterns(fa.ALPHA_ASCII, 62, 500, 19, 40)
    o = SplitOracle(W, k, s, c, e, sb64)
    F = o.build(pb, po)
    lines = _lines(rng, fa.ALPHA_ASCII, 300, 300, (pb, po))
    for l in lines:
        l[l >= 0x80] = 0x41
    b, off = _pack(lines)
    hits = o.query_lines(b, off, F)
    want = [1 if o.scan(l, F)[1].size else 0 for l in lines]
    assert hits.tolist() == want and sum(want) > 10
    # a pattern whose first W bytes hold a newline: operator<< hashes it
    # through the newline, so does operator[]; the text scan never tests
    # such a window
    pat = bytearray(b"abcdefghij\nklmnopqrstuv")
    nb, noff = _pack([np.frombuffer(bytes(pat), dtype=np.uint8).copy()])
    o2 = SplitOracle(W, k, s, c, e, sb64)
    F2 = o2.build(nb, noff)
    line = np.frombuffer(b"--" + bytes(pat) + b"--", dtype=np.uint8).copy()
    qb, qoff = _pack([line])
    assert o2.query_lines(qb, qoff, F2).tolist() == [1]
    assert o2.scan(line, F2)[1].size == 0
    # high bytes hash as 0: the same line with 0x00 and with 0xC3 agree
    l0 = np.frombuffer(b"q" * 30, dtype=np.uint8).copy()
    l0[5] = 0x00
    l1 = l0.copy()
    l1[5] = 0xC3
    hb, hoff = _pack([l0, l1])
    F3 = o2.build(*_pack([l0[:W]]))
    assert o2.query_lines(hb, hoff, F3).tolist() == [1, 1]
