"""Pins the oracle's split method (the CPU binary's filter,
src/rabin-karp/ffbf.cpp with src/hashes/hash_rot_sbox_pre_2.h) before it is
trusted as the checker: golden vectors from the closed form over the
reference's 64-bit sbox table, identities the mathematics fixes, hand values
of the index families, brute-force line semantics on small texts, and planted
mistakes the fixtures must reject.  CPU only."""
import random

import numpy as np
import pytest

from conftest import bits_of, load_golden
from oracle import SplitOracle, brute_matches, filter_save_deltas_bits

M32 = 0xFFFFFFFF


def test_reference_sbox64_fixture(ref_sbox64):
    # entries quoted from src/hashes/hash_rot_sbox_pre_2.h:56-120, including
    # ones written with fewer than 16 hex digits
    assert ref_sbox64.shape == (256,)
    assert int(ref_sbox64[0]) == 0x198C9C06F9ED325B
    assert int(ref_sbox64[14]) == 0x2BFA5B92EF9F5D5
    assert int(ref_sbox64[16]) == 0x969BCBEAE5DDB10
    assert int(ref_sbox64[255]) == 0x6D943F9CBE2423C6
    assert len(set(int(x) for x in ref_sbox64)) == 256


def test_hash_and_indices_match_closed_form_vectors(ref_sbox64):
    d = load_golden("split_vectors.json")
    for v in d["vectors"]:
        w = bytes.fromhex(v["window"])
        for shape, fwd in v["fwd"].items():
            k, s, c, e = (int(x) for x in shape.split(","))
            o = SplitOracle(len(w), k, s, c, e, ref_sbox64)
            h = o.hash_window(w)
            assert h == (v["h1"], v["h2"]), v["window"]
            assert o.fwd_indices(*h) == fwd, (v["window"], shape)
            assert o.rev_indices(*h) == v["rev"][shape], (v["window"], shape)


def test_all_rotations_identity(ref_sbox64):
    # W=64 copies of one byte c XOR all 64 rotations of S[c]: all-ones times
    # parity(S[c]); bytes >= 0x80 hash as byte 0
    o = SplitOracle(64, 5, 2, 24, 28, ref_sbox64)
    for c in range(256):
        sc = int(ref_sbox64[0 if c >= 0x80 else c])
        expect = M32 if bin(sc).count("1") % 2 else 0
        assert o.hash_window(bytes([c]) * 64) == (expect, expect), c


def test_one_byte_window_is_the_sbox_entry(ref_sbox64):
    o = SplitOracle(1, 5, 2, 24, 28, ref_sbox64)
    for c in (0, 10, 65, 127, 128, 200, 255):
        sc = int(ref_sbox64[0 if c >= 0x80 else c])
        assert o.hash_window(bytes([c])) == (sc & M32, sc >> 32)
    # two bytes: rol64(S[a], 1) ^ S[b]
    o2 = SplitOracle(2, 5, 2, 24, 28, ref_sbox64)
    a, b = int(ref_sbox64[ord("x")]), int(ref_sbox64[ord("y")])
    h = (((a << 1) | (a >> 63)) & ((1 << 64) - 1)) ^ b
    assert o2.hash_window(b"xy") == (h & M32, h >> 32)


def test_high_bytes_hash_as_zero(ref_sbox64):
    rng = random.Random(4)
    o = SplitOracle(19, 5, 2, 24, 28, ref_sbox64)
    for _ in range(50):
        w = bytearray(rng.randrange(256) for _ in range(19))
        z = bytes(0 if x >= 0x80 else x for x in w)
        assert o.hash_window(bytes(w)) == o.hash_window(z)
    assert o.hash_window(b"a" * 18 + b"\x80") != o.hash_window(b"a" * 18 + b"\x7f")


def test_index_families_by_hand(ref_sbox64):
    # getHashValue / getHashValue_2 (ffbf.cpp:135-185) evaluated by hand at
    # unit hash halves, before placement (cache part 2^32 bits, s = k)
    o = SplitOracle(19, 11, 11, 31, 31, ref_sbox64)
    # placement with c = 31 keeps x & (2^31 - 1); inputs below stay unchanged
    assert o.fwd_indices(1, 0) == [1, 0, 0x10000, 0, 1, 1, 1, 1, 1, 1, 1]
    assert o.fwd_indices(0, 1) == [0, 1, 0, 0x10000, 1, 2, 4, 3, 5, 6, 7]
    assert o.rev_indices(1, 0) == [1, 1, 1, 1, 1, 1, 1, 0, 0x10000, 0, 1]
    assert o.rev_indices(0, 1) == [2, 4, 3, 5, 6, 7, 0, 1, 0, 0x10000, 1]
    # placement: the first s functions in [0, 2^c), the rest in [2^c, 2^c + 2^e)
    p = SplitOracle(19, 5, 2, 24, 28, ref_sbox64)
    ones = p.fwd_indices(M32, M32)
    # h1 = h2 = 2^32-1: h1, h2 -> cache 2^24-1; rot16 -> ext 2^28-1 + 2^24;
    # h1 + h2 = 2^32 - 2 -> ext (2^28 - 2) + 2^24
    assert ones == [(1 << 24) - 1, (1 << 24) - 1, (1 << 28) - 1 + (1 << 24), (1 << 28) - 1 + (1 << 24),
                    (1 << 28) - 2 + (1 << 24)]
    p0 = SplitOracle(19, 3, 0, 10, 12, ref_sbox64)
    assert p0.fwd_indices(5, 0) == [5 + 1024, 0 + 1024, 0x50000 % 4096 + 1024]
    assert p0.rev_indices(5, 0) == [5 + 1024] * 3


def _case_arrays(case):
    pats = [bytes.fromhex(p) for p in case["patterns"]]
    b = np.frombuffer(b"".join(pats), dtype=np.uint8).copy()
    off = np.zeros(len(pats) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(p) for p in pats])
    text = np.frombuffer(bytes.fromhex(case["text"]), dtype=np.uint8).copy()
    return pats, b, off, text


def _oracle(case, sbox):
    return SplitOracle(case["window"], case["k"], case["s"], case["log2_cache_bits"], case["log2_ext_bits"], sbox)


@pytest.mark.parametrize("ci", [0, 1, 2])
def test_e2e_golden(ref_sbox64, ci):
    case = load_golden("split_e2e_small.json")["cases"][ci]
    pats, b, off, text = _case_arrays(case)
    o = _oracle(case, ref_sbox64)
    r = o.run(text, b, off)
    assert bits_of(r["F"]) == case["forward_bits"]
    assert r["cand"].tolist() == case["candidates"]
    assert bits_of(r["R"]) == case["reverse_bits"]
    assert [i for i in range(len(pats)) if (r["surv"][i >> 5] >> (i & 31)) & 1] == case["survivors"]
    assert list(zip(r["ids"].tolist(), r["offs"].tolist())) == [tuple(m) for m in case["matches"]]
    # threaded scan: same candidates and R
    R4, c4 = o.scan(text, r["F"], threads=4)
    assert np.array_equal(R4, r["R"]) and np.array_equal(c4, r["cand"])


def test_scan_line_semantics_brute_force(ref_sbox64):
    # random F, text of short lines with high bytes: candidates are exactly
    # the newline-free windows whose k bits are set
    rng = np.random.default_rng(12)
    o = SplitOracle(9, 4, 1, 9, 10, ref_sbox64)
    nw = o.n_words
    F = rng.integers(0, 2**32, size=nw, dtype=np.uint32) | rng.integers(0, 2**32, size=nw, dtype=np.uint32)
    text = rng.integers(0, 256, size=4000, dtype=np.uint8)
    text[rng.integers(0, 4000, size=300)] = 10
    R, cand = o.scan(text, F)
    Fb = set(bits_of(F))
    exp_c, exp_r = [], set()
    for i in range(text.size - 9 + 1):
        w = text[i:i + 9].tobytes()
        if b"\n" in w:
            continue
        h = o.hash_window(w)
        if all(x in Fb for x in o.fwd_indices(*h)):
            exp_c.append(i)
            exp_r.update(o.rev_indices(*h))
    assert len(exp_c) > 50
    assert cand.tolist() == exp_c
    assert bits_of(R) == sorted(exp_r)
    R3, c3 = o.scan(text, F, threads=3)
    assert np.array_equal(R3, R) and np.array_equal(c3, cand)


def test_verify_equals_brute_force(ref_sbox64):
    rng = random.Random(8)
    alpha = b"ACGT\n"
    pats = [bytes(rng.choice(b"ACGT") for _ in range(rng.randint(10, 30))) for _ in range(150)]
    pats += [b"ACGTACGTAC\nACGT", b"ACGTACGTACGTA\nCG"]
    text = bytearray(rng.choice(alpha) for _ in range(30000))
    for _ in range(80):
        p = rng.choice(pats)
        o_ = rng.randint(0, len(text) - len(p))
        text[o_:o_ + len(p)] = p
    text = np.frombuffer(bytes(text), dtype=np.uint8).copy()
    b = np.frombuffer(b"".join(pats), dtype=np.uint8).copy()
    off = np.zeros(len(pats) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(p) for p in pats])
    for (W, k, s, c, e) in [(10, 5, 2, 14, 16), (10, 2, 1, 6, 6), (8, 11, 4, 12, 12)]:
        o = SplitOracle(W, k, s, c, e, ref_sbox64)
        r = o.run(text, b, off)
        bi, bo = brute_matches(text, b, off, W)
        keep = [j for j in range(bi.size) if b"\n" not in pats[int(bi[j])][:W]]
        assert r["ids"].tolist() == bi[keep].tolist() and r["offs"].tolist() == bo[keep].tolist()
        assert len(keep) >= 60


def test_fixtures_reject_plausible_mistakes(ref_sbox64):
    """Mistakes planted in a Python copy of the method (never in the
    fixtures) must disagree with the committed fixtures: a hash without the
    newline restart, a 32-bit rotation, unfolded high bytes, swapped halves,
    and the external part placed at 0."""
    case = load_golden("split_e2e_small.json")["cases"][0]
    pats, _, _, text = _case_arrays(case)
    S = [int(x) for x in ref_sbox64]
    W, k, s, c, e = case["window"], case["k"], case["s"], case["log2_cache_bits"], case["log2_ext_bits"]
    t = bytes(text)
    M64 = (1 << 64) - 1

    def rol(x, r, bits=64):
        r %= bits
        m = (1 << bits) - 1
        return ((x << r) | (x >> (bits - r))) & m if r else x

    def h64(win, fold=True, bits=64):
        h = 0
        for j, ch in enumerate(win):
            h ^= rol(S[0 if (fold and ch >= 0x80) else ch] & ((1 << bits) - 1) if bits < 64 else
                     S[0 if (fold and ch >= 0x80) else ch], (W - 1 - j), bits)
        return h & M64

    def cands(hash_fn, skip_newline=True, swap=False, ext_base=True):
        def place(vals):
            return [(x & ((1 << c) - 1)) if i < s else (x & ((1 << e) - 1)) + ((1 << c) if ext_base else 0)
                    for i, x in enumerate(vals[:k])]

        def halves(h):
            return (h >> 32, h & M32) if swap else (h & M32, h >> 32)

        def fwd(h):
            h1, h2 = halves(h)
            f = [h1, h2, rol(h1, 16, 32), rol(h2, 16, 32), (h1 + h2) & M32]
            return place(f)

        def rev(h):
            h1, h2 = halves(h)
            return place([(h1 + m * h2) & M32 for m in (2, 4, 3, 5, 6)])

        F = set()
        for p in pats:
            if len(p) >= W:
                F.update(fwd(hash_fn(p[:W])))
        out, R = [], set()
        for o in range(len(t) - W + 1):
            win = t[o:o + W]
            if skip_newline and b"\n" in win:
                continue
            h = hash_fn(win)
            if all(x in F for x in fwd(h)):
                out.append(o)
                R.update(rev(h))
        return sorted(F), out, sorted(R)

    want = (case["forward_bits"], case["candidates"], case["reverse_bits"])
    assert cands(h64) == want
    assert cands(h64, skip_newline=False) != want
    assert cands(lambda w: h64(w, fold=False)) != want
    assert cands(lambda w: h64(w, bits=32)) != want
    assert cands(h64, swap=True) != want
    assert cands(h64, ext_base=False) != want
    # the fold alone: same filter bits, one candidate more or less
    assert cands(lambda w: h64(w, fold=False))[1] != case["candidates"]


def test_split_filter_delta_format(ref_sbox64):
    # the split filter's 2^c + 2^e bits (not a power of two) in the delta format
    o = SplitOracle(19, 5, 2, 10, 12, ref_sbox64)
    F = np.zeros(o.n_words, dtype=np.uint32)
    for bit in (0, 1023, 1024, 1024 + 4095):
        F[bit >> 5] |= np.uint32(1 << (bit & 31))
    assert filter_save_deltas_bits(F, o.n_bits).tolist() == [0, 1023, 1, 4095]
