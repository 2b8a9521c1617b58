"""Pins the CPU oracle (oracle/ffbf_oracle.c) before it is trusted as the
checker: golden vectors from the closed form, identities the mathematics
fixes, brute force on small inputs, and planted mistakes that the fixtures
must reject.  CPU only."""
import random

import numpy as np
import pytest

from conftest import bits_of, load_golden, words_of
from oracle import Oracle, brute_matches, candidate_lines, filter_load_deltas, filter_save_deltas, survivor_files


def rotl(x, r):
    r %= 32
    return ((x << r) | (x >> (32 - r))) & 0xFFFFFFFF if r else x


def test_reference_sbox_fixture(ref_sbox):
    # values quoted from src/gpgpu/sbox.h:19-24 for the DNA letters
    assert ref_sbox.shape == (256,)
    assert ref_sbox[ord("A")] == 0x504172DE
    assert ref_sbox[ord("C")] == 0x13AC3CF6
    assert ref_sbox[ord("G")] == 0x21786B7B
    assert ref_sbox[ord("T")] == 0x7C3E73A6
    assert ref_sbox[0] == 0xF53E1837 and ref_sbox[255] != 0


def test_hash_window_matches_closed_form_vectors(ref_sbox):
    for v in load_golden("hash_vectors.json")["vectors"]:
        w = bytes.fromhex(v["window"])
        o = Oracle(len(w), 4, 28, ref_sbox)
        assert o.hash_window(w) == (v["h1"], v["h2"]), v["window"]


def test_all_rotations_identity(ref_sbox):
    # W=32 copies of one byte c: h1 and h2 XOR all 32 rotations of S[c]
    # (3 is odd, so 3j mod 32 also visits every rotation): all-ones times parity(S[c]).
    o = Oracle(32, 4, 28, ref_sbox)
    for c in range(256):
        expect = 0xFFFFFFFF if bin(int(ref_sbox[c])).count("1") % 2 else 0
        assert o.hash_window(bytes([c]) * 32) == (expect, expect)
    # parity(S['A']) = 0 -> the all-A window hashes to zero
    assert o.hash_window(b"A" * 32) == (0, 0)


def test_short_windows_by_hand(ref_sbox):
    S = [int(x) for x in ref_sbox]
    o1 = Oracle(1, 4, 28, ref_sbox)
    for c in (0, 65, 200, 255):
        assert o1.hash_window(bytes([c])) == (S[c], S[c])
    o2 = Oracle(2, 4, 28, ref_sbox)
    a, b = ord("x"), ord("y")
    assert o2.hash_window(b"xy") == (rotl(S[a], 1) ^ S[b], rotl(S[a], 3) ^ S[b])


@pytest.mark.parametrize("W", [1, 2, 5, 19, 24, 31, 32, 33, 40, 64])
def test_rolling_equals_window(ref_sbox, W):
    rng = np.random.default_rng(W)
    text = rng.integers(0, 256, size=700, dtype=np.uint8)
    o = Oracle(W, 4, 20, ref_sbox)
    h1, h2 = o.hash_rolling(text)
    assert h1.size == text.size - W + 1
    for i in range(0, h1.size, 7):
        assert (int(h1[i]), int(h2[i])) == o.hash_window(text[i:i + W].tobytes())


def test_hash_depends_only_on_window(ref_sbox):
    # src/hashes/validate.rb:23-33: equal last-W bytes give equal hashes,
    # whatever precedes them in the rolled stream.
    W = 19
    o = Oracle(W, 4, 20, ref_sbox)
    tail = b"feed-forward filter"
    for prefix in (b"", b"x", b"some other prefix bytes" * 3, bytes(range(256))):
        h1, h2 = o.hash_rolling(prefix + tail)
        assert (int(h1[-1]), int(h2[-1])) == o.hash_window(tail)


def test_index_families_vectors(ref_sbox):
    for v in load_golden("hash_vectors.json")["vectors"]:
        for log2, expect in v["fwd"].items():
            o = Oracle(4, 5, int(log2), ref_sbox)
            assert o.fwd_indices(v["h1"], v["h2"]) == expect
            assert o.rev_indices(v["h1"], v["h2"]) == v["rev"][log2]
        # k<5 takes the first k functions, in order
        o3 = Oracle(4, 3, 28, ref_sbox)
        assert o3.fwd_indices(v["h1"], v["h2"]) == v["fwd"]["28"][:3]


def _case_arrays(case):
    pats = [bytes.fromhex(p) for p in case["patterns"]]
    off = np.zeros(len(pats) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(p) for p in pats])
    b = np.frombuffer(b"".join(pats), dtype=np.uint8).copy()
    text = np.frombuffer(bytes.fromhex(case["text"]), dtype=np.uint8).copy()
    return pats, b, off, text


@pytest.mark.parametrize("ci", [0, 1, 2])
def test_e2e_golden(ref_sbox, ci):
    case = load_golden("e2e_small.json")["cases"][ci]
    pats, b, off, text = _case_arrays(case)
    o = Oracle(case["window"], case["k"], case["log2_bits"], ref_sbox)
    r = o.run(text, b, off)
    assert bits_of(r["F"]) == case["forward_bits"]
    assert r["cand"].tolist() == case["candidates"]
    assert bits_of(r["R"]) == case["reverse_bits"]
    assert [i for i in range(len(pats)) if (r["surv"][i >> 5] >> (i & 31)) & 1] == case["survivors"]
    assert list(zip(r["ids"].tolist(), r["offs"].tolist())) == [tuple(m) for m in case["matches"]]


def test_build_is_order_independent_and_bounded(ref_sbox):
    rng = random.Random(3)
    pats = [bytes(rng.choice(b"ACGT") for _ in range(rng.randint(5, 40))) for _ in range(300)]
    o = Oracle(12, 4, 14, ref_sbox)

    def build(ps):
        b = np.frombuffer(b"".join(ps), dtype=np.uint8).copy()
        off = np.zeros(len(ps) + 1, dtype=np.uint64)
        off[1:] = np.cumsum([len(p) for p in ps])
        return o.build(b, off)

    F = build(pats)
    shuffled = pats[:]
    rng.shuffle(shuffled)
    assert np.array_equal(F, build(shuffled))
    n_long = sum(len(p) >= 12 for p in pats)
    assert 0 < len(bits_of(F)) <= 4 * n_long
    # short patterns contribute nothing
    assert np.array_equal(build([p for p in pats if len(p) >= 12]), F)


def test_scan_brute_force_semantics(ref_sbox):
    rng = np.random.default_rng(11)
    W, k, log2 = 7, 3, 11
    o = Oracle(W, k, log2, ref_sbox)
    F = rng.integers(0, 2**32, size=1 << (log2 - 5), dtype=np.uint32) & rng.integers(0, 2**32, size=1 << (log2 - 5), dtype=np.uint32)
    text = rng.integers(60, 70, size=3000, dtype=np.uint8)
    R, cand = o.scan(text, F)
    Fb = set(bits_of(F))
    exp_c, exp_r = [], set()
    for i in range(text.size - W + 1):
        h = o.hash_window(text[i:i + W].tobytes())
        if all(b in Fb for b in o.fwd_indices(*h)):
            exp_c.append(i)
            exp_r.update(o.rev_indices(*h))
    assert cand.tolist() == exp_c
    assert bits_of(R) == sorted(exp_r)
    R4, cand4 = o.scan(text, F, threads=4)
    assert np.array_equal(R4, R) and np.array_equal(cand4, cand)


def test_verify_equals_brute_force(ref_sbox):
    rng = random.Random(5)
    pats = [bytes(rng.choice(b"ACGT") for _ in range(rng.randint(8, 30))) for _ in range(200)]
    pats += [pats[0], pats[1][:9], b"AC"]
    text = bytearray(rng.choice(b"ACGT") for _ in range(20000))
    for _ in range(60):
        p = rng.choice(pats)
        o_ = rng.randint(0, len(text) - len(p))
        text[o_:o_ + len(p)] = p
    text = np.frombuffer(bytes(text), dtype=np.uint8).copy()
    b = np.frombuffer(b"".join(pats), dtype=np.uint8).copy()
    off = np.zeros(len(pats) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(p) for p in pats])
    for (W, k, log2) in [(8, 4, 16), (8, 2, 10), (5, 1, 10)]:
        o = Oracle(W, k, log2, [int(x) for x in ref_sbox])
        r = o.run(text, b, off)
        bi, bo = brute_matches(text, b, off, W)
        assert r["ids"].tolist() == bi.tolist() and r["offs"].tolist() == bo.tolist()
        assert bi.size >= 60


def test_filter_delta_format():
    F = words_of([0, 5, 37, 1023], 32)
    d = filter_save_deltas(F, 10)
    assert d.tolist() == [0, 5, 32, 986]  # ffbf.cpp:518-520: delta to the previous set index
    assert np.array_equal(filter_load_deltas(d, 10), F)
    rng = np.random.default_rng(1)
    G = rng.integers(0, 2**32, size=1 << 11, dtype=np.uint32)
    assert np.array_equal(filter_load_deltas(filter_save_deltas(G, 16), 16), G)


def test_fixtures_reject_plausible_mistakes(ref_sbox):
    """A hash with a plausible mistake (wrong rotation, dropped term, reversed
    window, wrong reverse function) must disagree with the committed vectors."""
    S = [int(x) for x in ref_sbox]
    vecs = load_golden("hash_vectors.json")["vectors"]

    def good(w):
        h1 = h2 = 0
        for c in w:
            h1, h2 = rotl(h1, 1) ^ S[c], rotl(h2, 3) ^ S[c]
        return h1, h2

    mutants = {
        "h2 rotates by 2": lambda w: (good(w)[0], _loop(S, w, 1, 2)[1]),
        "window reversed": lambda w: good(w[::-1]),
        "rotate right": lambda w: _loop(S, w, 31, 29),
        "last byte dropped": lambda w: good(w[:-1]) if len(w) > 1 else (0, 0),
    }
    for v in vecs:
        assert good(bytes.fromhex(v["window"])) == (v["h1"], v["h2"])
    for name, fn in mutants.items():
        assert any(fn(bytes.fromhex(v["window"])) != (v["h1"], v["h2"]) for v in vecs), name
    # reverse family with a swapped coefficient
    v = vecs[-1]
    bad = [(v["h1"] + 3 * v["h2"]) & 0xFFFFFFFF, (v["h1"] + 7 * v["h2"]) & 0xFFFFFFFF]
    assert [x & ((1 << 28) - 1) for x in bad] != v["rev"]["28"][:2]


def _loop(S, w, r1, r2):
    h1 = h2 = 0
    for c in w:
        h1, h2 = rotl(h1, r1) ^ S[c], rotl(h2, r2) ^ S[c]
    return h1, h2


def test_candidate_lines_hand_cases():
    # lines: "ab" [0,2), "\n" 2, "cd" [3,5), "\n" 5, "" (6 is '\n'), "efg" [7,10) without '\n'
    text = np.frombuffer(b"ab\ncd\n\nefg", dtype=np.uint8)
    assert candidate_lines(text, []) == b""
    assert candidate_lines(text, [0, 1]) == b"ab\n"           # two candidates, one line
    assert candidate_lines(text, [2]) == b""                  # on a line break
    assert candidate_lines(text, [6]) == b""                  # on an empty line's break
    assert candidate_lines(text, [4, 3, 9]) == b"cd\nefg\n"    # order-free; last line ends at the end
    assert candidate_lines(text, [1, 3, 8, 8]) == b"ab\ncd\nefg\n"
    # a text without any line break is one line
    assert candidate_lines(np.frombuffer(b"xyz", dtype=np.uint8), [2]) == b"xyz\n"


def test_survivor_files_hand_cases():
    b = np.frombuffer(b"alphabetagamma", dtype=np.uint8)
    off = np.array([0, 5, 9, 14], dtype=np.uint64)
    bits = np.array([0b101], dtype=np.uint32)
    pats, idx = survivor_files(b, off, bits)
    assert pats == b"alpha\ngamma\n" and idx == b"0\n2\n"
    assert survivor_files(b, off, np.array([0], dtype=np.uint32)) == (b"", b"")
