"""Writes the golden fixtures under tests/golden/ (run once; outputs are committed).

* ref_sbox.json -- the 256-entry uint32 substitution table the reference ships
  as data in src/gpgpu/sbox.h:2-67 (identical to the tables in
  src/hashes/hash_rot_sbox_2.h:56-121).  Copyright 2011 Carnegie Mellon
  University, Apache License 2.0.  Read from the header as data; no code.
* hash_vectors.json -- hash pairs and forward/reverse bit indices of fixed
  windows, computed here from the closed form
      h1 = XOR_j rotl(S[w_j], (W-1-j) mod 32),  h2 = XOR_j rotl(S[w_j], 3(W-1-j) mod 32)
  (greptest.cu:106-114 unrolled), which is a different formulation from the
  oracle's loop and rolling forms.
* e2e_small.json -- a small end-to-end case (patterns incl. short ones,
  duplicates and shared prefixes; a text with planted copies) whose filter
  bits, candidates, survivors and matches are derived by brute-force set
  semantics in pure Python.

* ref_sbox64.json -- the 256-entry uint64 substitution table of the CPU
  binary's hash, data in src/hashes/hash_rot_sbox_pre_2.h:55-120 (same
  copyright and licence).  Read from the header as data; no code.
* split_vectors.json -- hashes and bit indices of the CPU binary's split
  filter for fixed windows, from the closed form
      h = XOR_j rol64(S64[c_j], (W-1-j) mod 64),  c_j = 0 where the byte is >= 0x80,
      h1 = low 32 bits, h2 = high 32 bits
  (the ring-buffer update of hash_rot_sbox_pre_2.h:41-50 unrolled), placed
  in the cache part for the first s functions and the external part for the
  rest (ffbf.cpp:135-185).
* split_e2e_small.json -- small end-to-end cases of the split method over
  texts with newlines and bytes >= 0x80, derived by brute-force set
  semantics: a window is tested only if it holds no newline (ffbf.cpp:340-358).

Usage: python tests/golden/make_golden.py [path/to/reference]
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
M32 = 0xFFFFFFFF


def parse_sbox(ref_root: str) -> list:
    text = open(os.path.join(ref_root, "src/gpgpu/sbox.h")).read()
    vals = [int(x, 16) for x in re.findall(r"0x[0-9A-Fa-f]{8}", text)]
    assert len(vals) == 256, len(vals)
    return vals


def rotl(x: int, r: int) -> int:
    r %= 32
    return ((x << r) | (x >> (32 - r))) & M32 if r else x


def closed_form(S: list, win: bytes):
    W = len(win)
    h1 = h2 = 0
    for j, c in enumerate(win):
        h1 ^= rotl(S[c], (W - 1 - j) % 32)
        h2 ^= rotl(S[c], (3 * (W - 1 - j)) % 32)
    return h1, h2


def fwd(h1: int, h2: int, k: int, log2: int) -> list:
    m = (1 << log2) - 1
    f = [h1, h2, (h1 + h2) & M32, (h1 + 5 * h2) & M32, ((h1 << 16) | (h2 >> 16)) & M32]
    return [x & m for x in f[:k]]


def rev(h1: int, h2: int, k: int, log2: int) -> list:
    m = (1 << log2) - 1
    r = [(h1 + 7 * h2) & M32, (h1 + 3 * h2) & M32, ((h1 << 1) + h2) & M32, ((h1 << 2) + h2) & M32,
         (h1 * 11 + h2) & M32]
    return [x & m for x in r[:k]]


def lcg_bytes(seed: int, n: int, alphabet: bytes) -> bytes:
    # tiny self-contained generator for fixture inputs (Knuth MMIX LCG)
    out = bytearray()
    x = seed & 0xFFFFFFFFFFFFFFFF
    for _ in range(n):
        x = (x * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        out.append(alphabet[(x >> 33) % len(alphabet)])
    return bytes(out)


def e2e_case(S: list, W: int, k: int, log2: int, patterns: list, text: bytes) -> dict:
    F = set()
    for p in patterns:
        if len(p) >= W:
            F.update(fwd(*closed_form(S, p[:W]), k, log2))
    cands = []
    R = set()
    for o in range(len(text) - W + 1):
        h = closed_form(S, text[o:o + W])
        if all(b in F for b in fwd(*h, k, log2)):
            cands.append(o)
            R.update(rev(*h, k, log2))
    surv = [i for i, p in enumerate(patterns)
            if len(p) >= W and all(b in R for b in rev(*closed_form(S, p[:W]), k, log2))]
    cset = set(cands)
    matches = sorted((i, o) for i in surv for o in cset
                     if text[o:o + len(patterns[i])] == patterns[i])
    # every occurrence of every long-enough pattern, found without hashing
    occ = sorted((i, o) for i, p in enumerate(patterns) if len(p) >= W
                 for o in range(len(text) - len(p) + 1) if text[o:o + len(p)] == p)
    assert matches == occ, "no-false-negative invariant violated in fixture"
    return {
        "window": W, "k": k, "log2_bits": log2,
        "patterns": [p.hex() for p in patterns],
        "text": text.hex(),
        "forward_bits": sorted(F),
        "reverse_bits": sorted(R),
        "candidates": cands,
        "survivors": surv,
        "matches": matches,
    }


M64 = (1 << 64) - 1


def parse_sbox64(ref_root: str) -> list:
    text = open(os.path.join(ref_root, "src/hashes/hash_rot_sbox_pre_2.h")).read()
    body = text[text.index("sbox[] = {"):]
    body = body[:body.index("};")]
    vals = [int(x, 16) for x in re.findall(r"0x[0-9A-Fa-f]+", body)]
    assert len(vals) == 256, len(vals)
    return vals


def rol64(x: int, r: int) -> int:
    r %= 64
    return ((x << r) | (x >> (64 - r))) & M64 if r else x


def closed_form64(S64: list, win: bytes):
    W = len(win)
    h = 0
    for j, c in enumerate(win):
        h ^= rol64(S64[0 if c >= 0x80 else c], (W - 1 - j) % 64)
    return h & M32, h >> 32


def rot16(x: int) -> int:
    return ((x << 16) | (x >> 16)) & M32


def split_funcs(h1: int, h2: int):
    f = [h1, h2, rot16(h1), rot16(h2), h1 + h2, h1 + (h2 << 1), h1 + (h2 << 2), h1 + 3 * h2, h1 + 5 * h2,
         h1 + 6 * h2, h1 + 7 * h2]
    r = [h1 + (h2 << 1), h1 + (h2 << 2), h1 + 3 * h2, h1 + 5 * h2, h1 + 6 * h2, h1 + 7 * h2, h1, h2, rot16(h1),
         rot16(h2), h1 + h2]
    return [x & M32 for x in f], [x & M32 for x in r]


def split_place(vals: list, k: int, s: int, c: int, e: int) -> list:
    return [(x & ((1 << c) - 1)) if i < s else (x & ((1 << e) - 1)) + (1 << c) for i, x in enumerate(vals[:k])]


def split_e2e_case(S64: list, W: int, k: int, s: int, c: int, e: int, patterns: list, text: bytes) -> dict:
    def fbits(h):
        return split_place(split_funcs(*h)[0], k, s, c, e)

    def rbits(h):
        return split_place(split_funcs(*h)[1], k, s, c, e)

    F = set()
    for p in patterns:
        if len(p) >= W:
            F.update(fbits(closed_form64(S64, p[:W])))
    cands, R = [], set()
    for o in range(len(text) - W + 1):
        win = text[o:o + W]
        if b"\n" in win:
            continue
        h = closed_form64(S64, win)
        if all(b in F for b in fbits(h)):
            cands.append(o)
            R.update(rbits(h))
    surv = [i for i, p in enumerate(patterns)
            if len(p) >= W and all(b in R for b in rbits(closed_form64(S64, p[:W])))]
    cset = set(cands)
    matches = sorted((i, o) for i in surv for o in cset if text[o:o + len(patterns[i])] == patterns[i])
    # every occurrence of every long-enough pattern whose prefix window holds
    # no newline (such a window is never tested), found without hashing
    occ = sorted((i, o) for i, p in enumerate(patterns) if len(p) >= W and b"\n" not in p[:W]
                 for o in range(len(text) - len(p) + 1) if text[o:o + len(p)] == p)
    assert matches == occ, "no-false-negative invariant violated in split fixture"
    return {
        "window": W, "k": k, "s": s, "log2_cache_bits": c, "log2_ext_bits": e,
        "patterns": [p.hex() for p in patterns],
        "text": text.hex(),
        "forward_bits": sorted(F),
        "reverse_bits": sorted(R),
        "candidates": cands,
        "survivors": surv,
        "matches": matches,
    }


def split_main(ref: str) -> None:
    S64 = parse_sbox64(ref)
    with open(os.path.join(HERE, "ref_sbox64.json"), "w") as f:
        json.dump({"source": "efficient/ffbf src/hashes/hash_rot_sbox_pre_2.h:55-120 (Copyright 2011 Carnegie "
                             "Mellon University, Apache-2.0)", "sbox": [str(x) for x in S64]}, f, indent=1)
    windows = [
        b"A" * 19, b"The quick brown fox", b"feed-forward bloom filter"[:19], bytes(range(19)),
        bytes(range(200, 219)), b"ACGT" * 16, b"x" * 64, b"A", b"\x80", b"\xff" * 19,
        lcg_bytes(21, 19, bytes(range(32, 127))), lcg_bytes(22, 33, bytes(range(256))),
        lcg_bytes(23, 64, bytes(range(256))), lcg_bytes(24, 24, b"ACGT"),
    ]
    shapes = [(5, 2, 24, 28), (11, 3, 10, 12), (4, 0, 16, 20), (3, 3, 20, 8)]
    vecs = []
    for w in windows:
        h1, h2 = closed_form64(S64, w)
        fv, rv = split_funcs(h1, h2)
        vecs.append({"window": w.hex(), "h1": h1, "h2": h2,
                     "fwd": {"%d,%d,%d,%d" % sh: split_place(fv, *sh) for sh in shapes},
                     "rev": {"%d,%d,%d,%d" % sh: split_place(rv, *sh) for sh in shapes}})
    with open(os.path.join(HERE, "split_vectors.json"), "w") as f:
        json.dump({"sbox": "ref_sbox64.json", "shapes": "k,s,log2_cache_bits,log2_ext_bits", "vectors": vecs}, f,
                  indent=1)

    # texts of short lines with some high bytes; patterns include one with a
    # newline inside its prefix window (never a candidate), one with a newline
    # after it (matches across a line break), a short one and a duplicate
    alpha = bytes(range(32, 127)) + b"\xc3\xa9\x80\xff"
    pats = [lcg_bytes(300 + i, 19 + i % 13, alpha) for i in range(30)]
    pats += [pats[2], b"0123456789\n0123456789abc", b"0123456789abcdefghij\nklm", b"short", pats[4][:19]]
    text = bytearray(lcg_bytes(777, 6000, alpha))
    for i in range(0, len(text), 37):
        text[i] = 0x0A
    for pi, o in [(0, 5), (2, 800), (4, 1500), (7, 2600), (31, 3000), (32, 3500), (33, 4000), (9, 5980 - 40)]:
        p = pats[pi]
        o = min(o, len(text) - len(p))
        text[o:o + len(p)] = p
    # one plant straddling a line break inside its first 19 bytes: not found
    text[4500:4500 + len(pats[11])] = pats[11]
    text[4505] = 0x0A
    # a copy of a pattern whose bytes >= 0x80 are swapped for other such
    # bytes: the same hash (they all hash as 0), so a candidate, but no match
    hp = next(p for p in pats if len(p) >= 19 and any(x >= 0x80 for x in p[:12]))
    text[5200:5200 + len(hp)] = bytes((0xFF if x == 0x80 else 0x80) if x >= 0x80 else x for x in hp)
    cases = [split_e2e_case(S64, 19, 5, 2, 12, 14, pats, bytes(text)),
             split_e2e_case(S64, 19, 11, 3, 10, 12, pats, bytes(text)),
             split_e2e_case(S64, 12, 3, 0, 8, 9, pats, bytes(text))]
    with open(os.path.join(HERE, "split_e2e_small.json"), "w") as f:
        json.dump({"sbox": "ref_sbox64.json", "cases": cases}, f)


def main() -> None:
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    S = parse_sbox(ref)
    with open(os.path.join(HERE, "ref_sbox.json"), "w") as f:
        json.dump({"source": "efficient/ffbf src/gpgpu/sbox.h:2-67 (Copyright 2011 Carnegie Mellon University, "
                             "Apache-2.0)", "sbox": S}, f, indent=1)

    windows = [
        b"A" * 32, b"C" * 32, b"ACGT" * 8, b"A", b"C", b"G", b"T", b"AC", b"TTAGGGTTAGGGTTAGGGT",
        b"The quick brown fox", b"feed-forward bloom filter", bytes(range(32)), bytes(range(224, 256)) + b"xyz",
        lcg_bytes(7, 19, b"ACGT"), lcg_bytes(8, 24, b"ACGT"), lcg_bytes(9, 32, b"ACGT"),
        lcg_bytes(10, 64, bytes(range(32, 127))), lcg_bytes(11, 33, bytes(range(256))),
    ]
    vecs = []
    for w in windows:
        h1, h2 = closed_form(S, w)
        vecs.append({"window": w.hex(), "h1": h1, "h2": h2,
                     "fwd": {str(l): fwd(h1, h2, 5, l) for l in (10, 24, 26, 28, 32)},
                     "rev": {str(l): rev(h1, h2, 5, l) for l in (10, 24, 26, 28, 32)}})
    with open(os.path.join(HERE, "hash_vectors.json"), "w") as f:
        json.dump({"sbox": "ref_sbox.json", "vectors": vecs}, f, indent=1)

    # end-to-end cases
    cases = []
    alpha = b"ACGT"
    pats = [lcg_bytes(100 + i, 12 + (i * 7) % 20, alpha) for i in range(40)]
    pats += [pats[3], pats[5][:10] + b"TTTTTTTTTT", b"ACG", b"", pats[7][:12]]  # duplicate, shared prefix, short, empty
    text = bytearray(lcg_bytes(555, 3000, alpha))
    for j, (pi, o) in enumerate([(0, 10), (3, 500), (5, 1200), (7, 2900 - 40), (41, 2000), (9, 0), (12, 2970)]):
        p = pats[pi]
        o = min(o, len(text) - len(p))
        text[o:o + len(p)] = p
    cases.append(e2e_case(S, 10, 4, 12, pats, bytes(text)))
    # small filter, many false positives: exercises R pruning and exact verification
    cases.append(e2e_case(S, 6, 2, 10, pats, bytes(text)))
    # printable ASCII, k=5 (both commented-out fifth functions)
    apats = [lcg_bytes(900 + i, 19 + i % 9, bytes(range(32, 127))) for i in range(25)]
    atext = bytearray(lcg_bytes(901, 5000, bytes(range(32, 127))))
    for pi, o in [(0, 0), (4, 777), (9, 4500), (24, 5000 - 27)]:
        p = apats[pi]
        o = min(o, len(atext) - len(p))
        atext[o:o + len(p)] = p
    cases.append(e2e_case(S, 19, 5, 14, apats, bytes(atext)))
    with open(os.path.join(HERE, "e2e_small.json"), "w") as f:
        json.dump({"sbox": "ref_sbox.json", "cases": cases}, f)
    split_main(ref)


if __name__ == "__main__":
    main()
