"""CPU oracle of the reference FFBF path -- TEST INFRASTRUCTURE ONLY.

Binds ``oracle/libffbf_oracle.so`` (built from ``ffbf_oracle.c``, a plain C
restatement of efficient/ffbf's GPU arithmetic and CPU flow; each C function
cites the reference lines it follows).  Only tests/, the smoke check and
bench.py's cpu_baseline leg may import this package, and only as the checker
or the timed CPU baseline, never as part of the product path.

Parity status of each function (how it is pinned; of the reference only its
rolling-hash headers build here, see oracle/ref.mk):
  hash_window       the reference's own hash_rot_sbox_2.h compiled unmodified
                    (oracle/_ref, every window of random texts, W 1..64),
                    closed form XOR of rotations, the W=32 all-rotations
                    parity identity, W=1/W=2 hand values from the reference sbox
  hash_rolling      equals hash_window on every window; tail-only dependence
                    (the reference's validate.rb invariant)
  fwd/rev_indices   hand-evaluated fixtures (golden JSON)
  build/scan/pattern_pass   brute-force Python set semantics on small inputs,
                    order independence, popcount bounds
  verify            independent naive matcher (brute_matches) and planted
                    occurrences
  query_lines       a line's answer equals "a scan of the line as a text has
                    a candidate" (scan is pinned above), lines shorter than W
                    are false, hand lines holding a pattern; split method:
                    the same over newline-free lines, and a line with a '\n'
                    inside is hashed through it (hand-computed)
  save/load deltas  round trip and hand-built vectors
  candidate_lines / survivor_files   hand-built texts and pattern sets
                    (newline runs, a line without '\n', a candidate on a line
                    break, several candidates in one line)
  split method (SplitOracle, the CPU binary's filter)
    hash            the reference's own hash_rot_sbox_pre_2.h compiled
                    unmodified (oracle/_ref, W 1..64, bytes below 0x80),
                    closed form XOR of 64-bit rotations (golden vectors from
                    the reference's 64-bit sbox table), the W=64 all-rotations
                    parity identity, W=1 = the sbox entry, 0x80..0xFF hashing
                    as 0x00
    indices         hand-evaluated values at (h1,h2) = (1,0), (0,1) and
                    all-ones, cache/external placement by s
    build/scan/pass brute-force Python set semantics (windows without a
                    newline) on small texts with newlines and high bytes;
                    threaded scan equals the single-threaded one
    verify          naive matcher restricted to patterns whose prefix window
                    has no newline
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libffbf_oracle.so")


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    return ctypes.CDLL(LIB_PATH)


_lib = _load()
_P = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32


class Params(ctypes.Structure):
    _fields_ = [
        ("window", ctypes.c_uint32),
        ("k", ctypes.c_uint32),
        ("log2_bits", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("sbox", ctypes.c_uint32 * 256),
    ]


_sig = {
    "orc_hash_window": (None, [_P, _P, _P, _P]),
    "orc_hash_rolling": (None, [_P, _P, _u64, _P, _P]),
    "orc_fwd_indices": (None, [_P, _u32, _u32, _P]),
    "orc_rev_indices": (None, [_P, _u32, _u32, _P]),
    "orc_build": (_u64, [_P, _P, _P, _u64, _P]),
    "orc_scan": (_u64, [_P, _P, _u64, _P, _P, _P, _u64]),
    "orc_scan_mt": (_u64, [_P, _P, _u64, _P, _P, _P, _u64, ctypes.c_int]),
    "orc_pattern_pass": (_u64, [_P, _P, _P, _u64, _P, _P]),
    "orc_verify": (_u64, [_P, _P, _u64, _P, _P, _u64, _P, _u64, _P, _P, _P, _u64]),
    "orc_brute_matches": (_u64, [_P, _u64, _P, _P, _u64, _u32, _P, _P, _u64]),
    "orc_filter_save_deltas": (_u64, [_P, _u32, _P, _u64]),
    "orc_filter_load_deltas": (None, [_P, _u64, _u32, _P]),
    "orc_filter_save_deltas_bits": (_u64, [_P, _u64, _P, _u64]),
    "orc_split_bits": (_u64, [_P]),
    "orc_split_hash_window": (None, [_P, _P, _P, _P]),
    "orc_split_fwd_indices": (None, [_P, _u32, _u32, _P]),
    "orc_split_rev_indices": (None, [_P, _u32, _u32, _P]),
    "orc_split_build": (_u64, [_P, _P, _P, _u64, _P]),
    "orc_split_scan": (_u64, [_P, _P, _u64, _P, _P, _P, _u64]),
    "orc_split_scan_mt": (_u64, [_P, _P, _u64, _P, _P, _P, _u64, ctypes.c_int]),
    "orc_split_pattern_pass": (_u64, [_P, _P, _P, _u64, _P, _P]),
    "orc_split_verify": (_u64, [_P, _P, _u64, _P, _P, _u64, _P, _u64, _P, _P, _P, _u64]),
    "orc_query_lines": (_u64, [_P, _P, _P, _u64, _P, _P]),
    "orc_split_query_lines": (_u64, [_P, _P, _P, _u64, _P, _P]),
}
for _n, (_r, _a) in _sig.items():
    _f = getattr(_lib, _n)
    _f.restype = _r
    _f.argtypes = _a


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data if a.size else 0


class Oracle:
    """Parameter set (W, k, log2 M, sbox) plus the reference stages on the CPU."""

    def __init__(self, window: int, k: int, log2_bits: int, sbox):
        self.p = Params()
        self.p.window, self.p.k, self.p.log2_bits, self.p.reserved = window, k, log2_bits, 0
        sb = np.asarray(sbox, dtype=np.uint32)
        for i in range(256):
            self.p.sbox[i] = int(sb[i])
        self.window, self.k, self.log2_bits = window, k, log2_bits
        self.n_words = max(1, 1 << max(log2_bits - 5, 0))
        self._pp = ctypes.byref(self.p)

    def hash_window(self, win: bytes):
        w = np.frombuffer(bytes(win), dtype=np.uint8)
        h1, h2 = ctypes.c_uint32(0), ctypes.c_uint32(0)
        _lib.orc_hash_window(self._pp, _ptr(w), ctypes.byref(h1), ctypes.byref(h2))
        return h1.value, h2.value

    def hash_rolling(self, text) -> tuple:
        t = np.ascontiguousarray(np.frombuffer(bytes(text), dtype=np.uint8) if isinstance(text, (bytes, bytearray)) else text)
        m = max(0, t.size - self.window + 1)
        h1 = np.zeros(m, dtype=np.uint32)
        h2 = np.zeros(m, dtype=np.uint32)
        _lib.orc_hash_rolling(self._pp, _ptr(t), t.size, _ptr(h1), _ptr(h2))
        return h1, h2

    def fwd_indices(self, h1: int, h2: int) -> list:
        out = np.zeros(5, dtype=np.uint32)
        _lib.orc_fwd_indices(self._pp, h1, h2, _ptr(out))
        return out[: self.k].tolist()

    def rev_indices(self, h1: int, h2: int) -> list:
        out = np.zeros(5, dtype=np.uint32)
        _lib.orc_rev_indices(self._pp, h1, h2, _ptr(out))
        return out[: self.k].tolist()

    def build(self, pat_bytes: np.ndarray, pat_off: np.ndarray) -> np.ndarray:
        F = np.zeros(self.n_words, dtype=np.uint32)
        _lib.orc_build(self._pp, _ptr(pat_bytes), _ptr(pat_off), pat_off.size - 1, _ptr(F))
        return F

    def scan(self, text: np.ndarray, F: np.ndarray, threads: int = 1):
        """Returns (R, candidate offsets ascending)."""
        R = np.zeros(self.n_words, dtype=np.uint32)
        cap = max(1, min(text.size, 1 << 22))
        while True:
            cand = np.zeros(cap, dtype=np.uint64)
            if threads > 1:
                n = _lib.orc_scan_mt(self._pp, _ptr(text), text.size, _ptr(F), _ptr(R), _ptr(cand), cap, threads)
            else:
                n = _lib.orc_scan(self._pp, _ptr(text), text.size, _ptr(F), _ptr(R), _ptr(cand), cap)
            if n <= cap:
                return R, cand[:n]
            cap = int(n)

    def pattern_pass(self, pat_bytes: np.ndarray, pat_off: np.ndarray, R: np.ndarray) -> np.ndarray:
        n = pat_off.size - 1
        bm = np.zeros(max(1, (n + 31) // 32), dtype=np.uint32)
        _lib.orc_pattern_pass(self._pp, _ptr(pat_bytes), _ptr(pat_off), n, _ptr(R), _ptr(bm))
        return bm[: (n + 31) // 32]

    def verify(self, text, pat_bytes, pat_off, cand, surv_bitmap):
        """Sorted (pattern id, offset) matches as two arrays."""
        n = pat_off.size - 1
        sb = surv_bitmap if surv_bitmap.size else np.zeros(1, dtype=np.uint32)
        cap = 1 << 16
        while True:
            ids = np.zeros(cap, dtype=np.uint32)
            offs = np.zeros(cap, dtype=np.uint64)
            m = _lib.orc_verify(self._pp, _ptr(text), text.size, _ptr(pat_bytes), _ptr(pat_off), n,
                                _ptr(cand), cand.size, _ptr(sb), _ptr(ids), _ptr(offs), cap)
            if m <= cap:
                return ids[:m], offs[:m]
            cap = int(m)

    def query_lines(self, line_bytes: np.ndarray, line_off: np.ndarray, F: np.ndarray) -> np.ndarray:
        """Filter::operator[] per line (ffbf.cpp:285-308): uint8 0/1 per line."""
        n = line_off.size - 1
        hits = np.zeros(max(n, 1), dtype=np.uint8)
        fn = _lib.orc_split_query_lines if isinstance(self, SplitOracle) else _lib.orc_query_lines
        fn(self._pp, _ptr(line_bytes), _ptr(line_off), n, _ptr(F), _ptr(hits))
        return hits[:n]

    def run(self, text, pat_bytes, pat_off, threads: int = 1) -> dict:
        F = self.build(pat_bytes, pat_off)
        R, cand = self.scan(text, F, threads)
        surv = self.pattern_pass(pat_bytes, pat_off, R)
        ids, offs = self.verify(text, pat_bytes, pat_off, cand, surv)
        return {"F": F, "R": R, "cand": cand, "surv": surv, "ids": ids, "offs": offs}


class SplitParams(ctypes.Structure):
    _fields_ = [
        ("window", ctypes.c_uint32),
        ("k", ctypes.c_uint32),
        ("s", ctypes.c_uint32),
        ("log2_cache_bits", ctypes.c_uint32),
        ("log2_ext_bits", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("sbox", ctypes.c_uint64 * 256),
    ]


class SplitOracle(Oracle):
    """The CPU binary's split filter (src/rabin-karp/ffbf.cpp): 64-bit sbox
    hash, k functions of which the first s address a 2^c-bit cache part and
    the rest a 2^e-bit external part, hashing restarted at every newline and
    bytes >= 0x80 hashed as 0.  Same stage interface as Oracle."""

    def __init__(self, window: int, k: int, s: int, log2_cache_bits: int, log2_ext_bits: int, sbox):
        self.p = SplitParams()
        self.p.window, self.p.k, self.p.s = window, k, s
        self.p.log2_cache_bits, self.p.log2_ext_bits, self.p.reserved = log2_cache_bits, log2_ext_bits, 0
        sb = np.asarray(sbox, dtype=np.uint64)
        for i in range(256):
            self.p.sbox[i] = int(sb[i])
        self.window, self.k, self.s = window, k, s
        self.log2_cache_bits, self.log2_ext_bits = log2_cache_bits, log2_ext_bits
        self.n_bits = (1 << log2_cache_bits) + (1 << log2_ext_bits)
        self.n_words = self.n_bits // 32
        self._pp = ctypes.byref(self.p)

    def hash_window(self, win: bytes):
        w = np.frombuffer(bytes(win), dtype=np.uint8)
        h1, h2 = ctypes.c_uint32(0), ctypes.c_uint32(0)
        _lib.orc_split_hash_window(self._pp, _ptr(w), ctypes.byref(h1), ctypes.byref(h2))
        return h1.value, h2.value

    def hash_rolling(self, text):
        raise NotImplementedError("the split method hashes line by line; use scan")

    def fwd_indices(self, h1: int, h2: int) -> list:
        out = np.zeros(11, dtype=np.uint32)
        _lib.orc_split_fwd_indices(self._pp, h1, h2, _ptr(out))
        return out[: self.k].tolist()

    def rev_indices(self, h1: int, h2: int) -> list:
        out = np.zeros(11, dtype=np.uint32)
        _lib.orc_split_rev_indices(self._pp, h1, h2, _ptr(out))
        return out[: self.k].tolist()

    def build(self, pat_bytes: np.ndarray, pat_off: np.ndarray) -> np.ndarray:
        F = np.zeros(self.n_words, dtype=np.uint32)
        _lib.orc_split_build(self._pp, _ptr(pat_bytes), _ptr(pat_off), pat_off.size - 1, _ptr(F))
        return F

    def scan(self, text: np.ndarray, F: np.ndarray, threads: int = 1):
        R = np.zeros(self.n_words, dtype=np.uint32)
        cap = max(1, min(text.size, 1 << 22))
        while True:
            cand = np.zeros(cap, dtype=np.uint64)
            if threads > 1:
                n = _lib.orc_split_scan_mt(self._pp, _ptr(text), text.size, _ptr(F), _ptr(R), _ptr(cand), cap,
                                           threads)
            else:
                n = _lib.orc_split_scan(self._pp, _ptr(text), text.size, _ptr(F), _ptr(R), _ptr(cand), cap)
            if n <= cap:
                return R, cand[:n]
            cap = int(n)

    def pattern_pass(self, pat_bytes: np.ndarray, pat_off: np.ndarray, R: np.ndarray) -> np.ndarray:
        n = pat_off.size - 1
        bm = np.zeros(max(1, (n + 31) // 32), dtype=np.uint32)
        _lib.orc_split_pattern_pass(self._pp, _ptr(pat_bytes), _ptr(pat_off), n, _ptr(R), _ptr(bm))
        return bm[: (n + 31) // 32]

    def verify(self, text, pat_bytes, pat_off, cand, surv_bitmap):
        n = pat_off.size - 1
        sb = surv_bitmap if surv_bitmap.size else np.zeros(1, dtype=np.uint32)
        cap = 1 << 16
        while True:
            ids = np.zeros(cap, dtype=np.uint32)
            offs = np.zeros(cap, dtype=np.uint64)
            m = _lib.orc_split_verify(self._pp, _ptr(text), text.size, _ptr(pat_bytes), _ptr(pat_off), n,
                                      _ptr(cand), cand.size, _ptr(sb), _ptr(ids), _ptr(offs), cap)
            if m <= cap:
                return ids[:m], offs[:m]
            cap = int(m)


def filter_save_deltas_bits(F: np.ndarray, n_bits: int) -> np.ndarray:
    n = _lib.orc_filter_save_deltas_bits(_ptr(F), n_bits, None, 0)
    out = np.zeros(max(int(n), 1), dtype=np.uint32)
    _lib.orc_filter_save_deltas_bits(_ptr(F), n_bits, _ptr(out), n)
    return out[: int(n)]


def brute_matches(text: np.ndarray, pat_bytes: np.ndarray, pat_off: np.ndarray, min_len: int):
    cap = 1 << 16
    n = pat_off.size - 1
    while True:
        ids = np.zeros(cap, dtype=np.uint32)
        offs = np.zeros(cap, dtype=np.uint64)
        m = _lib.orc_brute_matches(_ptr(text), text.size, _ptr(pat_bytes), _ptr(pat_off), n, min_len,
                                   _ptr(ids), _ptr(offs), cap)
        if m <= cap:
            return ids[:m], offs[:m]
        cap = int(m)


def filter_save_deltas(F: np.ndarray, log2_bits: int) -> np.ndarray:
    n = _lib.orc_filter_save_deltas(_ptr(F), log2_bits, None, 0)
    out = np.zeros(max(int(n), 1), dtype=np.uint32)
    _lib.orc_filter_save_deltas(_ptr(F), log2_bits, _ptr(out), n)
    return out[: int(n)]


def filter_load_deltas(deltas: np.ndarray, log2_bits: int) -> np.ndarray:
    F = np.zeros(max(1, 1 << max(log2_bits - 5, 0)), dtype=np.uint32)
    d = np.ascontiguousarray(deltas, dtype=np.uint32)
    _lib.orc_filter_load_deltas(_ptr(d), d.size, log2_bits, _ptr(F))
    return F


def candidate_lines(text, cand) -> bytes:
    """Filtered corpus: each line holding a candidate window start, once, in
    text order, with a trailing newline (printpositions, greptest.cu:353-398:
    lines are delimited by '\n' (:375-390); a candidate on a '\n' byte or in an
    already printed line prints nothing (:373); the reference's memchr past the
    end (:387) is read as "the line ends at the end of the text")."""
    t = bytes(np.asarray(text, dtype=np.uint8))
    out = []
    prev_end = -1
    for o in sorted(set(int(x) for x in cand)):
        if o <= prev_end or t[o:o + 1] == b"\n":
            continue
        start = t.rfind(b"\n", 0, o) + 1
        end = t.find(b"\n", o)
        if end < 0:
            end = len(t)
        out.append(t[start:end] + b"\n")
        prev_end = end
    return b"".join(out)


def survivor_files(pat_bytes, pat_off, surv_bitmap):
    """(filtered patterns, index) file contents: every surviving pattern on
    its own line in id order, and its 0-based line number (filterPhrases,
    ffbf.cpp:430-440 writes the line and lineNumber; ids count from 0 at
    :404)."""
    b = bytes(np.asarray(pat_bytes, dtype=np.uint8))
    off = [int(x) for x in pat_off]
    bits = np.asarray(surv_bitmap, dtype=np.uint32)
    pats, idx = [], []
    for i in range(len(off) - 1):
        if (int(bits[i >> 5]) >> (i & 31)) & 1:
            pats.append(b[off[i]:off[i + 1]] + b"\n")
            idx.append(b"%d\n" % i)
    return b"".join(pats), b"".join(idx)
