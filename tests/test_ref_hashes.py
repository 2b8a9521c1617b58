"""The oracle's window hashes against the reference's own code.

oracle/ref.mk compiles the reference's unmodified rolling-hash headers
(src/hashes/hash_rot_sbox_2.h, the CPU form of the GPU hash pair, and
src/hashes/hash_rot_sbox_pre_2.h, the CPU binary's 64-bit hash) behind a small
C ABI into oracle/_ref/libref_hashes.so.  Their rolling update, fed from a
reset state, must give on every window the pair the oracle computes in closed
form, for every W from 1 to 64, with the reference's own tables (kept here as
data in tests/golden/ref_sbox.json and ref_sbox64.json).  The reference tree
exists only where the library is built, not where the GPU tests run: there
this module is skipped."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle import Oracle, SplitOracle

REF = "/root/reference"
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_hashes.so")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src", "hashes")),
                                reason="the reference tree is not present on this machine")


@pytest.fixture(scope="module")
def ref():
    subprocess.run(["make", "-s", "-f", "ref.mk", "-C", os.path.join(ROOT, "oracle")], check=True,
                   capture_output=True, text=True)
    lib = ctypes.CDLL(LIB)
    for f in (lib.ref_pair_hashes, lib.ref_split_hashes):
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
    return lib


def _ref_hashes(fn, W, text):
    m = text.size - W + 1
    h1 = np.zeros(m, dtype=np.uint32)
    h2 = np.zeros(m, dtype=np.uint32)
    assert fn(W, text.ctypes.data, text.size, h1.ctypes.data, h2.ctypes.data) == 0
    return h1, h2


@pytest.mark.parametrize("W", list(range(1, 65)))
def test_pair_hash_equals_reference(ref, ref_sbox, W):
    rng = np.random.default_rng(1000 + W)
    text = rng.integers(0, 256, size=4 * W + 97, dtype=np.uint8)
    r1, r2 = _ref_hashes(ref.ref_pair_hashes, W, text)
    o = Oracle(W, 4, 20, ref_sbox)
    o1, o2 = o.hash_rolling(text)
    assert np.array_equal(r1, o1) and np.array_equal(r2, o2)
    for i in (0, 1, text.size - W):
        assert (int(r1[i]), int(r2[i])) == o.hash_window(text[i:i + W].tobytes())


@pytest.mark.parametrize("W", list(range(1, 65)))
def test_split_hash_equals_reference(ref, ref_sbox64, W):
    # bytes below 0x80: the CPU binary folds higher bytes to 0 before hashing
    # (ffbf.cpp:124-133); the oracle does the same, pinned separately
    rng = np.random.default_rng(2000 + W)
    text = rng.integers(0, 128, size=3 * W + 61, dtype=np.uint8)
    r1, r2 = _ref_hashes(ref.ref_split_hashes, W, text)
    o = SplitOracle(W, 5, 2, 10, 12, ref_sbox64)
    for i in range(text.size - W + 1):
        assert (int(r1[i]), int(r2[i])) == o.hash_window(text[i:i + W].tobytes()), i


def test_reference_hashes_reject_other_window_lengths(ref):
    t = np.zeros(8, dtype=np.uint8)
    out = np.zeros(8, dtype=np.uint32)
    for W in (0, 65):
        assert ref.ref_pair_hashes(W, t.ctypes.data, t.size, out.ctypes.data, out.ctypes.data) == -1
        assert ref.ref_split_hashes(W, t.ctypes.data, t.size, out.ctypes.data, out.ctypes.data) == -1


def test_a_planted_oracle_mistake_is_caught(ref, ref_sbox):
    # rotating h2 by 2 instead of 3 per byte: the reference disagrees
    W = 19
    text = np.random.default_rng(7).integers(0, 256, size=200, dtype=np.uint8)
    r1, r2 = _ref_hashes(ref.ref_pair_hashes, W, text)
    sb = np.asarray(ref_sbox, dtype=np.uint64)

    def rotl(x, r):
        r %= 32
        return ((x << r) | (x >> (32 - r))) & 0xFFFFFFFF if r else x

    bad = 0
    for j in range(W):
        bad = rotl(bad, 2) ^ int(sb[text[j]])
    assert bad != int(r2[0])
    good = 0
    for j in range(W):
        good = rotl(good, 3) ^ int(sb[text[j]])
    assert good == int(r2[0])
