"""Shared fixtures.  Tests marked `gpu` need an MI355X; everything else runs on
the CPU (oracle, fixtures, generators, library loading)."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def ref_sbox():
    return np.array(load_golden("ref_sbox.json")["sbox"], dtype=np.uint32)


@pytest.fixture(scope="session")
def ref_sbox64():
    return np.array([int(x) for x in load_golden("ref_sbox64.json")["sbox"]], dtype=np.uint64)


def bits_of(words: np.ndarray) -> list:
    """Sorted indices of set bits of a uint32 word array (bit b in word b>>5)."""
    w = np.asarray(words, dtype=np.uint32)
    bits = np.unpackbits(w.view(np.uint8), bitorder="little")
    return np.nonzero(bits)[0].tolist()


def words_of(bits, n_words: int) -> np.ndarray:
    out = np.zeros(n_words, dtype=np.uint32)
    for b in bits:
        out[b >> 5] |= np.uint32(1 << (b & 31))
    return out
