"""C-ABI library checks that need no GPU: it loads, exports every declared
symbol, reports errors instead of crashing, and its seeded workload
generators are deterministic and have the stated distributions."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import ffbf_amd as fa


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(ffbf_\w+)\s*\(", text))


def test_every_declared_symbol_is_exported_and_bound():
    lib = ctypes.CDLL(fa.LIB_PATH)
    declared = _declared("ffbf.h") | _declared("ffbf_workload.h")
    assert len(declared) >= 35
    for name in sorted(declared):
        assert hasattr(lib, name), name
        assert name in fa.ABI, f"{name} not bound in ffbf_amd.ABI"


def test_oracle_library_loads():
    import oracle

    assert os.path.exists(oracle.LIB_PATH)


def test_version_and_errors_without_device():
    assert b"gfx950" in fa.lib().ffbf_version()
    n = fa.device_count()
    if n:
        pytest.skip("a HIP device is present")
    with pytest.raises(fa.FFBFError) as e:
        fa.Matcher(19, 4, 24, np.arange(256, dtype=np.uint32))
    assert e.value.status == fa.EDEVICE
    assert "device" in str(e.value)


def test_bad_params_rejected_before_device_use():
    for kw in ({"window": 0}, {"window": 65}, {"k": 0}, {"k": 6}, {"log2_bits": 9}, {"log2_bits": 33}):
        args = {"window": 19, "k": 4, "log2_bits": 24}
        args.update(kw)
        with pytest.raises(fa.FFBFError) as e:
            fa.Matcher(args["window"], args["k"], args["log2_bits"], np.zeros(256, dtype=np.uint32))
        assert e.value.status == fa.EINVAL


def test_null_context_is_an_error_not_a_crash():
    lib = fa.lib()
    assert lib.ffbf_build(None) == fa.EINVAL
    assert lib.ffbf_scan(None, None, 0) == fa.EINVAL
    assert b"null" in lib.ffbf_last_error()
    lib.ffbf_destroy(None)


def test_text_generator_is_counter_based_and_thread_independent():
    a = fa.gen_bytes(fa.ALPHA_DNA_GC40, 42, 1 << 21, threads=1)
    b = fa.gen_bytes(fa.ALPHA_DNA_GC40, 42, 1 << 21, threads=7)
    assert np.array_equal(a, b)
    # any window of the stream can be generated on its own
    c = fa.gen_bytes(fa.ALPHA_DNA_GC40, 42, 1001, start=123457)
    assert np.array_equal(c, a[123457:124458])
    assert not np.array_equal(a[:4096], fa.gen_bytes(fa.ALPHA_DNA_GC40, 43, 4096))


def test_dna_composition_gc40():
    a = fa.gen_bytes(fa.ALPHA_DNA_GC40, 7, 1 << 22)
    counts = {ch: int(np.count_nonzero(a == ord(ch))) / a.size for ch in "ACGT"}
    assert set(np.unique(a).tolist()) == {ord(c) for c in "ACGT"}
    assert abs(counts["A"] - 0.3) < 0.003 and abs(counts["T"] - 0.3) < 0.003
    assert abs(counts["C"] - 0.2) < 0.003 and abs(counts["G"] - 0.2) < 0.003


def test_ascii_uniform_printable():
    a = fa.gen_bytes(fa.ALPHA_ASCII, 9, 1 << 21)
    assert a.min() == 0x20 and a.max() == 0x7E
    h = np.bincount(a, minlength=256)[0x20:0x7F] / a.size
    assert np.all(np.abs(h - 1 / 95) < 0.0015)


def test_patterns_plants_and_repeats():
    b, off = fa.gen_patterns(fa.ALPHA_DNA_GC40, 5, 5000, 32, 100)
    lens = np.diff(off)
    assert lens.min() >= 32 and lens.max() <= 100 and abs(lens.mean() - 66) < 1.5
    assert b.size == int(off[-1])
    text = fa.gen_bytes(fa.ALPHA_DNA_GC40, 6, 1 << 20)
    ids, offs = fa.plant(8, 100, text, b, off)
    last = {}
    for j, (i, o) in enumerate(zip(ids.tolist(), offs.tolist())):
        last[o] = (i, j)
    # the last plant is intact
    i, o = int(ids[-1]), int(offs[-1])
    assert text[o:o + (off[i + 1] - off[i])].tobytes() == b[off[i]:off[i + 1]].tobytes()
    t2 = text.copy()
    fa.tandem_repeats(3, t2, block=4096)
    changed = np.count_nonzero(t2 != text) / t2.size
    assert 0.03 < changed < 0.12  # ~10% of bases rewritten, some unchanged by chance


def test_sbox_seeded():
    s1, s2 = fa.sbox_from_seed(1), fa.sbox_from_seed(2)
    assert s1.shape == (256,) and not np.array_equal(s1, s2)
    assert np.array_equal(s1, fa.sbox_from_seed(1))
    assert len(np.unique(s1)) == 256


def test_lognormal_sizes_deterministic_clipped_and_shaped():
    a = fa.lognormal_sizes(41, 20000, 1024 * 1024.0, 1.0, 4096, 64 << 20)
    b = fa.lognormal_sizes(41, 20000, 1024 * 1024.0, 1.0, 4096, 64 << 20)
    assert np.array_equal(a, b)  # same seed, same list
    assert not np.array_equal(a, fa.lognormal_sizes(42, 20000, 1024 * 1024.0, 1.0, 4096, 64 << 20))
    assert a.min() >= 4096 and a.max() <= 64 << 20
    # log-sizes: median ~ log(median), spread ~ sigma (clipping trims little at sigma 1)
    la = np.log(a.astype(np.float64))
    assert abs(np.median(la) - np.log(1024 * 1024.0)) < 0.05
    assert 0.9 < np.std(la) < 1.1
    # sigma 0 gives the median itself; a narrow clip range clamps everything
    assert np.all(fa.lognormal_sizes(1, 100, 5000.0, 0.0, 1, 1 << 20) == 5000)
    assert np.all(fa.lognormal_sizes(1, 100, 5000.0, 2.0, 777, 777) == 777)
    with pytest.raises(fa.FFBFError):
        fa.lognormal_sizes(1, 10, 5000.0, 1.0, 10, 5)  # lo > hi
