"""ffbf_run_device_noalloc: the allocation-free variant of ffbf_run_device.

Its scratch (counters, candidate staging, candidate and match lists, sort
buffers) is a fixed device array of the library, shared by every call on a
GPU.  Each call's matches, reverse filter and survivor counts must equal
ffbf_run_device's on the same text, on consecutive calls (the second one on
a smaller text, so stale contents of the array would show), and against the
oracle."""
import ctypes

import numpy as np
import pytest

import ffbf_amd as fa
from oracle import Oracle, SplitOracle

pytestmark = pytest.mark.gpu

def noalloc(m, dev, n, out, cap):
    """ffbf_run_device_noalloc through the C ABI: (status, total matches)."""
    nm = ctypes.c_uint64(0)
    st = fa.lib().ffbf_run_device_noalloc(m._h, dev.ptr, n, out.ptr if out is not None else None, cap,
                                          ctypes.byref(nm))
    return st, nm.value


def read_matches(out, n):
    raw = out.download(n * fa.MATCH_DTYPE.itemsize) if n else np.zeros(0, dtype=np.uint8)
    return raw.view(fa.MATCH_DTYPE)


def same(a, b):
    return np.array_equal(a["pattern_id"], b["pattern_id"]) and np.array_equal(a["offset"], b["offset"])


def _case(alpha, W, k, log2, n_pat, lmin, lmax, n_text, plants, seed):
    sbox = fa.sbox_from_seed(seed)
    pb, po = fa.gen_patterns(alpha, seed * 10 + 1, n_pat, lmin, lmax)
    text = fa.gen_bytes(alpha, seed * 10 + 2, n_text)
    fa.plant(seed * 10 + 3, plants, text, pb, po)
    return sbox, pb, po, text


@pytest.mark.parametrize("mode", [-1, 0, 1, 3])
def test_noalloc_equals_run_device_twice(mode):
    W, k, log2 = 32, 4, 22
    sbox, pb, po, text = _case(fa.ALPHA_DNA_GC40, W, k, log2, 20000, 32, 100, 6 << 20, 3000, 71)
    n1, n2 = text.size, text.size // 3 + 5
    print(f"noalloc mode {mode}: {po.size - 1} patterns, texts {n1} then {n2} bytes", flush=True)
    dev = fa.DeviceBuffer.from_host(text)
    out = fa.DeviceBuffer(1 << 20)
    with fa.Matcher(W, k, log2, sbox) as m:
        m.set_patterns(pb, po)
        m.build()
        m.set_scan_mode(mode)
        for n in (n1, n2):
            m.run_device(dev.ptr, n)
            ref = m.matches()
            ref_R = m.filter(fa.FILTER_REVERSE)
            st, nm = noalloc(m, dev, n, out, (1 << 20) // 16)
            assert st == fa.OK, fa.lib().ffbf_last_error()
            got = read_matches(out, nm)
            assert nm == ref.size and ref.size > 100
            assert same(got, ref)
            assert np.array_equal(m.filter(fa.FILTER_REVERSE), ref_R)
            o = Oracle(W, k, log2, sbox).run(text[:n], pb, po)
            assert np.array_equal(got["pattern_id"], o["ids"]) and np.array_equal(got["offset"], o["offs"])
    out.free()
    dev.free()


def test_noalloc_lds_filter_and_split_method():
    # whole filter in LDS with R gathered in LDS (2^15 bits), ASCII
    sbox, pb, po, text = _case(fa.ALPHA_ASCII, 19, 3, 15, 800, 19, 40, 3 << 20, 500, 72)
    dev = fa.DeviceBuffer.from_host(text)
    out = fa.DeviceBuffer(1 << 20)
    with fa.Matcher(19, 3, 15, sbox) as m:
        m.set_patterns(pb, po)
        m.build()
        for n in (text.size, text.size // 2 + 3):
            m.run_device(dev.ptr, n)
            ref = m.matches()
            ref_R = m.filter(fa.FILTER_REVERSE)
            st, nm = noalloc(m, dev, n, out, (1 << 20) // 16)
            assert st == fa.OK, fa.lib().ffbf_last_error()
            assert same(read_matches(out, nm), ref) and ref.size > 100
            assert np.array_equal(m.filter(fa.FILTER_REVERSE), ref_R)
    dev.free()
    # the CPU binary's split filter, text in lines
    sb64 = np.random.default_rng(73).integers(0, 2**64, size=256, dtype=np.uint64)
    b2, o2 = fa.gen_patterns(fa.ALPHA_ASCII, 74, 3000, 19, 40)
    t2 = fa.gen_bytes(fa.ALPHA_ASCII, 75, 2 << 20)
    t2[np.random.default_rng(76).integers(0, t2.size, size=t2.size // 60)] = 10
    fa.plant(77, 400, t2, b2, o2)
    dev = fa.DeviceBuffer.from_host(t2)
    with fa.Matcher.split(19, 5, 2, 16, 22, sb64) as m:
        m.set_patterns(b2, o2)
        m.build()
        for n in (t2.size, t2.size // 2 + 1):
            m.run_device(dev.ptr, n)
            ref = m.matches()
            st, nm = noalloc(m, dev, n, out, (1 << 20) // 16)
            assert st == fa.OK, fa.lib().ffbf_last_error()
            got = read_matches(out, nm)
            assert same(got, ref) and ref.size > 50
            o = SplitOracle(19, 5, 2, 16, 22, sb64).run(t2[:n], b2, o2)
            assert np.array_equal(got["pattern_id"], o["ids"]) and np.array_equal(got["offset"], o["offs"])
    out.free()
    dev.free()


def test_noalloc_limits_are_errors():
    sbox, pb, po, text = _case(fa.ALPHA_DNA_GC40, 24, 3, 20, 5000, 24, 60, 2 << 20, 800, 78)
    dev = fa.DeviceBuffer.from_host(text)
    out = fa.DeviceBuffer(1 << 20)
    cap = (1 << 20) // 16
    with fa.Matcher(24, 3, 20, sbox) as m:
        m.set_patterns(pb, po)
        # before a build
        st, _ = noalloc(m, dev, text.size, out, cap)
        assert st == fa.ESTATE
        m.build()
        # an output buffer smaller than the match list: the total is reported
        m.run_device(dev.ptr, text.size)
        ref = m.matches()
        assert ref.size > 10
        st, nm = noalloc(m, dev, text.size, out, 4)
        assert st == fa.ECAPACITY and nm == ref.size
        st, nm = noalloc(m, dev, text.size, out, cap)
        assert st == fa.OK and same(read_matches(out, nm), ref)
    # more candidate windows than the scratch holds: every window of a
    # saturated filter is a candidate (2^20 > 2^18 windows)
    with fa.Matcher(24, 1, 10, sbox) as m:
        pats = [bytes(text[i * 40:i * 40 + 24]) for i in range(2000)]
        m.set_pattern_list(pats)
        m.build()
        st, _ = noalloc(m, dev, 1 << 20, out, cap)
        assert st == fa.ECAPACITY
        assert b"candidate" in fa.lib().ffbf_last_error()
        # the same context still runs a small text afterwards
        m.run_device(dev.ptr, 50000)
        ref = m.matches()
        st, nm = noalloc(m, dev, 50000, out, cap)
        assert st == fa.OK and nm == ref.size and same(read_matches(out, nm), ref)
    out.free()
    dev.free()
