"""Filters of 2^29 to 2^32 bits (64 MiB to 512 MiB of F), the top of the
range ffbf_create accepts: F, R and the match list against the oracle in
every scan mode that applies (F probed directly, the LDS fold, the
L2-resident fold, automatic).  At 2^32 bits the index mask is all ones and
word indices reach 2^27 - 1.  Needs an MI355X."""
import numpy as np
import pytest

import ffbf_amd as fa
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log2", [29, 30, 31, 32])
def test_large_filters_match_oracle(log2):
    W, k = 24, 4
    sbox = fa.sbox_from_seed(400 + log2)
    pb, po = fa.gen_patterns(fa.ALPHA_DNA_GC40, 401 + log2, 40000, 24, 60)
    text = fa.gen_bytes(fa.ALPHA_DNA_GC40, 402 + log2, 6 << 20)
    fa.plant(403 + log2, 3000, text, pb, po)
    print(f"filter 2^{log2} bits: 40000 patterns, 6 MiB text", flush=True)
    ref = Oracle(W, k, log2, sbox).run(text, pb, po, threads=8)
    assert ref["ids"].size > 2000
    dev = fa.DeviceBuffer.from_host(text)
    with fa.Matcher(W, k, log2, sbox) as m:
        m.set_patterns(pb, po)
        m.build()
        assert np.array_equal(m.filter(fa.FILTER_FORWARD), ref["F"])
        for mode in (fa.MODE_AUTO, fa.MODE_GLOBAL, fa.MODE_FOLD, fa.MODE_GFOLD, fa.MODE_HYBRID):
            m.set_scan_mode(mode)
            m.run_device(dev.ptr, text.size)
            got = m.matches()
            assert np.array_equal(got["pattern_id"], ref["ids"]), mode
            assert np.array_equal(got["offset"], ref["offs"]), mode
            assert np.array_equal(m.filter(fa.FILTER_REVERSE), ref["R"]), mode
            assert m.n_candidates() == ref["cand"].size, mode
    dev.free()
