"""Text ingest (host memory, files, pipes through the pinned staging ring with
per-piece scans) and the reference's text outputs (filtered patterns with
their index, filtered corpus lines) against the oracle and against the
device-resident scan.  Needs an MI355X (-m gpu)."""
import os
import sys
import threading

import numpy as np
import pytest

import ffbf_amd as fa
from oracle import Oracle, candidate_lines, survivor_files

pytestmark = pytest.mark.gpu

W, K = 19, 4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_matcher(seed, n_pat=3000, log2=20, alpha=fa.ALPHA_ASCII):
    sbox = fa.sbox_from_seed(seed)
    b, off = fa.gen_patterns(alpha, seed + 1, n_pat, W, 48)
    m = fa.Matcher(W, K, log2, sbox)
    m.set_patterns(b, off)
    m.build()
    return m, sbox, b, off


def make_text(seed, n, b, off, alpha=fa.ALPHA_ASCII, newline_every=0):
    t = fa.gen_bytes(alpha, seed, n)
    if n > 100:
        fa.plant(seed + 7, 1 + n // 20000, t, b, off)
    if newline_every:
        rng = np.random.default_rng(seed)
        t[rng.integers(0, max(n, 1), size=max(1, n // newline_every))] = ord("\n")
    return t


def device_reference(m, t):
    """Candidates, R and matches of the same text scanned from device memory."""
    dev = fa.DeviceBuffer.from_host(t if t.size else np.zeros(1, np.uint8))
    m.run_device(dev.ptr, t.size)
    out = (m.candidates().copy(), m.filter(fa.FILTER_REVERSE).copy(), m.matches().copy())
    dev.free()
    return out


def same_as(m, ref):
    cand, R, mt = ref
    assert np.array_equal(m.candidates(), cand)
    assert np.array_equal(m.filter(fa.FILTER_REVERSE), R)
    m.pattern_pass()
    m.verify()
    got = m.matches()
    assert np.array_equal(got["pattern_id"], mt["pattern_id"]) and np.array_equal(got["offset"], mt["offset"])


@pytest.mark.parametrize("chunk_mib,depth,threads", [(1, 2, 1), (1, 4, 8), (2, 3, 4)])
def test_host_scan_in_pieces_equals_device_scan(chunk_mib, depth, threads):
    say(f"host ingest chunk={chunk_mib} MiB depth={depth} threads={threads}")
    m, sbox, b, off = make_matcher(11)
    m.set_ingest(chunk_mib << 20, depth, threads)
    # sizes around piece boundaries, a piece shorter than W, and empty/short texts
    sizes = [0, 5, W - 1, W, (1 << 20) - 3, (1 << 20) + W // 2, 3 * (1 << 20) + 7, (5 << 20) + 1]
    for i, n in enumerate(sizes):
        t = make_text(100 + i, n, b, off)
        ref = device_reference(m, t)
        m.scan(t)
        same_as(m, ref)
    # one text checked against the oracle as well
    t = make_text(150, (2 << 20) + 11, b, off)
    r = Oracle(W, K, 20, sbox).run(t, b, off)
    m.scan(t)
    assert np.array_equal(m.candidates(), np.sort(r["cand"]))
    assert np.array_equal(m.filter(fa.FILTER_REVERSE), r["R"])
    m.pattern_pass()
    m.verify()
    got = m.matches()
    assert np.array_equal(got["pattern_id"], r["ids"]) and np.array_equal(got["offset"], r["offs"])
    assert m.timings()["ingest_ms"] > 0
    m.close()


def test_scan_file_fd_and_pipe(tmp_path):
    say("file / fd / pipe ingest")
    m, sbox, b, off = make_matcher(21, n_pat=5000, log2=22, alpha=fa.ALPHA_DNA_GC40)
    m.set_ingest(1 << 20, 4, 4)
    t = make_text(200, (6 << 20) + 333, b, off, alpha=fa.ALPHA_DNA_GC40)
    ref = device_reference(m, t)
    path = tmp_path / "text.bin"
    t.tofile(path)
    # by name, and scan + pass + verify in one call
    m.scan_file(path)
    same_as(m, ref)
    assert m.run_file(path) == ref[2].size
    # from an open descriptor at a position: only the rest of the file
    cut = (1 << 20) + 5
    ref_tail = device_reference(m, t[cut:])
    with open(path, "rb") as f:
        f.seek(cut)
        m.scan_fd(f.fileno())
        assert f.tell() == t.size
    same_as(m, ref_tail)
    # from a pipe, written in uneven pieces by another thread (size unknown:
    # the device text grows while pieces arrive)
    r_fd, w_fd = os.pipe()
    rng = np.random.default_rng(5)

    def writer():
        pos = 0
        while pos < t.size:
            step = int(rng.integers(1, 700_000))
            os.write(w_fd, t[pos:pos + step].tobytes())
            pos += step
        os.close(w_fd)

    th = threading.Thread(target=writer)
    th.start()
    try:
        m.scan_fd(r_fd)
    finally:
        th.join()
        os.close(r_fd)
    same_as(m, ref)
    # errors come back as status codes
    with pytest.raises(fa.FFBFError):
        m.scan_file(tmp_path / "missing.bin")
    m.close()


def test_reference_text_outputs(tmp_path):
    say("filtered patterns, index and corpus lines")
    m, sbox, b, off = make_matcher(31, n_pat=2000, log2=15)  # small filter: many false candidates
    t = make_text(300, 400_000, b, off, newline_every=90)
    r = Oracle(W, K, 15, sbox).run(t, b, off)
    assert r["cand"].size > 300
    m.scan(t)
    m.pattern_pass()
    pats, idx = tmp_path / "phrases.out", tmp_path / "__index_phrases.out"
    m.write_survivors(pats, idx)
    want_p, want_i = survivor_files(b, off, r["surv"])
    assert pats.read_bytes() == want_p and idx.read_bytes() == want_i
    assert len(want_i) > 0
    corpus = tmp_path / "corpus.out"
    m.write_candidate_lines(corpus, t)
    assert corpus.read_bytes() == candidate_lines(t, r["cand"])
    # the device copy of the text gives the same file
    m.write_candidate_lines(corpus)
    assert corpus.read_bytes() == candidate_lines(t, r["cand"])
    with pytest.raises(fa.FFBFError):
        m.write_candidate_lines(corpus, t[:-1])  # not the scanned text
    m.close()
