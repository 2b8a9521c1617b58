"""The C ABI from plain C: tests/c/ffbf_c_client.c is compiled with gcc
(C99, warnings as errors) against include/ffbf.h and linked to
libffbf_amd.so.  Without a GPU it must report the device error and exit
cleanly; on an MI355X (-m gpu) its matches must equal the oracle's."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

import ffbf_amd as fa
from oracle import Oracle

SRC = os.path.join(ROOT, "tests", "c", "ffbf_c_client.c")
LIBDIR = os.path.dirname(fa.LIB_PATH)


def build_client(tmp_path):
    exe = str(tmp_path / "ffbf_c_client")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-O1", "-I", os.path.join(ROOT, "include"),
                    SRC, "-o", exe, "-L", LIBDIR, "-lffbf_amd", "-Wl,-rpath," + LIBDIR],
                   check=True, capture_output=True, text=True)
    return exe


def test_c_client_builds_and_reports_errors_without_device(tmp_path):
    if fa.device_count():
        pytest.skip("a device is present: covered by the gpu test")
    exe = build_client(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert lines[0].startswith("version ") and "gfx950" in lines[0]
    assert lines[1].startswith("create status -2: ")  # FFBF_EDEVICE, with a message
    assert len(lines[1]) > len("create status -2: ")


@pytest.mark.gpu
def test_c_client_matches_equal_oracle(tmp_path):
    exe = build_client(tmp_path)
    out = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    got = [tuple(int(x) for x in l.split()[1:]) for l in out.stdout.splitlines() if l.startswith("match ")]
    sbox = np.array([((i * 2654435761) ^ 0x9E3779B9) & 0xFFFFFFFF for i in range(256)], dtype=np.uint32)
    pats = [b"needle in", b"haystack", b"abc", b"rolling hash"]
    b, off = fa.patterns_from_list(pats)
    text = np.frombuffer(b"a haystack with a needle in it, and a rolling hash over a haystack", dtype=np.uint8)
    r = Oracle(8, 4, 16, sbox).run(text.copy(), b, off)
    expect = list(zip(r["ids"].tolist(), r["offs"].tolist()))
    assert got == expect
    assert expect == [(0, 18), (1, 2), (1, 58), (3, 38)]
