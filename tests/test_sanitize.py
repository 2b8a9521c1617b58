"""Host code under sanitizers, on the CPU.

* The serving scheduler (ffbf_amd/csrc/ffbf_sched.cpp) with a fake executor
  (tests/c/sched_test.cpp): pieces and their ownership, admission by slot size
  and HBM budget, an executor that fails a batch, cancellation before, during
  and after a run, draining at shutdown -- under AddressSanitizer +
  UndefinedBehaviorSanitizer and under ThreadSanitizer.
* The host radix sort of match records (tests/c/hostsort_test.cpp) against
  std::stable_sort, under AddressSanitizer + UndefinedBehaviorSanitizer.
* The C oracle (tests/c/oracle_sanitize.c drives every function of
  oracle/ffbf_oracle.c over W 1..64, empty inputs and short caps) under
  AddressSanitizer + UndefinedBehaviorSanitizer.
Each is built with g++/gcc (no device code) and must exit 0."""
import os
import subprocess

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "ffbf_amd", "csrc")
INC = os.path.join(ROOT, "include")
SCHED_TEST = os.path.join(ROOT, "tests", "c", "sched_test.cpp")


def _run(exe, timeout):
    out = subprocess.run([exe], capture_output=True, text=True, timeout=timeout,
                         env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1"))
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


@pytest.mark.parametrize("san", ["address,undefined", "thread"])
def test_scheduler_under_sanitizers(tmp_path, san):
    exe = str(tmp_path / "sched_test")
    flags = ["-fsanitize=" + san, "-fno-omit-frame-pointer"]
    if san != "thread":
        flags.append("-fno-sanitize-recover=all")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-pthread", *flags, "-I", INC, "-I", CSRC,
                    SCHED_TEST, os.path.join(CSRC, "ffbf_sched.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    assert "all checks passed" in _run(exe, 300)


def test_host_match_sort_under_asan_ubsan(tmp_path):
    """The radix sort that orders match lists on the host (tests/c/hostsort_test.cpp)."""
    exe = str(tmp_path / "hostsort_test")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-I", CSRC,
                    os.path.join(ROOT, "tests", "c", "hostsort_test.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    assert "all checks passed" in _run(exe, 300)


def test_oracle_under_asan_ubsan(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize", f"OUT={tmp_path}"], check=True,
                   capture_output=True, text=True)
    assert "all checks passed" in _run(str(tmp_path / "oracle_sanitize"), 300)
