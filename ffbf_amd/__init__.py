"""MI355X feed-forward Bloom filter (FFBF) matcher -- Python side of the C ABI.

The work happens in ``libffbf_amd.so`` (HIP kernels for gfx950 plus the C++
context behind ``include/ffbf.h``); this module only binds it with ctypes and
offers a small object interface that follows the reference's flow
(efficient/ffbf):

=====================  ==============================================  ==================
this module            reference                                       file:line
=====================  ==============================================  ==================
``Matcher.build``      ``setup_bloom_search`` / ``Filter::operator<<``  greptest.cu:455-474, ffbf.cpp:268-283
``Matcher.scan``       ``executeGrepOverlap`` / ``Filter::processFile`` greptest.cu:505-576, ffbf.cpp:310-392
``Matcher.pattern_pass`` ``executePatternFiltering`` / ``filterPhrases`` greptest.cu:578-596, ffbf.cpp:394-485
``Matcher.verify``     ``grep -Ff`` on the filtered outputs             src/rabin-karp/README:49-52
``Matcher.filter``     ``bvDump``                                       greptest.cu:341-351
=====================  ==============================================  ==================

There is no CPU fallback: if the shared library is missing, importing this
package raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Iterable, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libffbf_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' "
        "or make -C ffbf_amd/csrc)"
    )

_lib = ctypes.CDLL(LIB_PATH)

OK = 0
EINVAL, EDEVICE, ENOMEM, ESTATE, ECAPACITY, ECANCELED = -1, -2, -3, -4, -5, -6
FILTER_FORWARD, FILTER_REVERSE, FILTER_FOLD, FILTER_L2_FOLD = 0, 1, 2, 3
MODE_AUTO, MODE_GLOBAL, MODE_FOLD, MODE_LDS, MODE_GFOLD, MODE_HYBRID = -1, 0, 1, 2, 3, 4
BUILD_AUTO, BUILD_ATOMIC, BUILD_BINNED = 0, 1, 2
ALPHA_ASCII, ALPHA_DNA_GC40 = 0, 1


class FFBFError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"ffbf status {status}: {message}")
        self.status = status


class Params(ctypes.Structure):
    _fields_ = [
        ("window", ctypes.c_uint32),
        ("k", ctypes.c_uint32),
        ("log2_bits", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("sbox", ctypes.c_uint32 * 256),
    ]


class SplitParams(ctypes.Structure):
    """ffbf_split_params: the CPU binary's split filter (64-bit sbox)."""
    _fields_ = [
        ("window", ctypes.c_uint32),
        ("k", ctypes.c_uint32),
        ("s", ctypes.c_uint32),
        ("log2_cache_bits", ctypes.c_uint32),
        ("log2_ext_bits", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("sbox", ctypes.c_uint64 * 256),
    ]


class Timings(ctypes.Structure):
    _fields_ = [
        ("build_ms", ctypes.c_float),
        ("scan_ms", ctypes.c_float),
        ("scan_kernel_ms", ctypes.c_float),
        ("pattern_pass_ms", ctypes.c_float),
        ("verify_ms", ctypes.c_float),
        ("scan_mode", ctypes.c_uint32),
        ("fold_log2_bits", ctypes.c_uint32),
        ("probes", ctypes.c_uint32),
        ("block", ctypes.c_uint32),
        ("ingest_ms", ctypes.c_float),
        ("l2_fold_log2_bits", ctypes.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


MATCH_DTYPE = np.dtype([("pattern_id", "<u4"), ("reserved", "<u4"), ("offset", "<u8")])
BATCH_MATCH_DTYPE = np.dtype([("text_id", "<u4"), ("pattern_id", "<u4"), ("offset", "<u8")])


class BatchTuning(ctypes.Structure):
    _fields_ = [
        ("chunk_iters", ctypes.c_uint32),
        ("probes", ctypes.c_uint32),
        ("block", ctypes.c_uint32),
        ("use_graph", ctypes.c_uint32),
        ("bytes_per_lane", ctypes.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}

_c = ctypes
_P = _c.c_void_p
_u32, _u64, _i32 = _c.c_uint32, _c.c_uint64, _c.c_int
_pu64 = _c.POINTER(_c.c_uint64)

# name: (restype, argtypes) -- every symbol include/ffbf.h and include/ffbf_workload.h declare
ABI = {
    "ffbf_last_error": (_c.c_char_p, []),
    "ffbf_version": (_c.c_char_p, []),
    "ffbf_create": (_i32, [_c.POINTER(Params), _i32, _c.POINTER(_P)]),
    "ffbf_create_split": (_i32, [_c.POINTER(SplitParams), _i32, _c.POINTER(_P)]),
    "ffbf_destroy": (None, [_P]),
    "ffbf_set_patterns": (_i32, [_P, _P, _P, _u64]),
    "ffbf_set_patterns_lines": (_i32, [_P, _c.c_char_p, _u64]),
    "ffbf_build": (_i32, [_P]),
    "ffbf_scan": (_i32, [_P, _P, _u64]),
    "ffbf_scan_device": (_i32, [_P, _P, _u64]),
    "ffbf_scan_fd": (_i32, [_P, _i32]),
    "ffbf_scan_file": (_i32, [_P, _c.c_char_p]),
    "ffbf_run_file": (_i32, [_P, _c.c_char_p, _pu64]),
    "ffbf_set_ingest": (_i32, [_P, _u64, _u32, _u32]),
    "ffbf_write_survivors": (_i32, [_P, _c.c_char_p, _c.c_char_p]),
    "ffbf_write_candidate_lines": (_i32, [_P, _P, _u64, _c.c_char_p]),
    "ffbf_run_device": (_i32, [_P, _P, _u64, _pu64]),
    "ffbf_run_device_noalloc": (_i32, [_P, _P, _u64, _P, _u64, _pu64]),
    "ffbf_pattern_pass": (_i32, [_P, _pu64]),
    "ffbf_verify": (_i32, [_P, _pu64]),
    "ffbf_get_matches": (_i32, [_P, _P, _u64, _pu64]),
    "ffbf_match_text": (_i32, [_P, _P, _u64, _c.POINTER(_P), _pu64]),
    "ffbf_free_matches": (None, [_P]),
    "ffbf_get_filter": (_i32, [_P, _i32, _P, _u64]),
    "ffbf_query_line": (_i32, [_P, _P, _u64, _c.POINTER(_c.c_int)]),
    "ffbf_query_lines": (_i32, [_P, _P, _P, _u64, _P]),
    "ffbf_get_candidates": (_i32, [_P, _P, _u64, _pu64]),
    "ffbf_get_survivors": (_i32, [_P, _P, _u64, _pu64]),
    "ffbf_filter_save": (_i32, [_P, _P, _u64, _pu64]),
    "ffbf_filter_load": (_i32, [_P, _P, _u64]),
    "ffbf_set_scan_mode": (_i32, [_P, _i32, _u32]),
    "ffbf_set_scan_tuning": (_i32, [_P, _u32, _u32]),
    "ffbf_set_build_method": (_i32, [_P, _i32]),
    "ffbf_set_timing": (_i32, [_P, _i32]),
    "ffbf_batch_set_timing": (_i32, [_P, _i32]),
    "ffbf_gather_bench": (_i32, [_i32, _u32, _u32, _c.POINTER(_c.c_float), _c.POINTER(_c.c_double)]),
    "ffbf_probe_bench": (_i32, [_i32, _u32, _u32, _u32, _c.POINTER(_c.c_float), _c.POINTER(_c.c_double)]),
    "ffbf_get_timings": (_i32, [_P, _c.POINTER(Timings)]),
    "ffbf_synchronize": (_i32, [_P]),
    "ffbf_get_stream": (_P, [_P]),
    "ffbf_device_alloc": (_i32, [_i32, _u64, _c.POINTER(_P)]),
    "ffbf_device_free": (_i32, [_P]),
    "ffbf_device_upload": (_i32, [_P, _P, _u64]),
    "ffbf_device_download": (_i32, [_P, _P, _u64]),
    "ffbf_device_count": (_i32, [_c.POINTER(_i32)]),
    "ffbf_device_copy": (_i32, [_P, _P, _u64]),
    "ffbf_host_alloc": (_i32, [_u64, _c.POINTER(_P)]),
    "ffbf_host_free": (_i32, [_P]),
    "ffbf_batch_create": (_i32, [_P, _u32, _c.POINTER(_P)]),
    "ffbf_batch_destroy": (None, [_P]),
    "ffbf_batch_set_texts": (_i32, [_P, _P, _P, _P, _u32]),
    "ffbf_batch_run": (_i32, [_P, _pu64]),
    "ffbf_batch_run_per_text": (_i32, [_P, _pu64]),
    "ffbf_batch_get_matches": (_i32, [_P, _P, _u64, _pu64]),
    "ffbf_batch_set_tuning": (_i32, [_P, _c.POINTER(BatchTuning)]),
    "ffbf_batch_get_tuning": (_i32, [_P, _c.POINTER(BatchTuning)]),
    "ffbf_batch_autotune": (_i32, [_P, _c.POINTER(BatchTuning), _c.POINTER(_c.c_float)]),
    "ffbf_batch_timings": (_i32, [_P, _c.POINTER(_c.c_float), _c.POINTER(_c.c_float)]),
    "ffbf_server_create": (_i32, [_P, _u32, _P, _c.POINTER(_P)]),
    "ffbf_server_destroy": (None, [_P]),
    "ffbf_server_submit": (_i32, [_P, _u32, _P, _u64, _pu64]),
    "ffbf_server_submit_many": (_i32, [_P, _u32, _P, _P, _P, _P]),
    "ffbf_server_cancel": (_i32, [_P, _u64]),
    "ffbf_server_poll": (_i32, [_P, _u64, _c.POINTER(_c.c_int)]),
    "ffbf_server_wait": (_i32, [_P, _u64, _c.POINTER(_P), _pu64, _c.POINTER(_c.c_double)]),
    "ffbf_server_get_stats": (_i32, [_P, _P]),
    # workload generators (host only)
    "ffbf_gen_bytes": (_i32, [_u32, _u64, _u64, _u64, _P, _i32]),
    "ffbf_gen_pattern_offsets": (_i32, [_u64, _u64, _u32, _u32, _P]),
    "ffbf_gen_pattern_bytes": (_i32, [_u32, _u64, _u64, _P, _i32]),
    "ffbf_gen_plants": (_i32, [_u64, _u64, _P, _u64, _P, _P, _u64, _P, _P]),
    "ffbf_gen_tandem_repeats": (_i32, [_u64, _P, _u64, _u32, _u32, _u32, _u32, _u32, _i32]),
    "ffbf_gen_sbox": (_i32, [_u64, _P]),
    "ffbf_gen_u64": (_u64, [_u64, _u64, _u64]),
    "ffbf_gen_lognormal_sizes": (_i32, [_u64, _u64, _c.c_double, _c.c_double, _u64, _u64, _P]),
}

for _name, (_res, _args) in ABI.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def lib() -> ctypes.CDLL:
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data if a.size else 0


def _check(status: int) -> None:
    if status != OK:
        msg = _lib.ffbf_last_error()
        raise FFBFError(status, msg.decode() if msg else "")


def device_count() -> int:
    n = _c.c_int(0)
    _check(_lib.ffbf_device_count(_c.byref(n)))
    return n.value


def gather_bench(log2_bytes: int, steps: int = 64, device: int = 0):
    """(ms, probes/s) of random uint32 gathers into a 2^log2_bytes array."""
    ms = _c.c_float(0)
    rate = _c.c_double(0)
    _check(_lib.ffbf_gather_bench(device, log2_bytes, steps, _c.byref(ms), _c.byref(rate)))
    return ms.value, rate.value


(PROBE_LOAD, PROBE_LOAD_NONTEMPORAL, PROBE_LOAD_AGENT, PROBE_LOAD_SYSTEM, PROBE_ATOMIC_OR, PROBE_STREAM,
 PROBE_UPLOAD) = range(7)
PROBE_KIND_NAMES = {PROBE_LOAD: "load", PROBE_LOAD_NONTEMPORAL: "non-temporal load",
                    PROBE_LOAD_AGENT: "agent-scope atomic load", PROBE_LOAD_SYSTEM: "system-scope atomic load",
                    PROBE_ATOMIC_OR: "device-scope atomicOr", PROBE_STREAM: "streaming read+write pass",
                    PROBE_UPLOAD: "pinned-host upload by the server's upload kernel"}


def probe_bench(kind: int, log2_bytes: int, steps: int = 64, device: int = 0):
    """(ms, rate) of random accesses of one kind (PROBE_*) into a
    2^log2_bytes array, in accesses/s; for PROBE_STREAM, `steps` passes that
    read and write back the whole array, in bytes/s; for PROBE_UPLOAD, the
    time of one upload of a 2^log2_bytes pinned buffer and bytes/s."""
    ms, rate = _c.c_float(0.0), _c.c_double(0.0)
    _check(_lib.ffbf_probe_bench(device, log2_bytes, steps, kind, _c.byref(ms), _c.byref(rate)))
    return ms.value, rate.value


# ---------------------------------------------------------------- workloads
def sbox_from_seed(seed: int) -> np.ndarray:
    out = np.zeros(256, dtype=np.uint32)
    _check(_lib.ffbf_gen_sbox(seed, _ptr(out)))
    return out


def gen_bytes(alphabet: int, seed: int, n: int, start: int = 0, threads: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    _check(_lib.ffbf_gen_bytes(alphabet, seed, start, n, _ptr(out), threads))
    return out


def gen_patterns(alphabet: int, seed: int, n: int, lmin: int, lmax: int, threads: int = 0):
    """Pattern set as (bytes uint8[], offsets uint64[n+1])."""
    off = np.empty(n + 1, dtype=np.uint64)
    _check(_lib.ffbf_gen_pattern_offsets(seed, n, lmin, lmax, _ptr(off)))
    total = int(off[-1])
    b = np.empty(total, dtype=np.uint8)
    _check(_lib.ffbf_gen_pattern_bytes(alphabet, seed, total, _ptr(b), threads))
    return b, off


def plant(seed: int, n_plants: int, text: np.ndarray, pat_bytes: np.ndarray, pat_off: np.ndarray):
    """Copies n_plants seeded patterns into text (in place); returns (ids, offsets)."""
    ids = np.empty(n_plants, dtype=np.uint32)
    offs = np.empty(n_plants, dtype=np.uint64)
    _check(
        _lib.ffbf_gen_plants(
            seed, n_plants, _ptr(text), text.size, _ptr(pat_bytes), _ptr(pat_off), pat_off.size - 1, _ptr(ids), _ptr(offs)
        )
    )
    return ids, offs


def tandem_repeats(seed: int, text: np.ndarray, block: int = 4096, min_len: int = 205, max_len: int = 614,
                   min_period: int = 1, max_period: int = 6, threads: int = 0) -> None:
    _check(_lib.ffbf_gen_tandem_repeats(seed, _ptr(text), text.size, block, min_len, max_len, min_period, max_period,
                                        threads))


def lognormal_sizes(seed: int, n: int, median: float, sigma: float, lo: int, hi: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    _check(_lib.ffbf_gen_lognormal_sizes(seed, n, median, sigma, lo, hi, _ptr(out)))
    return out


def gen_u64(seed: int, tag: int, i: int) -> int:
    return int(_lib.ffbf_gen_u64(seed, tag, i))


def patterns_from_list(patterns: Sequence[bytes]):
    off = np.zeros(len(patterns) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(p) for p in patterns], dtype=np.uint64) if patterns else []
    b = np.frombuffer(b"".join(patterns), dtype=np.uint8).copy()
    return b, off


# ---------------------------------------------------------------- device memory
class DeviceBuffer:
    """Plain device allocation on one HIP device (plumbing for tests/bench)."""

    def __init__(self, nbytes: int, device: int = 0):
        p = _c.c_void_p(0)
        _check(_lib.ffbf_device_alloc(device, nbytes, _c.byref(p)))
        self.ptr = p.value or 0
        self.nbytes = nbytes

    @classmethod
    def from_host(cls, a: np.ndarray, device: int = 0) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes, device)
        buf.upload(a)
        return buf

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        if offset + a.nbytes > self.nbytes:
            raise ValueError("upload past the end of the buffer")
        _check(_lib.ffbf_device_upload(self.ptr + offset, _ptr(a), a.nbytes))

    def download(self, nbytes: int | None = None, offset: int = 0) -> np.ndarray:
        nbytes = self.nbytes - offset if nbytes is None else nbytes
        out = np.empty(nbytes, dtype=np.uint8)
        _check(_lib.ffbf_device_download(_ptr(out), self.ptr + offset, nbytes))
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: int) -> None:
        """Device-to-device copy with the library's plain copy kernel."""
        _check(_lib.ffbf_device_copy(self.ptr, other.ptr, nbytes))

    def free(self) -> None:
        if self.ptr:
            _lib.ffbf_device_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedBuffer:
    """Page-locked host bytes with a numpy view (uploads at full link speed)."""

    def __init__(self, nbytes: int):
        p = _c.c_void_p(0)
        _check(_lib.ffbf_host_alloc(nbytes, _c.byref(p)))
        self.ptr = p.value or 0
        self.nbytes = nbytes
        raw = (_c.c_uint8 * max(nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(raw, dtype=np.uint8)[:nbytes]

    def free(self) -> None:
        if self.ptr:
            self.array = None
            _lib.ffbf_host_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- matcher
@dataclass
class Result:
    matches: np.ndarray  # MATCH_DTYPE, sorted by (pattern_id, offset)

    @property
    def pairs(self) -> list:
        return list(zip(self.matches["pattern_id"].tolist(), self.matches["offset"].tolist()))


class Matcher:
    """One resident filter pair (F, R) and pattern set on one GPU."""

    def __init__(self, window: int, k: int, log2_bits: int, sbox: Iterable[int], device: int = 0):
        sb = np.asarray(list(sbox) if not isinstance(sbox, np.ndarray) else sbox, dtype=np.uint32)
        if sb.shape != (256,):
            raise ValueError("sbox must have 256 entries")
        p = Params()
        p.window, p.k, p.log2_bits, p.flags = window, k, log2_bits, 0
        for i in range(256):
            p.sbox[i] = int(sb[i])
        h = _c.c_void_p(0)
        _check(_lib.ffbf_create(_c.byref(p), device, _c.byref(h)))
        self._h = h
        self.window, self.k, self.log2_bits = window, k, log2_bits
        self.n_words = 1 << (log2_bits - 5)
        self.n_patterns = 0

    @classmethod
    def split(cls, window: int, k: int, s: int, log2_cache_bits: int, log2_ext_bits: int, sbox64: Iterable[int],
              device: int = 0) -> "Matcher":
        """A context of the CPU binary's split filter (ffbf_create_split):
        64-bit sbox hash, the first s of k functions in a 2^c-bit cache part,
        the rest in a 2^e-bit external part, lines hashed separately."""
        sb = np.asarray(list(sbox64) if not isinstance(sbox64, np.ndarray) else sbox64, dtype=np.uint64)
        if sb.shape != (256,):
            raise ValueError("sbox must have 256 entries")
        p = SplitParams()
        p.window, p.k, p.s, p.flags = window, k, s, 0
        p.log2_cache_bits, p.log2_ext_bits = log2_cache_bits, log2_ext_bits
        for i in range(256):
            p.sbox[i] = int(sb[i])
        h = _c.c_void_p(0)
        _check(_lib.ffbf_create_split(_c.byref(p), device, _c.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        self.window, self.k, self.log2_bits = window, k, 0
        self.n_words = ((1 << log2_cache_bits) + (1 << log2_ext_bits)) // 32
        self.n_patterns = 0
        return self

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.ffbf_destroy(self._h)
            self._h = _c.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- patterns / build
    def set_patterns(self, pat_bytes: np.ndarray, pat_off: np.ndarray) -> None:
        pat_bytes = np.ascontiguousarray(pat_bytes, dtype=np.uint8)
        pat_off = np.ascontiguousarray(pat_off, dtype=np.uint64)
        _check(_lib.ffbf_set_patterns(self._h, _ptr(pat_bytes), _ptr(pat_off), pat_off.size - 1))
        self.n_patterns = pat_off.size - 1

    def set_pattern_list(self, patterns: Sequence[bytes]) -> None:
        self.set_patterns(*patterns_from_list(patterns))

    def set_patterns_lines(self, text: bytes) -> None:
        _check(_lib.ffbf_set_patterns_lines(self._h, text, len(text)))
        self.n_patterns = text.count(b"\n") + (0 if (not text or text.endswith(b"\n")) else 1)

    def build(self) -> None:
        _check(_lib.ffbf_build(self._h))

    def set_scan_mode(self, mode: int, fold_log2_bits: int = 0) -> None:
        _check(_lib.ffbf_set_scan_mode(self._h, mode, fold_log2_bits))

    def set_scan_tuning(self, probes: int = 0, block: int = 0) -> None:
        _check(_lib.ffbf_set_scan_tuning(self._h, probes, block))

    def set_build_method(self, method: int) -> None:
        """BUILD_AUTO, BUILD_ATOMIC or BUILD_BINNED, from the next build()."""
        _check(_lib.ffbf_set_build_method(self._h, method))

    def set_timing(self, on: bool) -> None:
        """Stage timing events (default on; each costs ~4-5 us between kernels)."""
        _check(_lib.ffbf_set_timing(self._h, 1 if on else 0))

    # -- scan / pass / verify
    def scan(self, text) -> int:
        t = np.frombuffer(text, dtype=np.uint8) if isinstance(text, (bytes, bytearray)) else np.ascontiguousarray(text)
        _check(_lib.ffbf_scan(self._h, _ptr(t), t.size))
        return self.n_candidates()

    def scan_device(self, ptr: int, n: int) -> int:
        _check(_lib.ffbf_scan_device(self._h, ptr, n))
        return self.n_candidates()

    def scan_fd(self, fd: int) -> int:
        """Scan everything readable from fd (a file from its position, a pipe)."""
        _check(_lib.ffbf_scan_fd(self._h, fd))
        return self.n_candidates()

    def scan_file(self, path) -> int:
        _check(_lib.ffbf_scan_file(self._h, os.fsencode(path)))
        return self.n_candidates()

    def run_file(self, path) -> int:
        nm = _c.c_uint64(0)
        _check(_lib.ffbf_run_file(self._h, os.fsencode(path), _c.byref(nm)))
        return nm.value

    def set_ingest(self, chunk_bytes: int = 0, depth: int = 0, threads: int = 0) -> None:
        _check(_lib.ffbf_set_ingest(self._h, chunk_bytes, depth, threads))

    def write_survivors(self, patterns_path, index_path=None) -> None:
        """Surviving patterns (one per line) and their 0-based ids, as the
        reference's filtered-phrases and __index_ files."""
        _check(_lib.ffbf_write_survivors(self._h, os.fsencode(patterns_path),
                                         None if index_path is None else os.fsencode(index_path)))

    def write_candidate_lines(self, out_path, text=None) -> None:
        """Lines of the scanned text holding a candidate window (the filtered
        corpus); text: host copy of the scanned text, else read back."""
        if text is None:
            _check(_lib.ffbf_write_candidate_lines(self._h, None, 0, os.fsencode(out_path)))
        else:
            t = np.frombuffer(text, dtype=np.uint8) if isinstance(text, (bytes, bytearray)) else np.ascontiguousarray(text)
            _check(_lib.ffbf_write_candidate_lines(self._h, _ptr(t), t.size, os.fsencode(out_path)))

    def run_device(self, ptr: int, n: int) -> int:
        nm = _c.c_uint64(0)
        _check(_lib.ffbf_run_device(self._h, ptr, n, _c.byref(nm)))
        return nm.value

    def pattern_pass(self) -> int:
        n = _c.c_uint64(0)
        _check(_lib.ffbf_pattern_pass(self._h, _c.byref(n)))
        return n.value

    def verify(self) -> int:
        n = _c.c_uint64(0)
        _check(_lib.ffbf_verify(self._h, _c.byref(n)))
        return n.value

    def match(self, text) -> Result:
        t = np.frombuffer(text, dtype=np.uint8) if isinstance(text, (bytes, bytearray)) else np.ascontiguousarray(text)
        out = _c.c_void_p(0)
        n = _c.c_uint64(0)
        _check(_lib.ffbf_match_text(self._h, _ptr(t), t.size, _c.byref(out), _c.byref(n)))
        try:
            if n.value:
                raw = (_c.c_uint8 * (n.value * MATCH_DTYPE.itemsize)).from_address(out.value)
                arr = np.frombuffer(raw, dtype=MATCH_DTYPE).copy()
            else:
                arr = np.zeros(0, dtype=MATCH_DTYPE)
        finally:
            _lib.ffbf_free_matches(out)
        return Result(arr)

    # -- line queries (Filter::operator[])
    def query_line(self, line) -> bool:
        t = np.frombuffer(bytes(line), dtype=np.uint8) if isinstance(line, (bytes, bytearray)) else np.ascontiguousarray(line)
        h = _c.c_int(0)
        _check(_lib.ffbf_query_line(self._h, _ptr(t), t.size, _c.byref(h)))
        return bool(h.value)

    def query_lines(self, line_bytes: np.ndarray, line_off: np.ndarray) -> np.ndarray:
        """uint8 0/1 per line: line i is line_bytes[line_off[i]:line_off[i+1]]."""
        line_bytes = np.ascontiguousarray(line_bytes, dtype=np.uint8)
        line_off = np.ascontiguousarray(line_off, dtype=np.uint64)
        n = line_off.size - 1
        hits = np.zeros(max(n, 1), dtype=np.uint8)
        _check(_lib.ffbf_query_lines(self._h, _ptr(line_bytes), _ptr(line_off), n, _ptr(hits)))
        return hits[:n]

    # -- results
    def matches(self) -> np.ndarray:
        n = _c.c_uint64(0)
        st = _lib.ffbf_get_matches(self._h, None, 0, _c.byref(n))
        if st not in (OK, ECAPACITY):
            _check(st)
        out = np.zeros(n.value, dtype=MATCH_DTYPE)
        _check(_lib.ffbf_get_matches(self._h, _ptr(out), n.value, _c.byref(n)))
        return out

    def n_candidates(self) -> int:
        n = _c.c_uint64(0)
        st = _lib.ffbf_get_candidates(self._h, None, 0, _c.byref(n))
        if st not in (OK, ECAPACITY):
            _check(st)
        return n.value

    def candidates(self) -> np.ndarray:
        n = self.n_candidates()
        out = np.zeros(n, dtype=np.uint64)
        m = _c.c_uint64(0)
        _check(_lib.ffbf_get_candidates(self._h, _ptr(out), n, _c.byref(m)))
        return out

    def survivors(self) -> np.ndarray:
        words = (self.n_patterns + 31) // 32
        out = np.zeros(max(words, 1), dtype=np.uint32)
        n = _c.c_uint64(0)
        _check(_lib.ffbf_get_survivors(self._h, _ptr(out), words, _c.byref(n)))
        return out[:words]

    def filter(self, which: int = FILTER_FORWARD) -> np.ndarray:
        """F, R, or one of the scan's folds (FILTER_FOLD, FILTER_L2_FOLD)."""
        words = self.n_words
        if which in (FILTER_FOLD, FILTER_L2_FOLD):
            t = self.timings()
            bits = t["fold_log2_bits"] if which == FILTER_FOLD else t["l2_fold_log2_bits"]
            words = max(1, (1 << bits) // 32) if bits else 1
        out = np.zeros(words, dtype=np.uint32)
        _check(_lib.ffbf_get_filter(self._h, which, _ptr(out), words))
        return out

    def filter_save(self) -> np.ndarray:
        n = _c.c_uint64(0)
        st = _lib.ffbf_filter_save(self._h, None, 0, _c.byref(n))
        if st not in (OK, ECAPACITY):
            _check(st)
        out = np.zeros(n.value, dtype=np.uint32)
        _check(_lib.ffbf_filter_save(self._h, _ptr(out), n.value, _c.byref(n)))
        return out

    def filter_load(self, deltas: np.ndarray) -> None:
        d = np.ascontiguousarray(deltas, dtype=np.uint32)
        _check(_lib.ffbf_filter_load(self._h, _ptr(d), d.size))

    def timings(self) -> dict:
        t = Timings()
        _check(_lib.ffbf_get_timings(self._h, _c.byref(t)))
        return t.as_dict()

    def synchronize(self) -> None:
        _check(_lib.ffbf_synchronize(self._h))

    @property
    def stream(self) -> int:
        return _lib.ffbf_get_stream(self._h) or 0


class Batch:
    """Many device texts, each routed to one of several resident Matchers,
    scanned in one persistent launch (replayed from a captured hipGraph)."""

    def __init__(self, matchers: Sequence[Matcher]):
        self._matchers = list(matchers)  # keep the contexts alive
        arr = (_c.c_void_p * len(self._matchers))(*[m._h.value for m in self._matchers])
        h = _c.c_void_p(0)
        _check(_lib.ffbf_batch_create(arr, len(self._matchers), _c.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.ffbf_batch_destroy(self._h)
            self._h = _c.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_texts(self, dev_ptr: int, offsets: np.ndarray, filters: np.ndarray) -> None:
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        fil = np.ascontiguousarray(filters, dtype=np.uint32)
        if off.size != fil.size + 1:
            raise ValueError("offsets needs one more entry than filters")
        _check(_lib.ffbf_batch_set_texts(self._h, dev_ptr, _ptr(off), _ptr(fil), fil.size))

    def run(self) -> int:
        n = _c.c_uint64(0)
        _check(_lib.ffbf_batch_run(self._h, _c.byref(n)))
        return n.value

    def run_per_text(self) -> int:
        n = _c.c_uint64(0)
        _check(_lib.ffbf_batch_run_per_text(self._h, _c.byref(n)))
        return n.value

    def matches(self) -> np.ndarray:
        n = _c.c_uint64(0)
        st = _lib.ffbf_batch_get_matches(self._h, None, 0, _c.byref(n))
        if st not in (OK, ECAPACITY):
            _check(st)
        out = np.zeros(n.value, dtype=BATCH_MATCH_DTYPE)
        _check(_lib.ffbf_batch_get_matches(self._h, _ptr(out), n.value, _c.byref(n)))
        return out

    def set_tuning(self, chunk_iters: int = 0, probes: int = 0, block: int = 0, use_graph: bool = True,
                   bytes_per_lane: int = 0) -> None:
        t = BatchTuning(chunk_iters, probes, block, 1 if use_graph else 0, bytes_per_lane)
        _check(_lib.ffbf_batch_set_tuning(self._h, _c.byref(t)))

    def tuning(self) -> dict:
        t = BatchTuning()
        _check(_lib.ffbf_batch_get_tuning(self._h, _c.byref(t)))
        return t.as_dict()

    def autotune(self):
        t = BatchTuning()
        ms = _c.c_float(0)
        _check(_lib.ffbf_batch_autotune(self._h, _c.byref(t), _c.byref(ms)))
        return t.as_dict(), ms.value

    def set_timing(self, on: bool) -> None:
        """Timing events around each run (default on; ~10 us per run)."""
        _check(_lib.ffbf_batch_set_timing(self._h, 1 if on else 0))

    def timings(self):
        run = _c.c_float(0)
        scan = _c.c_float(0)
        _check(_lib.ffbf_batch_timings(self._h, _c.byref(run), _c.byref(scan)))
        return run.value, scan.value


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("max_batch_bytes", ctypes.c_uint64),
        ("slots", ctypes.c_uint32),
        ("max_batch_requests", ctypes.c_uint32),
        ("upload", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


UPLOAD_DEFAULT, UPLOAD_COPY, UPLOAD_KERNEL = 0, 1, 2


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("submitted", "completed", "rejected", "batches", "bytes", "pending", "in_flight_bytes",
                 "kernel_upload_bytes")]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Server:
    """Continuous-batching request loop over resident filters (copies of the
    given Matchers).  submit() returns a request id; wait() blocks for its
    matches and the submit-to-completion latency."""

    def __init__(self, matchers: Sequence[Matcher], hbm_budget_bytes: int = 0, slots: int = 0,
                 max_batch_requests: int = 0, max_batch_bytes: int = 0, upload: int = UPLOAD_DEFAULT):
        arr = (_c.c_void_p * len(matchers))(*[m._h.value for m in matchers])
        cfg = ServerConfig(hbm_budget_bytes, max_batch_bytes, slots, max_batch_requests, upload, 0)
        h = _c.c_void_p(0)
        _check(_lib.ffbf_server_create(arr, len(matchers), _c.byref(cfg), _c.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, filter_id: int, text) -> int:
        t = np.frombuffer(text, dtype=np.uint8) if isinstance(text, (bytes, bytearray)) else np.ascontiguousarray(text)
        rid = _c.c_uint64(0)
        _check(_lib.ffbf_server_submit(self._h, filter_id, _ptr(t), t.size, _c.byref(rid)))
        self._keep[rid.value] = t  # the text must outlive the request
        return rid.value

    def submit_many(self, requests) -> list:
        """requests: iterable of (filter id, uint8 array); one C call."""
        reqs = [(f, np.frombuffer(t, dtype=np.uint8) if isinstance(t, (bytes, bytearray)) else np.ascontiguousarray(t))
                for f, t in requests]
        n = len(reqs)
        fil = np.array([f for f, _ in reqs], dtype=np.uint32)
        ptrs = np.array([_ptr(t) for _, t in reqs], dtype=np.uint64)
        sizes = np.array([t.size for _, t in reqs], dtype=np.uint64)
        ids = np.zeros(n, dtype=np.uint64)
        _check(_lib.ffbf_server_submit_many(self._h, n, _ptr(fil), _ptr(ptrs), _ptr(sizes), _ptr(ids)))
        for rid, (_, t) in zip(ids.tolist(), reqs):
            self._keep[rid] = t
        return ids.tolist()

    def poll(self, rid: int) -> bool:
        d = _c.c_int(0)
        _check(_lib.ffbf_server_poll(self._h, rid, _c.byref(d)))
        return bool(d.value)

    def cancel(self, rid: int) -> bool:
        """True if the request will complete cancelled, False if it had
        completed already (wait() then returns its matches)."""
        st = _lib.ffbf_server_cancel(self._h, rid)
        if st == ESTATE:
            return False
        _check(st)
        return True

    def wait(self, rid: int):
        out = _c.c_void_p(0)
        n = _c.c_uint64(0)
        lat = _c.c_double(0)
        st = _lib.ffbf_server_wait(self._h, rid, _c.byref(out), _c.byref(n), _c.byref(lat))
        self._keep.pop(rid, None)
        _check(st)
        try:
            if n.value:
                raw = (_c.c_uint8 * (n.value * MATCH_DTYPE.itemsize)).from_address(out.value)
                arr = np.frombuffer(raw, dtype=MATCH_DTYPE).copy()
            else:
                arr = np.zeros(0, dtype=MATCH_DTYPE)
        finally:
            _lib.ffbf_free_matches(out)
        return arr, lat.value

    def stats(self) -> dict:
        st = ServerStats()
        _check(_lib.ffbf_server_get_stats(self._h, _c.byref(st)))
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.ffbf_server_destroy(self._h)
            self._h = _c.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
