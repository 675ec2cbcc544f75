"""ctypes bindings of libkite_gpu.so (include/kite_gpu.h) and libkite_synth.so.

The GPU library is required: every entry point raises when it is missing or
when no HIP device is present.  There is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import weakref
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

_gpu = None
_synth = None
# live contexts and pinned arrays: released by an exit hook before the
# interpreter (and, after it, the HIP runtime) shuts down
_live: "weakref.WeakSet" = weakref.WeakSet()


def _release_all() -> None:
    for obj in list(_live):
        try:
            obj._release()
        except Exception:  # pragma: no cover - best effort at exit
            pass


class KiteError(RuntimeError):
    """Error reported by the native engine (message from kite_gpu_last_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class CountStats(C.Structure):
    _fields_ = [
        ("input_bytes", C.c_uint64), ("kmers", C.c_uint64), ("unique", C.c_uint64),
        ("buckets", C.c_uint64), ("batches", C.c_uint32), ("split_levels", C.c_uint32),
        ("skipped_chunks", C.c_uint32), ("words", C.c_uint32), ("graph", C.c_uint32),
        ("d2h_engine", C.c_uint32), ("device_ms", C.c_double),
        ("extract_ms", C.c_double), ("split_ms", C.c_double), ("sort_ms", C.c_double),
        ("compact_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class CorrStats(C.Structure):
    _fields_ = [
        ("gram_ms_pearson", C.c_double), ("gram_ms_spearman", C.c_double),
        ("fp64_ms_pearson", C.c_double), ("fp64_ms_spearman", C.c_double),
        ("rows", C.c_uint64), ("cols", C.c_uint32), ("code_bytes", C.c_uint32),
        ("gram_ms_total", C.c_double), ("fp64_ms_total", C.c_double), ("passes", C.c_uint32),
        ("fp64_passes", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


CORR_PEARSON, CORR_SPEARMAN = 1, 2
CORR_FP64_BASELINE, CORR_LOLO = 1, 2

APPEND_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64,
                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint8),
                        C.c_size_t)

class BatchSample(C.Structure):
    _fields_ = [("fastq", C.c_void_p), ("len", C.c_size_t)]


BATCH_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.POINTER(CountStats), C.POINTER(C.c_uint64),
                       C.POINTER(C.c_uint32), C.POINTER(C.c_uint64))
BATCH_CHUNKS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.POINTER(CountStats), C.c_uint32,
                              C.POINTER(C.c_uint8), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                              C.POINTER(C.c_uint64))
BATCH_PLAIN, BATCH_PIPELINED, BATCH_SERIAL = 0, 1, 2


class DumpRun(C.Structure):
    _fields_ = [("partition", C.c_uint32), ("codec", C.c_uint32), ("rows", C.c_uint64),
                ("data", C.c_void_p), ("bytes", C.c_uint64)]


class DumpStats(C.Structure):
    _fields_ = [("runs", C.c_uint64), ("rows_in", C.c_uint64), ("rows", C.c_uint64), ("lines", C.c_uint64),
                ("bytes", C.c_uint64), ("groups", C.c_uint64), ("decode_ms", C.c_double),
                ("merge_ms", C.c_double), ("format_ms", C.c_double), ("copy_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


DUMP_WRITE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_char), C.c_uint64)


class MergeInput(C.Structure):
    _fields_ = [("runs", C.POINTER(DumpRun)), ("n_runs", C.c_uint64), ("has_min", C.c_int),
                ("min_count", C.c_uint32), ("has_max", C.c_int), ("max_count", C.c_uint32)]


class MergeStats(C.Structure):
    _fields_ = [("runs", C.c_uint64), ("rows_in", C.c_uint64), ("rows_kept", C.c_uint64), ("rows", C.c_uint64),
                ("bytes", C.c_uint64), ("groups", C.c_uint64), ("decode_ms", C.c_double),
                ("merge_ms", C.c_double), ("encode_ms", C.c_double), ("in_ms", C.c_double),
                ("out_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}
HISTO_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_uint64)

# name -> (restype, argtypes); every symbol include/kite_gpu.h declares
GPU_SYMBOLS = {
    "kite_gpu_abi_version": (C.c_int, []),
    "kite_gpu_last_error": (C.c_char_p, []),
    "kite_gpu_open": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "kite_gpu_close": (None, [C.c_void_p]),
    "kite_gpu_count_begin": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_uint32]),
    "kite_gpu_count_feed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "kite_gpu_count_feed_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "kite_gpu_count_feed_bgzf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "kite_gpu_inflate_bgzf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                        C.POINTER(C.c_size_t), C.POINTER(C.c_double)]),
    "kite_gpu_count_finish": (C.c_int, [C.c_void_p, C.POINTER(CountStats)]),
    "kite_gpu_count_emit": (C.c_int, [C.c_void_p, APPEND_FN, C.c_void_p]),
    "kite_gpu_count_result_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "kite_gpu_count_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kite_gpu_count_device_result": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p),
                                               C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "kite_gpu_merge_tables": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_void_p]),
    "kite_gpu_union_tables": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_void_p]),
    "kite_gpu_corr_counts": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                       C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.POINTER(CorrStats)]),
    "kite_gpu_analyze_tables": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                          C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.POINTER(CorrStats)]),
    "kite_gpu_pca": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                               C.c_void_p]),
    "kite_gpu_pca_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                       C.c_void_p]),
    "kite_gpu_count_batch": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_uint32,
                                       C.POINTER(BatchSample), C.c_uint32, C.c_int, BATCH_FN,
                                       C.c_void_p]),
    "kite_gpu_count_batch_chunks": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_uint32,
                                              C.POINTER(BatchSample), C.c_uint32, C.c_int, BATCH_CHUNKS_FN,
                                              C.c_void_p]),
    "kite_gpu_dump": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(DumpRun),
                                C.c_uint64, C.c_int, C.c_uint32, C.c_int, C.c_uint32, DUMP_WRITE_FN,
                                C.c_void_p, C.POINTER(DumpStats)]),
    "kite_gpu_histo": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(DumpRun), C.c_uint64, C.c_int,
                                 C.c_uint32, C.c_int, C.c_uint32, HISTO_FN, C.c_void_p, C.POINTER(DumpStats)]),
    "kite_gpu_merge_runs": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(MergeInput), C.c_uint32,
                                      APPEND_FN, C.c_void_p, C.POINTER(MergeStats)]),
    "kite_gpu_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_uint32, C.c_void_p, C.c_size_t,
                                    C.c_char_p, C.c_size_t]),
    "kite_gpu_set_tuned": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_int]),
    "kite_gpu_copy_peak": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "kite_gpu_host_alloc": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "kite_gpu_host_free": (None, [C.c_void_p]),
    "kite_gpu_count_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "kite_gpu_count_kernel_stats": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "kite_gpu_batch_timeline": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
}


def gpu_lib_path() -> Path:
    return LIB_DIR / "libkite_gpu.so"


def load_gpu() -> C.CDLL:
    """Load libkite_gpu.so (raises if it has not been built)."""
    global _gpu
    if _gpu is None:
        path = gpu_lib_path()
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build it with `python -m kite_amd._build`")
        lib = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
        for name, (res, args) in GPU_SYMBOLS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _gpu = lib
        # registered after torch's own hooks when torch is already imported,
        # so it runs before them (atexit runs hooks last-in first-out)
        atexit.register(_release_all)
    return _gpu


def _check(rc: int) -> None:
    if rc != 0:
        msg = load_gpu().kite_gpu_last_error()
        raise KiteError(rc, msg.decode() if msg else "unknown error")


def _ptr(buf) -> int:
    if isinstance(buf, np.ndarray):
        assert buf.flags["C_CONTIGUOUS"]
        return buf.ctypes.data
    if isinstance(buf, (bytes, bytearray)):
        return C.cast(C.c_char_p(bytes(buf)), C.c_void_p).value  # pragma: no cover
    raise TypeError(type(buf))


class PinnedArray:
    """Page-locked host bytes (kite_gpu_host_alloc); `.array` is a uint8 view."""

    def __init__(self, nbytes: int):
        lib = load_gpu()
        p = C.c_void_p()
        _check(lib.kite_gpu_host_alloc(nbytes, C.byref(p)))
        self._lib, self.ptr, self.nbytes = lib, p.value, nbytes
        self.array = np.ctypeslib.as_array((C.c_uint8 * max(nbytes, 1)).from_address(p.value))[:nbytes]
        _live.add(self)

    def free(self) -> None:
        if self.ptr:
            self.array = None
            self._lib.kite_gpu_host_free(C.c_void_p(self.ptr))
            self.ptr = None

    _release = free

    def __del__(self, _finalizing=sys.is_finalizing):  # pragma: no cover
        # no HIP call while the interpreter finalizes: the exit hook has
        # already released what was still live (the check is bound at
        # definition: module globals are gone late in shutdown)
        if _finalizing():
            return
        try:
            self.free()
        except Exception:
            pass


class CountResult:
    """Reduced count table: per-partition ascending unique keys and counts."""

    def __init__(self, keys: np.ndarray, counts: np.ndarray, part_off: np.ndarray, words: int,
                 k: int, canonical: bool):
        self.keys = keys            # (n_unique, words) uint64, word 0 least significant
        self.counts = counts        # (n_unique,) uint32
        self.part_off = part_off    # (partitions + 1,) uint64
        self.words = words
        self.k = k
        self.canonical = canonical

    @property
    def partitions(self) -> int:
        return len(self.part_off) - 1

    def __len__(self) -> int:
        return len(self.counts)


class GpuContext:
    """One engine context on one HIP device (kite_gpu_open / kite_gpu_close)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        lib = load_gpu()
        h = C.c_void_p()
        _check(lib.kite_gpu_open(device, hbm_budget_bytes, C.byref(h)))
        self._h = h
        self._lib = lib
        _live.add(self)

    def close(self) -> None:
        if self._h:
            self._lib.kite_gpu_close(self._h)
            self._h = None

    _release = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self, _finalizing=sys.is_finalizing):  # pragma: no cover
        # (bound at definition: module globals are gone late in shutdown)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    # -- count session (kite_gpu_count_*) --------------------------------
    def count_begin(self, k: int, canonical: bool, partitions: int) -> None:
        _check(self._lib.kite_gpu_count_begin(self._h, k, int(bool(canonical)), partitions))
        self._k, self._canonical = k, bool(canonical)

    def count_feed(self, chunk: np.ndarray | bytes) -> None:
        arr = np.frombuffer(chunk, dtype=np.uint8) if isinstance(chunk, (bytes, bytearray)) else chunk
        arr = np.ascontiguousarray(arr, dtype=np.uint8)
        _check(self._lib.kite_gpu_count_feed(self._h, arr.ctypes.data, arr.size))

    def count_feed_device(self, dev_ptr: int, nbytes: int) -> None:
        _check(self._lib.kite_gpu_count_feed_device(self._h, C.c_void_p(dev_ptr), nbytes))

    def count_feed_bgzf(self, data) -> int:
        """Feed BGZF bytes (whole members are consumed); returns the bytes
        consumed."""
        arr = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data
        used = C.c_size_t(0)
        _check(self._lib.kite_gpu_count_feed_bgzf(self._h, arr.ctypes.data if arr.size else None, arr.size,
                                                  C.byref(used)))
        return used.value

    def inflate_bgzf(self, data) -> tuple[bytes, float]:
        """Inflate the whole BGZF members of `data` on the device."""
        arr = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data
        need = C.c_size_t(0)
        ms = C.c_double(0)
        rc = self._lib.kite_gpu_inflate_bgzf(self._h, arr.ctypes.data if arr.size else None, arr.size, None, 0,
                                             C.byref(need), None)
        if rc != 0 and need.value == 0:
            _check(rc)
        out = np.empty(max(need.value, 1), dtype=np.uint8)
        _check(self._lib.kite_gpu_inflate_bgzf(self._h, arr.ctypes.data if arr.size else None, arr.size,
                                               out.ctypes.data, out.size, C.byref(need), C.byref(ms)))
        return out[:need.value].tobytes(), ms.value

    def count_finish(self) -> CountStats:
        st = CountStats()
        _check(self._lib.kite_gpu_count_finish(self._h, C.byref(st)))
        return st

    def count_download(self) -> CountResult:
        n = C.c_uint64()
        w = C.c_uint32()
        p = C.c_uint32()
        _check(self._lib.kite_gpu_count_result_size(self._h, C.byref(n), C.byref(w), C.byref(p)))
        keys = np.empty((n.value, w.value), dtype=np.uint64)
        counts = np.empty(n.value, dtype=np.uint32)
        part_off = np.empty(p.value + 1, dtype=np.uint64)
        _check(self._lib.kite_gpu_count_download(self._h, keys.ctypes.data, counts.ctypes.data,
                                                 part_off.ctypes.data))
        return CountResult(keys, counts, part_off, w.value, self._k, self._canonical)

    def count_emit(self) -> list[tuple]:
        """Collect the encoded run chunks: (partition, codec, rows, raw_bytes,
        key_min[4], key_max[4], payload bytes)."""
        out: list[tuple] = []

        def cb(user, part, codec, rows, raw, kmin, kmax, data, ln):
            out.append((part, codec, rows, raw, tuple(kmin[i] for i in range(4)),
                        tuple(kmax[i] for i in range(4)), C.string_at(data, ln)))
            return 0

        fn = APPEND_FN(cb)
        _check(self._lib.kite_gpu_count_emit(self._h, fn, None))
        return out

    # -- merge / union (kite_gpu_merge_tables, kite_gpu_union_tables) ----
    @staticmethod
    def _tables(tables):
        ks, cs, os_ = [], [], []
        for keys, counts, off in tables:
            ks.append(np.ascontiguousarray(keys, dtype=np.uint64))
            cs.append(np.ascontiguousarray(counts, dtype=np.uint64))
            os_.append(np.ascontiguousarray(off, dtype=np.uint64))
        arr = lambda xs: (C.c_void_p * len(xs))(*[x.ctypes.data for x in xs])
        return ks, cs, os_, arr(ks), arr(cs), arr(os_)

    def merge_tables(self, tables, words: int):
        """Sorted union with u64 count sums of [(keys, counts, part_off)]."""
        ks, cs, os_, kp, cp, op = self._tables(tables)
        P = len(os_[0]) - 1
        n = sum(int(o[-1]) for o in os_)
        ok = np.empty((max(n, 1), words), dtype=np.uint64)
        oc = np.empty(max(n, 1), dtype=np.uint64)
        oo = np.empty(P + 1, dtype=np.uint64)
        _check(self._lib.kite_gpu_merge_tables(self._h, words, P, len(tables), kp, cp, op,
                                               ok.ctypes.data, oc.ctypes.data, oo.ctypes.data))
        u = int(oo[-1])
        return ok[:u], oc[:u], oo

    def union_tables(self, tables, words: int):
        """Union keys, u32 count matrix (rows x tables) and offsets."""
        ks, cs, os_, kp, cp, op = self._tables(tables)
        P = len(os_[0]) - 1
        n = sum(int(o[-1]) for o in os_)
        ok = np.empty((max(n, 1), words), dtype=np.uint64)
        om = np.empty((max(n, 1), len(tables)), dtype=np.uint32)
        oo = np.empty(P + 1, dtype=np.uint64)
        _check(self._lib.kite_gpu_union_tables(self._h, words, P, len(tables), kp, cp, op,
                                               ok.ctypes.data, om.ctypes.data, oo.ctypes.data))
        u = int(oo[-1])
        return ok[:u], om[:u], oo

    # -- analyze (kite_gpu_corr_counts, kite_gpu_analyze_tables, kite_gpu_pca)
    def corr_counts(self, counts: np.ndarray, methods: int = 3, flags: int = 0) -> dict:
        """Correlation of the columns of an N x n u32 count matrix."""
        m = np.ascontiguousarray(counts, dtype=np.uint32)
        N, n = m.shape
        outs = {k: np.zeros((n, n)) for k in ("pearson", "spearman", "cov", "pearson_fp64",
                                               "spearman_fp64")}
        st = CorrStats()
        _check(self._lib.kite_gpu_corr_counts(
            self._h, m.ctypes.data if N else None, N, n, methods, flags,
            *[outs[k].ctypes.data for k in ("pearson", "spearman", "cov", "pearson_fp64",
                                             "spearman_fp64")], C.byref(st)))
        outs["stats"] = st.as_dict()
        return outs

    def analyze_tables(self, tables, words: int, methods: int = 3, flags: int = 0,
                       want_union: bool = False) -> dict:
        ks, cs, os_, kp, cp, op = self._tables(tables)
        P = len(os_[0]) - 1
        n = len(tables)
        rows = sum(int(o[-1]) for o in os_)
        uk = np.empty((max(rows, 1), words), dtype=np.uint64) if want_union else None
        um = np.empty((max(rows, 1), n), dtype=np.uint32) if want_union else None
        uo = np.empty(P + 1, dtype=np.uint64)
        outs = {k: np.zeros((n, n)) for k in ("pearson", "spearman", "cov", "pearson_fp64",
                                               "spearman_fp64")}
        st = CorrStats()
        _check(self._lib.kite_gpu_analyze_tables(
            self._h, words, P, n, kp, cp, op, methods, flags,
            uk.ctypes.data if want_union else None, um.ctypes.data if want_union else None,
            uo.ctypes.data,
            *[outs[k].ctypes.data for k in ("pearson", "spearman", "cov", "pearson_fp64",
                                             "spearman_fp64")], C.byref(st)))
        outs["stats"] = st.as_dict()
        outs["part_off"] = uo
        if want_union:
            u = int(uo[-1])
            outs["keys"], outs["matrix"] = uk[:u], um[:u]
        return outs

    def pca(self, cov: np.ndarray, dims: int) -> tuple[np.ndarray, np.ndarray]:
        c = np.ascontiguousarray(cov, dtype=np.float64)
        n = c.shape[0]
        scores = np.zeros((n, dims))
        ev = np.zeros(n)
        _check(self._lib.kite_gpu_pca(self._h, c.ctypes.data, n, dims, scores.ctypes.data,
                                      ev.ctypes.data))
        return scores, ev

    def pca_noalloc(self, cov: np.ndarray, dims: int) -> tuple[np.ndarray, np.ndarray]:
        """kite_gpu_pca_noalloc: the same PCA in fixed device memory (n <= 64)."""
        c = np.ascontiguousarray(cov, dtype=np.float64)
        n = c.shape[0]
        scores = np.zeros((n, dims))
        ev = np.zeros(n)
        _check(self._lib.kite_gpu_pca_noalloc(self._h, c.ctypes.data, n, dims, scores.ctypes.data,
                                              ev.ctypes.data))
        return scores, ev

    # -- batched count of many samples (kite_gpu_count_batch) ------------
    def count_batch(self, samples, k: int, canonical: bool, partitions: int,
                    mode: int = BATCH_PIPELINED, on_result=None, encoded: bool = False) -> list:
        """Count each FASTQ buffer in `samples` (uint8 arrays, ideally from
        PinnedArray) as its own table.  With on_result=None the tables are
        returned as [(CountResult, stats dict)]; otherwise
        on_result(i, stats dict, keys, counts, part_off) receives views that
        are only valid during the call, and the list holds its return values.

        encoded=True delivers chunk payloads encoded on the device instead:
        the default result is [(codec, bytes, byte_off, rows_off, minmax,
        stats dict)] and on_result(i, stats, codec, data, byte_off, rows_off,
        minmax) receives views valid during the call."""
        if encoded:
            return self._count_batch_chunks(samples, k, canonical, partitions, mode, on_result)
        arrs = [np.ascontiguousarray(a, dtype=np.uint8) for a in samples]
        descr = (BatchSample * max(len(arrs), 1))()
        for i, a in enumerate(arrs):
            descr[i].fastq = a.ctypes.data if a.size else None
            descr[i].len = a.size
        words = 1 if k <= 31 else (2 if k <= 64 else 4)
        P = partitions
        out: list = [None] * len(arrs)
        errors: list = []

        def cb(user, i, st, kp, cp, op):
            try:
                d = {f: getattr(st.contents, f) for f, _ in CountStats._fields_}
                u = d["unique"]
                keys = np.ctypeslib.as_array(kp, shape=(max(u, 1) * words,))[:u * words].reshape(u, words)
                counts = np.ctypeslib.as_array(cp, shape=(max(u, 1),))[:u]
                off = np.ctypeslib.as_array(op, shape=(P + 1,))
                if on_result is None:
                    out[i] = (CountResult(keys.copy(), counts.copy(), off.copy(), words, k, bool(canonical)), d)
                else:
                    out[i] = on_result(i, d, keys, counts, off)
                return 0
            except Exception as e:  # pragma: no cover - surfaced after the call
                errors.append(e)
                return 1

        fn = BATCH_FN(cb)
        rc = self._lib.kite_gpu_count_batch(self._h, k, int(bool(canonical)), P, descr, len(arrs), mode,
                                            fn, None)
        if errors:
            raise errors[0]
        _check(rc)
        return out

    def _count_batch_chunks(self, samples, k, canonical, partitions, mode, on_result):
        arrs = [np.ascontiguousarray(a, dtype=np.uint8) for a in samples]
        descr = (BatchSample * max(len(arrs), 1))()
        for i, a in enumerate(arrs):
            descr[i].fastq = a.ctypes.data if a.size else None
            descr[i].len = a.size
        P = partitions
        out: list = [None] * len(arrs)
        errors: list = []

        def cb(user, i, st, codec, dp, bo, ro, mm):
            try:
                d = {f: getattr(st.contents, f) for f, _ in CountStats._fields_}
                boff = np.ctypeslib.as_array(bo, shape=(P + 1,))
                total = int(boff[-1])
                data = np.ctypeslib.as_array(dp, shape=(max(total, 1),))[:total]
                rows = np.ctypeslib.as_array(ro, shape=(P + 1,))
                minmax = np.ctypeslib.as_array(mm, shape=(P * 8,)).reshape(P, 8)
                if on_result is None:
                    out[i] = (int(codec), data.tobytes(), boff.copy(), rows.copy(), minmax.copy(), d)
                else:
                    out[i] = on_result(i, d, int(codec), data, boff, rows, minmax)
                return 0
            except Exception as e:  # pragma: no cover - surfaced after the call
                errors.append(e)
                return 1

        fn = BATCH_CHUNKS_FN(cb)
        rc = self._lib.kite_gpu_count_batch_chunks(self._h, k, int(bool(canonical)), P, descr, len(arrs), mode,
                                                   fn, None)
        if errors:
            raise errors[0]
        _check(rc)
        return out

    def dump_runs(self, k: int, key_width: int, partitions: int, runs, min_count=None, max_count=None,
                  write=None) -> tuple[bytes, dict]:
        """`kite dump` of stored runs on the device (kite_gpu_dump).  runs:
        (partition, codec, rows, payload bytes) in partition order.  The
        text is returned, or passed piece by piece to write(memoryview)."""
        arr, keep = self._runs_array(runs)
        pieces: list = []
        errors: list = []

        def cb(user, text, n):
            try:
                chunk = C.string_at(text, n)
                if write is None:
                    pieces.append(chunk)
                else:
                    write(chunk)
                return 0
            except Exception as e:  # pragma: no cover - surfaced after the call
                errors.append(e)
                return 1

        fn = DUMP_WRITE_FN(cb)
        st = DumpStats()
        rc = self._lib.kite_gpu_dump(self._h, k, key_width, partitions, arr, len(runs),
                                     int(min_count is not None), int(min_count or 0),
                                     int(max_count is not None), int(max_count or 0), fn, None, C.byref(st))
        if errors:
            raise errors[0]
        _check(rc)
        return b"".join(pieces), st.as_dict()

    @staticmethod
    def _runs_array(runs):
        arr = (DumpRun * max(len(runs), 1))()
        keep = []
        for i, (part, codec, rows, data) in enumerate(runs):
            buf = np.frombuffer(bytes(data), dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
            keep.append(buf)
            arr[i].partition, arr[i].codec, arr[i].rows = part, codec, rows
            arr[i].data = buf.ctypes.data
            arr[i].bytes = len(data)
        return arr, keep

    def histo_runs(self, key_width: int, partitions: int, runs, min_count=None,
                   max_count=None) -> tuple[dict, dict]:
        """`kite histo` of one input's stored runs on the device
        (kite_gpu_histo): {total: frequency}, and the stats."""
        arr, keep = self._runs_array(runs)
        out: dict = {}

        def cb(user, vp, fp, n):
            for i in range(n):
                out[int(vp[i])] = int(fp[i])
            return 0

        fn = HISTO_FN(cb)
        st = DumpStats()
        rc = self._lib.kite_gpu_histo(self._h, key_width, partitions, arr, len(runs),
                                      int(min_count is not None), int(min_count or 0),
                                      int(max_count is not None), int(max_count or 0), fn, None, C.byref(st))
        del keep
        _check(rc)
        return out, st.as_dict()

    def merge_runs(self, key_width: int, partitions: int, inputs) -> tuple[list[tuple], dict]:
        """`kite merge` of stored runs on the device (kite_gpu_merge_runs).
        inputs: [(runs, min_count, max_count)], runs as for dump_runs.
        Returns the chunks (partition, codec, rows, raw_bytes, key_min[4],
        key_max[4], payload bytes) in partition order, and the stats."""
        arrs, keeps = [], []
        ins = (MergeInput * len(inputs))()
        for i, (runs, mn, mx) in enumerate(inputs):
            arr, keep = self._runs_array(runs)
            arrs.append(arr)
            keeps.append(keep)
            ins[i].runs = C.cast(arr, C.POINTER(DumpRun))
            ins[i].n_runs = len(runs)
            ins[i].has_min, ins[i].min_count = int(mn is not None), int(mn or 0)
            ins[i].has_max, ins[i].max_count = int(mx is not None), int(mx or 0)
        out: list[tuple] = []

        def cb(user, part, codec, rows, raw, kmin, kmax, data, ln):
            out.append((part, codec, rows, raw, tuple(kmin[i] for i in range(4)),
                        tuple(kmax[i] for i in range(4)), C.string_at(data, ln)))
            return 0

        fn = APPEND_FN(cb)
        st = MergeStats()
        rc = self._lib.kite_gpu_merge_runs(self._h, key_width, partitions, ins, len(inputs), fn, None, C.byref(st))
        del keeps, arrs
        _check(rc)
        return out, st.as_dict()

    def count_profile(self, enable: bool, timeline: bool = False) -> None:
        """Kernel events on the engine stream (enable) and per-sample stage
        events of a pipelined batch (timeline)."""
        _check(self._lib.kite_gpu_count_profile(self._h, (1 if enable else 0) | (2 if timeline else 0)))

    def count_kernel_stats(self) -> dict:
        import json
        buf = C.create_string_buffer(1 << 16)
        n = self._lib.kite_gpu_count_kernel_stats(self._h, buf, len(buf))
        if n < 0:
            _check(n)
        return json.loads(buf.value.decode())

    def autotune(self, fastq: np.ndarray, k: int, canonical: bool, partitions: int) -> dict:
        """kite_gpu_autotune on one sample: the chosen count settings, kept
        on this context for inputs of the sample's size class."""
        import json
        arr = np.ascontiguousarray(fastq, dtype=np.uint8)
        buf = C.create_string_buffer(1 << 16)
        _check(self._lib.kite_gpu_autotune(self._h, k, int(bool(canonical)), partitions, arr.ctypes.data, arr.size,
                                           buf, len(buf)))
        return json.loads(buf.value.decode())

    def set_tuned(self, k: int, partitions: int, size_class: int, chosen: dict) -> None:
        """kite_gpu_set_tuned: re-apply settings an autotune run chose
        (`chosen` as in its report) for one size class."""
        _check(self._lib.kite_gpu_set_tuned(self._h, k, partitions, size_class, int(chosen["split_target_div"]),
                                            int(chosen["max_digit_bits"]), int(chosen["hash_classes"]),
                                            int(chosen["extract_subs"]), int(chosen["split_above_div"])))

    def copy_peak(self, nbytes: int = 1 << 30) -> dict:
        """kite_gpu_copy_peak: page-locked H2D, D2H and both-at-once GB/s."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(self._lib.kite_gpu_copy_peak(self._h, nbytes, C.byref(a), C.byref(b), C.byref(c)))
        return {"h2d_GBps": a.value, "d2h_GBps": b.value, "both_GBps": c.value}

    def batch_timeline(self) -> np.ndarray:
        """Per-sample stage times of the last profiled pipelined batch, ms
        since its start: columns h2d0, h2d1, run0, run1, d2h0, d2h1."""
        n = self._lib.kite_gpu_batch_timeline(self._h, None, 0)
        if n < 0:
            _check(n)
        out = np.zeros((max(n, 1), 6))
        self._lib.kite_gpu_batch_timeline(self._h, out.ctypes.data, n)
        return out[:n]

    def count(self, fastq: np.ndarray | bytes, k: int, canonical: bool, partitions: int,
              chunks: list[tuple[int, int]] | None = None) -> tuple[CountResult, CountStats]:
        """Count a FASTQ image held in host memory.  `chunks` lists
        (offset, length) record-aligned pieces to feed; default: one chunk."""
        arr = np.frombuffer(fastq, dtype=np.uint8) if isinstance(fastq, (bytes, bytearray)) else fastq
        self.count_begin(k, canonical, partitions)
        if chunks is None:
            chunks = [(0, arr.size)]
        for off, ln in chunks:
            self.count_feed(arr[off:off + ln])
        st = self.count_finish()
        return self.count_download(), st


# ---- synthetic data ---------------------------------------------------------
class SynthSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("genome_len", C.c_uint64), ("read_len", C.c_uint32),
        ("both_strands", C.c_uint32), ("err_rate", C.c_double), ("n_rate", C.c_double),
        ("repeat_segments", C.c_uint32), ("repeat_seg_len", C.c_uint32),
        ("repeat_copies", C.c_uint32), ("_pad", C.c_uint32), ("repeat_div", C.c_double),
        ("variant_seed", C.c_uint64), ("snv_rate", C.c_double),
    ]


def load_synth() -> C.CDLL:
    global _synth
    if _synth is None:
        path = LIB_DIR / "libkite_synth.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build it with `python -m kite_amd._build`")
        lib = C.CDLL(str(path))
        lib.kite_synth_genome.restype = C.c_int
        lib.kite_synth_genome.argtypes = [C.POINTER(SynthSpec), C.c_void_p]
        lib.kite_synth_offset.restype = C.c_uint64
        lib.kite_synth_offset.argtypes = [C.POINTER(SynthSpec), C.c_uint64]
        lib.kite_synth_reads.restype = C.c_int64
        lib.kite_synth_reads.argtypes = [C.POINTER(SynthSpec), C.c_void_p, C.c_uint64, C.c_uint64,
                                         C.c_void_p, C.c_uint64, C.c_int]
        _synth = lib
    return _synth


def synth_genome(spec: SynthSpec) -> np.ndarray:
    g = np.empty(spec.genome_len, dtype=np.uint8)
    if load_synth().kite_synth_genome(C.byref(spec), g.ctypes.data) != 0:
        raise RuntimeError("genome generation failed")
    return g


def synth_reads(spec: SynthSpec, genome: np.ndarray, r0: int, r1: int,
                threads: int | None = None, out: np.ndarray | None = None) -> np.ndarray:
    lib = load_synth()
    need = lib.kite_synth_reads(C.byref(spec), genome.ctypes.data, r0, r1, None, 0, 1)
    if out is None:
        out = np.empty(need, dtype=np.uint8)
    assert out.size >= need
    threads = threads or min(16, os.cpu_count() or 1)
    n = lib.kite_synth_reads(C.byref(spec), genome.ctypes.data, r0, r1, out.ctypes.data, out.size,
                             threads)
    if n != need:
        raise RuntimeError("read generation failed")
    return out[:need]
