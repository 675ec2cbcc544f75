"""ctypes access to the CPU oracle (oracle/build/libkite_oracle.so).

Test infrastructure only: the oracle is the checker, never the thing
measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_SO = ORACLE_DIR / "build" / "libkite_oracle.so"
ORACLE_BIN = ORACLE_DIR / "build" / "kite_oracle"

_lib = None


class OkTable(C.Structure):
    _fields_ = [
        ("k", C.c_int), ("canonical", C.c_int), ("words", C.c_int), ("partitions", C.c_uint32),
        ("part_off", C.POINTER(C.c_uint64)), ("keys", C.POINTER(C.c_uint64)),
        ("counts", C.POINTER(C.c_uint64)), ("n_unique", C.c_uint64), ("n_kmers", C.c_uint64),
        ("n_reads", C.c_uint64),
    ]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists() or not ORACLE_BIN.exists():
            subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
        L = C.CDLL(str(ORACLE_SO))
        L.ok_key_words.restype = C.c_int
        L.ok_key_words.argtypes = [C.c_int]
        L.ok_extract.restype = C.c_size_t
        L.ok_extract.argtypes = [C.c_char_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                 C.c_size_t]
        L.ok_bucket.restype = C.c_uint32
        L.ok_bucket.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.ok_parse_sequences.restype = C.c_long
        L.ok_parse_sequences.argtypes = [C.c_char_p, C.c_size_t, C.c_void_p, C.c_size_t]
        L.ok_fastq_split_4lines.restype = C.c_size_t
        L.ok_fastq_split_4lines.argtypes = [C.c_char_p, C.c_size_t]
        L.ok_plain_chunks.restype = C.c_size_t
        L.ok_plain_chunks.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_size_t]
        L.ok_count_fastq.restype = C.c_int
        L.ok_count_fastq.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_uint32,
                                     C.c_size_t, C.c_int, C.POINTER(OkTable)]
        L.ok_table_free.argtypes = [C.POINTER(OkTable)]
        L.ok_count_ref.restype = C.c_int
        L.ok_count_ref.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_uint32, C.c_size_t, C.c_int,
                                   C.POINTER(OkTable), C.POINTER(OkRefStats)]
        L.ok_sort_rle.restype = C.c_size_t
        L.ok_sort_rle.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_void_p]
        L.ok_dvle_encode.restype = C.c_size_t
        L.ok_dvle_encode.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        L.ok_kite_write_table.restype = C.c_int
        L.ok_kite_write_table.argtypes = [C.c_char_p, C.POINTER(OkTable), C.c_uint16, C.c_int,
                                          C.c_uint32, C.c_int, C.c_uint32, C.c_size_t, C.c_int]
        L.ok_kite_load.restype = C.c_int
        L.ok_kite_load.argtypes = [C.c_char_p, C.POINTER(OkTable), C.POINTER(C.c_int),
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_int), C.POINTER(C.c_uint32)]
        L.ok_kite_dump.restype = C.c_int
        L.ok_kite_dump.argtypes = [C.c_char_p, C.c_char_p]
        L.ok_kite_histo.restype = C.c_int
        L.ok_kite_histo.argtypes = [C.POINTER(C.c_char_p), C.c_int, C.POINTER(C.c_char_p),
                                    C.c_char_p]
        L.ok_kite_merge.restype = C.c_int
        L.ok_kite_merge.argtypes = [C.POINTER(C.c_char_p), C.c_int, C.c_char_p, C.c_int]
        L.ok_kite_union.restype = C.c_int
        L.ok_kite_union.argtypes = [C.POINTER(C.c_char_p), C.c_int, C.POINTER(C.c_uint64),
                                    C.POINTER(C.POINTER(C.c_uint64)),
                                    C.POINTER(C.POINTER(C.c_uint64)), C.POINTER(C.c_int)]
        L.ok_decode_key.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p]
        L.ok_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _err() -> str:
    return lib().ok_last_error().decode()


def words_for(k: int) -> int:
    return lib().ok_key_words(k)


def key_to_words(v: int, words: int) -> np.ndarray:
    return np.array([(v >> (64 * i)) & (2**64 - 1) for i in range(words)], dtype=np.uint64)


def words_to_int(w) -> int:
    return sum(int(x) << (64 * i) for i, x in enumerate(w))


def extract(seq: bytes, k: int, canonical: bool, words: int | None = None) -> list[int]:
    words = words or words_for(k)
    n = lib().ok_extract(seq, len(seq), k, int(canonical), words, None, 0)
    out = np.zeros(max(1, n) * words, dtype=np.uint64)
    lib().ok_extract(seq, len(seq), k, int(canonical), words, out.ctypes.data, n)
    return [words_to_int(out[i * words:(i + 1) * words]) for i in range(n)]


def bucket(key: int, words: int, bits: int) -> int:
    w = key_to_words(key, words)
    return lib().ok_bucket(w.ctypes.data, words, bits)


def parse_sequences(data: bytes) -> list[bytes] | None:
    cap = data.count(b"\n") + 2
    spans = np.zeros(2 * cap, dtype=np.uint64)
    n = lib().ok_parse_sequences(data, len(data), spans.ctypes.data, cap)
    if n < 0:
        return None
    return [data[int(spans[2 * i]):int(spans[2 * i] + spans[2 * i + 1])] for i in range(n)]


def plain_chunks(data: bytes | np.ndarray, chunk_size: int) -> list[tuple[int, int]]:
    arr = np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray)) else data
    cap = arr.size // max(1, chunk_size) + 4
    out = np.zeros(2 * cap, dtype=np.uint64)
    n = lib().ok_plain_chunks(arr.ctypes.data, arr.size, chunk_size, out.ctypes.data, cap)
    assert n <= cap
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(n)]


class OkRefStats(C.Structure):
    _fields_ = [("reads", C.c_uint64), ("kmers", C.c_uint64), ("run_rows", C.c_uint64),
                ("encoded_bytes", C.c_uint64)]


class Table:
    """Oracle count table copied into numpy arrays."""

    def __init__(self, t: OkTable):
        self.k, self.canonical, self.words, self.partitions = t.k, t.canonical, t.words, t.partitions
        self.n_unique, self.n_kmers, self.n_reads = t.n_unique, t.n_kmers, t.n_reads
        P = t.partitions
        self.part_off = np.ctypeslib.as_array(t.part_off, shape=(P + 1,)).copy()
        n = max(1, t.n_unique)
        self.keys = np.ctypeslib.as_array(t.keys, shape=(n * t.words,)).copy()[
            :t.n_unique * t.words].reshape(-1, t.words)
        self.counts = np.ctypeslib.as_array(t.counts, shape=(n,)).copy()[:t.n_unique]


def count_fastq(data: bytes | np.ndarray, k: int, canonical: bool, partitions: int,
                chunk_size: int = 8 << 20, threads: int = 4) -> Table:
    arr = np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray)) else data
    arr = np.ascontiguousarray(arr)
    t = OkTable()
    rc = lib().ok_count_fastq(arr.ctypes.data if arr.size else None, arr.size, k, int(canonical),
                              partitions, chunk_size, threads, C.byref(t))
    if rc:
        raise RuntimeError(_err())
    out = Table(t)
    lib().ok_table_free(C.byref(t))
    return out


def count_ref(data: bytes | np.ndarray, k: int, canonical: bool, partitions: int,
              chunk_size: int = 8 << 20, threads: int = 4) -> tuple[Table, dict]:
    """ok_count_ref (the reference-shaped count of the CPU baseline): its
    per-worker runs merged into a table, and its run statistics."""
    arr = np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray)) else data
    arr = np.ascontiguousarray(arr)
    t, st = OkTable(), OkRefStats()
    rc = lib().ok_count_ref(arr.ctypes.data if arr.size else None, arr.size, k, int(canonical), partitions,
                            chunk_size, threads, C.byref(t), C.byref(st))
    if rc:
        raise RuntimeError(_err())
    out = Table(t)
    lib().ok_table_free(C.byref(t))
    return out, {"reads": st.reads, "kmers": st.kmers, "run_rows": st.run_rows,
                 "encoded_bytes": st.encoded_bytes}


def sort_rle(keys: np.ndarray, words: int, k: int) -> tuple[np.ndarray, np.ndarray]:
    keys = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    n = keys.size // words
    counts = np.zeros(max(1, n), dtype=np.uint32)
    u = lib().ok_sort_rle(keys.ctypes.data, n, words, k, counts.ctypes.data)
    return keys[:u * words].reshape(-1, words), counts[:u]


def dvle_encode(keys: np.ndarray, counts: np.ndarray, words: int) -> bytes:
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    n = counts.size
    out = np.zeros(max(64, n * 48), dtype=np.uint8)
    ln = lib().ok_dvle_encode(keys.ctypes.data, counts.ctypes.data, n, words, out.ctypes.data)
    return out[:ln].tobytes()


def _paths(paths) -> C.Array:
    arr = (C.c_char_p * len(paths))(*[str(p).encode() for p in paths])
    return arr


def kite_dump(path, out) -> None:
    if lib().ok_kite_dump(str(path).encode(), str(out).encode()):
        raise RuntimeError(_err())


def kite_histo(paths, out, labels=None) -> None:
    lab = _paths(labels) if labels else None
    if lib().ok_kite_histo(_paths(paths), len(paths), lab, str(out).encode()):
        raise RuntimeError(_err())


def kite_merge(paths, out, lanes: int = 2) -> None:
    if lib().ok_kite_merge(_paths(paths), len(paths), str(out).encode(), lanes):
        raise RuntimeError(_err())


def kite_union(paths) -> tuple[np.ndarray, np.ndarray]:
    rows = C.c_uint64()
    kp = C.POINTER(C.c_uint64)()
    cp = C.POINTER(C.c_uint64)()
    w = C.c_int()
    if lib().ok_kite_union(_paths(paths), len(paths), C.byref(rows), C.byref(kp), C.byref(cp),
                           C.byref(w)):
        raise RuntimeError(_err())
    n = rows.value
    keys = np.ctypeslib.as_array(kp, shape=(max(1, n) * w.value,)).copy()[:n * w.value]
    counts = np.ctypeslib.as_array(cp, shape=(max(1, n) * len(paths),)).copy()[:n * len(paths)]
    libc = C.CDLL(None)
    libc.free(kp)
    libc.free(cp)
    return keys.reshape(-1, w.value), counts.reshape(-1, len(paths))


def kite_load(path) -> tuple[Table, dict]:
    t = OkTable()
    hmin, hmax = C.c_int(), C.c_int()
    mn, mx = C.c_uint32(), C.c_uint32()
    if lib().ok_kite_load(str(path).encode(), C.byref(t), C.byref(hmin), C.byref(mn),
                          C.byref(hmax), C.byref(mx)):
        raise RuntimeError(_err())
    out = Table(t)
    lib().ok_table_free(C.byref(t))
    return out, {"min": mn.value if hmin.value else None, "max": mx.value if hmax.value else None}


def write_table(path, table: Table, codec_mask: int = 7, min_count=None, max_count=None,
                staging: int = 16 << 20, lanes: int = 2) -> None:
    t = OkTable()
    t.k, t.canonical, t.words, t.partitions = table.k, table.canonical, table.words, table.partitions
    po = np.ascontiguousarray(table.part_off, dtype=np.uint64)
    ks = np.ascontiguousarray(table.keys, dtype=np.uint64).reshape(-1)
    cs = np.ascontiguousarray(table.counts, dtype=np.uint64)
    if ks.size == 0:
        ks = np.zeros(4, dtype=np.uint64)
    if cs.size == 0:
        cs = np.zeros(1, dtype=np.uint64)
    t.part_off = po.ctypes.data_as(C.POINTER(C.c_uint64))
    t.keys = ks.ctypes.data_as(C.POINTER(C.c_uint64))
    t.counts = cs.ctypes.data_as(C.POINTER(C.c_uint64))
    t.n_unique = len(table.counts)
    rc = lib().ok_kite_write_table(str(path).encode(), C.byref(t), codec_mask,
                                   int(min_count is not None), min_count or 0,
                                   int(max_count is not None), max_count or 0, staging, lanes)
    if rc:
        raise RuntimeError(_err())


def decode_key(key: int, words: int, k: int) -> str:
    w = key_to_words(key, words)
    buf = C.create_string_buffer(k + 1)
    lib().ok_decode_key(w.ctypes.data, words, k, buf)
    return buf.raw[:k].decode()


def dump_tsv_from_table(table: Table, min_count=None, max_count=None) -> bytes:
    """Dump text of an oracle table (u32 saturating totals, header filter)."""
    out = []
    for i in range(len(table.counts)):
        c = min(int(table.counts[i]), 0xFFFFFFFF)
        if min_count is not None and c < min_count:
            continue
        if max_count is not None and c > max_count:
            continue
        out.append(decode_key(words_to_int(table.keys[i]), table.words, table.k) + f"\t{c}\n")
    return "".join(out).encode()
