"""Brute-force restatements used to pin the oracle on tiny inputs.

These follow the mathematics of the path, not the oracle's code: a k-mer is
the base-4 integer of its letters (A0 C1 G2 T3, first letter most
significant), its reverse complement is computed on the string, and counts
come from collections.Counter.
"""
from __future__ import annotations

from collections import Counter

CODE = {"A": 0, "C": 1, "G": 2, "T": 3}
COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}
GOLD = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


def kmer_value(s: str) -> int:
    v = 0
    for ch in s:
        v = v * 4 + CODE[ch]
    return v


def revcomp(s: str) -> str:
    return "".join(COMP[c] for c in reversed(s))


def kmers(seq: str, k: int, canonical: bool) -> list[int]:
    """All windows of k upper-cased ACGT letters; any other byte breaks windows."""
    out = []
    s = seq
    for i in range(len(s) - k + 1):
        w = s[i:i + k]
        if all(c in "ACGTacgt" for c in w):
            w = w.upper()
            v = kmer_value(w)
            if canonical:
                v = min(v, kmer_value(revcomp(w)))
            out.append(v)
    return out


def words_for(k: int) -> int:
    return 1 if k <= 31 else (2 if k <= 64 else 4)


def partition(v: int, k: int, P: int) -> int:
    W = words_for(k)
    f = 0
    for i in range(W):
        f ^= (v >> (64 * i)) & M64
    bits = P.bit_length() - 1
    if bits == 0:
        return 0
    return ((f * GOLD) & M64) >> (64 - bits)


def fastq_sequences(text: bytes) -> list[str]:
    """Line 2 of every 4-line record; the file must start with '@'."""
    if not text:
        return []
    if text[:1] != b"@":
        return []
    lines = text.split(b"\n")
    return [lines[i].decode("latin-1") for i in range(1, len(lines), 4)]


def count(text: bytes, k: int, canonical: bool, P: int) -> dict[int, list[tuple[int, int]]]:
    c = Counter()
    for s in fastq_sequences(text):
        c.update(kmers(s, k, canonical))
    parts: dict[int, list[tuple[int, int]]] = {}
    for v, n in c.items():
        parts.setdefault(partition(v, k, P), []).append((v, n))
    return {p: sorted(lst) for p, lst in parts.items()}


def decode(v: int, k: int) -> str:
    """Letters of a key the way the reference's dump writes them.  For k >= 65
    the reference decodes 64 letters from all 128 bits of the high half and
    then the last k-64 letters of the low half."""
    L = "ACGT"
    if k <= 64:
        return "".join(L[(v >> (2 * (k - 1 - i))) & 3] for i in range(k))
    lo = v & ((1 << 128) - 1)
    hi = v >> 128
    first = "".join(L[(hi >> (126 - 2 * i)) & 3] for i in range(64))
    lb = k - 64
    rest = "".join(L[(lo >> (2 * (lb - 1 - j))) & 3] for j in range(lb))
    return first + rest


def dump(text: bytes, k: int, canonical: bool, P: int, min_count=None, max_count=None) -> bytes:
    parts = count(text, k, canonical, P)
    out = []
    for p in range(P):
        for v, n in parts.get(p, []):
            n = min(n, 0xFFFFFFFF)
            if min_count is not None and n < min_count:
                continue
            if max_count is not None and n > max_count:
                continue
            out.append(f"{decode(v, k)}\t{n}\n")
    return "".join(out).encode()


def exact_windows(fq, k: int, reads_per_block: int = 100_000) -> int:
    """Valid k-mer windows of a FASTQ image (numpy uint8 array of whole
    newline-terminated 4-line records): sequence lines (line 2 of every
    record) split at every byte that is not A/C/G/T in either case; a run
    of L valid bases holds max(0, L - k + 1) windows.  Vectorised over
    reads of equal length, for workload-sized inputs."""
    import numpy as np
    nl = np.flatnonzero(fq == 0x0A)
    assert nl.size % 4 == 0
    starts = nl[0::4] + 1
    lens = nl[1::4] - starts
    valid = np.zeros(256, dtype=np.int32)
    for b in b"ACGTacgt":
        valid[b] = 1
    total = 0
    for L in np.unique(lens):
        L = int(L)
        if L < k:
            continue
        sel = np.flatnonzero(lens == L)
        for a in range(0, sel.size, reads_per_block):
            s = starts[sel[a:a + reads_per_block]]
            bad = 1 - valid[fq[s[:, None] + np.arange(L)[None, :]]]
            cs = np.zeros((s.size, L + 1), dtype=np.int32)
            np.cumsum(bad, axis=1, out=cs[:, 1:])
            total += int(np.count_nonzero(cs[:, k:] - cs[:, :L - k + 1] == 0))
    return total
