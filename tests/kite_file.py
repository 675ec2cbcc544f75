"""A minimal, independent Python reading of the .kite byte layout
(src/format/header.rs:62-151, bucket_dir.rs:5-140, run_block.rs:1-95).

Used by tests to build files with arbitrary chunk structure and to inspect
the headers, chunk indexes and partition tables that the engine and the
oracle write.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field


def varint(v: int) -> bytes:
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        if v == 0:
            out.append(b)
            return bytes(out)
        out.append(b | 0x80)


def read_varint(buf: bytes, pos: int) -> tuple[int, int]:
    r = 0
    sh = 0
    while True:
        b = buf[pos]
        pos += 1
        r |= (b & 0x7F) << sh
        if not b & 0x80:
            return r, pos
        sh += 7


def encode_run(rows: list[tuple[int, int]], key_width: int) -> tuple[int, bytes]:
    """(codec, payload) of one sorted run."""
    if key_width == 256:
        data = b"".join((k & ((1 << 128) - 1)).to_bytes(16, "little") + (k >> 128).to_bytes(16, "little")
                        + c.to_bytes(4, "little") for k, c in rows)
        return 1, data
    kb = key_width // 8
    k0, c0 = rows[0]
    out = bytearray(k0.to_bytes(kb, "little") + c0.to_bytes(4, "little"))
    prev = k0
    for k, c in rows[1:]:
        out += varint(k - prev) + varint(c)
        prev = k
    return 2, bytes(out)


def decode_run(codec: int, data: bytes, rows: int, key_width: int) -> list[tuple[int, int]]:
    out = []
    if key_width == 256 or not codec & 2:
        es = {64: 12, 128: 20, 256: 36}[key_width]
        kb = es - 4
        for i in range(rows):
            e = data[i * es:(i + 1) * es]
            if key_width == 256:
                k = int.from_bytes(e[:16], "little") | (int.from_bytes(e[16:32], "little") << 128)
            else:
                k = int.from_bytes(e[:kb], "little")
            out.append((k, int.from_bytes(e[kb:kb + 4], "little")))
        return out
    kb = key_width // 8
    k = int.from_bytes(data[:kb], "little")
    out.append((k, int.from_bytes(data[kb:kb + 4], "little")))
    pos = kb + 4
    for _ in range(rows - 1):
        d, pos = read_varint(data, pos)
        c, pos = read_varint(data, pos)
        k += d
        out.append((k, c))
    return out


@dataclass
class Chunk:
    partition: int
    codec: int
    file_offset: int
    encoded_bytes: int
    raw_bytes: int
    rows: int
    key_min: tuple[int, int]
    key_max: tuple[int, int]


@dataclass
class KiteFile:
    k: int
    canonical: int
    partitions: int
    key_width: int
    codec_mask: int
    total_chunks: int
    total_rows: int
    min_count: int | None
    max_count: int | None
    chunks: list[Chunk] = field(default_factory=list)
    part_table: list[tuple[int, int, int, int, int]] = field(default_factory=list)
    raw: bytes = b""

    def runs(self, partition: int) -> list[list[tuple[int, int]]]:
        first, count = self.part_table[partition][:2]
        out = []
        for c in self.chunks[first:first + count]:
            data = self.raw[c.file_offset:c.file_offset + c.encoded_bytes]
            out.append(decode_run(c.codec, data, c.rows, self.key_width))
        return out


def read(path) -> KiteFile:
    raw = open(path, "rb").read()
    assert raw[:4] == b"KX2!", "magic"
    ver, k, canon, P, kw, cm, cs = struct.unpack_from("<HBBIHHI", raw, 4)
    assert ver == 1
    cio, pto, fo, tc, tr = struct.unpack_from("<QQQQQ", raw, 20)
    mn = mx = None
    if raw[60:64] == b"KMT1":
        fl, a, b = struct.unpack_from("<III", raw, 64)
        mn = a if fl & 1 else None
        mx = b if fl & 2 else None
    assert raw[fo + 16:fo + 24] == b"KX2FOOT!"
    f = KiteFile(k, canon, P, kw, cm, tc, tr, mn, mx, raw=raw)
    n = (pto - cio) // 128
    for i in range(n):
        e = raw[cio + 128 * i:cio + 128 * (i + 1)]
        pid, codec, off, enc, rawb, rows = struct.unpack_from("<IIQQQQ", e, 0)
        kminlo, kminhi, kmaxlo, kmaxhi = (int.from_bytes(e[40 + 16 * j:56 + 16 * j], "little") for j in range(4))
        f.chunks.append(Chunk(pid, codec, off, enc, rawb, rows, (kminlo, kminhi), (kmaxlo, kmaxhi)))
    for p in range(P):
        f.part_table.append(struct.unpack_from("<QQQQQ", raw, pto + 40 * p))
    return f


def write(path, k: int, canonical: bool, partitions: int, key_width: int,
          runs: list[tuple[int, list[tuple[int, int]]]], codec_mask: int = 7,
          min_count: int | None = None, max_count: int | None = None) -> None:
    """Write runs [(partition, rows)] sequentially in the given order."""
    data = bytearray()
    metas = []
    off = 128
    for p, rows in runs:
        codec, payload = encode_run(rows, key_width)
        k0, k1 = rows[0][0], rows[-1][0]
        split = (lambda v: (v & ((1 << 128) - 1), v >> 128)) if key_width == 256 else (lambda v: (v, 0))
        metas.append((p, codec, off, len(payload), len(payload), len(rows), split(k0), split(k1)))
        data += payload
        off += len(payload)
    by_part = {p: [] for p in range(partitions)}
    for m in metas:
        by_part[m[0]].append(m)
    index = bytearray()
    table = bytearray()
    first = 0
    for p in range(partitions):
        ms = sorted(by_part[p], key=lambda m: m[2])
        for m in ms:
            index += struct.pack("<IIQQQQ", m[0], m[1], m[2], m[3], m[4], m[5])
            for v in (m[6][0], m[6][1], m[7][0], m[7][1]):
                index += v.to_bytes(16, "little")
            index += bytes(24)
        table += struct.pack("<QQQQQ", first, len(ms), sum(m[5] for m in ms),
                             min((m[2] for m in ms), default=0), sum(m[3] for m in ms))
        first += len(ms)
    cio = 128 + len(data)
    pto = cio + len(index)
    fo = pto + len(table)
    tc = len(metas)
    tr = sum(m[5] for m in metas)
    hdr = bytearray(b"KX2!" + struct.pack("<HBBIHHI", 1, k, int(canonical), partitions, key_width,
                                          codec_mask, 0) + struct.pack("<QQQQQ", cio, pto, fo, tc, tr))
    res = bytearray(68)
    if min_count is not None or max_count is not None:
        res[0:4] = b"KMT1"
        fl = (1 if min_count is not None else 0) | (2 if max_count is not None else 0)
        res[4:16] = struct.pack("<III", fl, min_count or 0, max_count or 0)
    hdr += res
    assert len(hdr) == 128
    with open(path, "wb") as fh:
        fh.write(hdr + data + index + table + struct.pack("<QQ", tc, tr) + b"KX2FOOT!")
