"""BGZF input inflated on the device (kite_gpu_inflate_bgzf,
kite_gpu_count_feed_bgzf): the inflated bytes equal the input for streams of
every DEFLATE block type written by zlib; malformed members are refused; a
BGZF count equals the count of the same plain FASTQ, through the ABI with
small batches (record cuts and carried tails) and through the CLI."""
import gzip
import os
import random
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

import bgzf
import oracle_lib as ol
from kite_amd import GpuContext, KiteError

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
KITE = ROOT / "kite_amd" / "lib" / "kite"


def _fastq(n, L, seed, alphabet="ACGTACGTACGTN"):
    rng = random.Random(seed)
    recs = []
    for i in range(n):
        ln = rng.choice([L, L // 3, 2 * L, 0]) if L else 0
        s = "".join(rng.choice(alphabet) for _ in range(ln))
        q = "".join(rng.choice("#,:FIJ@") for _ in range(ln))
        recs.append(f"@r{i} x\n{s}\n+\n{q}\n")
    return "".join(recs).encode()


@pytest.mark.parametrize("kind", ["random", "fastq", "runs", "text"])
@pytest.mark.parametrize("level,strategy", [(6, zlib.Z_DEFAULT_STRATEGY), (9, zlib.Z_DEFAULT_STRATEGY),
                                            (1, zlib.Z_FIXED), (0, zlib.Z_DEFAULT_STRATEGY),
                                            (6, zlib.Z_HUFFMAN_ONLY), (6, zlib.Z_RLE)])
def test_inflate_matches_input(gpu_ctx, kind, level, strategy):
    rng = random.Random(hash((kind, level, strategy)) & 0xFFFF)
    if kind == "random":
        data = bytes(rng.randrange(256) for _ in range(150000))
    elif kind == "fastq":
        data = _fastq(800, 150, 3)
    elif kind == "runs":
        data = b"".join(bytes([rng.randrange(4) + 65]) * rng.randrange(1, 600) for _ in range(800))
    else:
        data = b" ".join(rng.choice([b"kmer", b"count", b"partition", b"ACGT" * 20, b"\n"]) for _ in range(40000))
    for piece in (0xFF00, 65536, 1000, 1):
        if piece == 65536 and (kind == "random" or level == 0):
            continue  # 64 KiB that does not compress does not fit one member (BSIZE is 16 bits)
        if piece == 1:
            data_p = data[:3000]
        else:
            data_p = data
        comp = bgzf.compress(data_p, piece=piece, level=level, strategy=strategy)
        assert gzip.decompress(comp) == data_p  # the writer is valid gzip
        out, ms = gpu_ctx.inflate_bgzf(comp)
        same = out == data_p
        assert same, f"{kind} level={level} strategy={strategy} piece={piece}"
        assert ms >= 0


def test_inflate_refuses_malformed(gpu_ctx):
    data = _fastq(300, 150, 9)
    comp = bgzf.compress(data, piece=20000, eof=False)
    # a truncated last member is not consumed
    members = -(-len(data) // 20000)
    assert members >= 2
    out, _ = gpu_ctx.inflate_bgzf(comp[:-10])
    assert out == data[:(members - 1) * 20000]
    # not BGZF
    with pytest.raises(KiteError, match="Non-BGZF"):
        gpu_ctx.inflate_bgzf(gzip.compress(data))
    # a member whose trailer claims another size
    m = bytearray(bgzf.member(data[:5000]))
    m[-4:] = (4999).to_bytes(4, "little")
    with pytest.raises(KiteError, match="ISIZE"):
        gpu_ctx.inflate_bgzf(bytes(m))
    # deflate data with an invalid block type (3)
    bad = bytearray(bgzf.member(data[:5000], level=0))
    bad[18] = 0x07  # BFINAL=1, BTYPE=11
    with pytest.raises(KiteError, match="deflate"):
        gpu_ctx.inflate_bgzf(bytes(bad))
    # a dynamic stream cut short: the member keeps its sizes, the data ends early
    m = bgzf.member(data[:20000], level=9)
    cut = bytearray(m[:18 + 200] + m[-8:])
    cut[16:18] = (len(cut) - 1).to_bytes(2, "little")
    with pytest.raises(KiteError):
        gpu_ctx.inflate_bgzf(bytes(cut))


class _Bits:
    """LSB-first DEFLATE bit writer (RFC 1951 3.1.1): fields LSB first,
    Huffman codes most significant bit first."""

    def __init__(self):
        self.v, self.n = 0, 0

    def put(self, value, nbits):
        self.v |= value << self.n
        self.n += nbits

    def huff(self, code, nbits):
        self.put(int(format(code, f"0{nbits}b")[::-1], 2), nbits)

    def bytes(self):
        return self.v.to_bytes((self.n + 7) // 8, "little")


def _fixed_stream(symbols):
    """One final fixed-Huffman block (RFC 1951 3.2.6): ("lit", byte),
    ("len", symbol 257..287), ("dist", code 0..31), ("eob",); no extra bits
    (the lengths and distances used here have none)."""
    w = _Bits()
    w.put(1, 1)
    w.put(1, 2)
    for s in symbols:
        if s[0] == "dist":
            w.huff(s[1], 5)
            continue
        v = {"lit": lambda: s[1], "len": lambda: s[1], "eob": lambda: 256}[s[0]]()
        if v < 144:
            w.huff(0x30 + v, 8)
        elif v < 256:
            w.huff(0x190 + v - 144, 9)
        elif v < 280:
            w.huff(v - 256, 7)
        else:
            w.huff(0xC0 + v - 280, 8)
    return w.bytes()


def _raw_member(deflate, isize):
    import struct
    bsize = 12 + 6 + len(deflate) + 8
    head = struct.pack("<BBBBIBBH", 0x1F, 0x8B, 8, 4, 0, 0, 0xFF, 6) + b"BC" + struct.pack("<HH", 2, bsize - 1)
    return head + deflate + struct.pack("<II", 0, isize)


def test_inflate_refuses_invalid_codes(gpu_ctx):
    """Symbols a fixed-Huffman block can encode but RFC 1951 forbids:
    literal/length 286-287, distance codes 30-31, and a distance beyond the
    bytes produced.  Each crafted stream is pinned by zlib first (the valid
    one inflates to ABABA, the others raise)."""
    ok = _fixed_stream([("lit", 65), ("lit", 66), ("len", 257), ("dist", 1), ("eob",)])
    assert zlib.decompress(ok, -15) == b"ABABA"
    out, _ = gpu_ctx.inflate_bgzf(_raw_member(ok, 5))
    assert out == b"ABABA"
    bad = {
        "literal/length 286": _fixed_stream([("lit", 65), ("len", 286), ("eob",)]),
        "literal/length 287": _fixed_stream([("lit", 65), ("len", 287), ("eob",)]),
        "distance code 30": _fixed_stream([("lit", 65), ("lit", 66), ("len", 257), ("dist", 30), ("eob",)]),
        "distance code 31": _fixed_stream([("lit", 65), ("lit", 66), ("len", 257), ("dist", 31), ("eob",)]),
        "distance too far": _fixed_stream([("lit", 65), ("len", 257), ("dist", 1), ("eob",)]),
    }
    for name, raw in bad.items():
        with pytest.raises(zlib.error):
            zlib.decompress(raw, -15)
        with pytest.raises(KiteError, match="deflate"):
            gpu_ctx.inflate_bgzf(_raw_member(raw, 5))


@pytest.mark.parametrize("k,canon,P", [(21, True, 64), (31, False, 16), (45, True, 8), (77, False, 4)])
def test_count_bgzf_equals_plain(k, canon, P):
    """A tiny device budget makes many batches, each cut after its last
    record with the rest carried; the emitted runs, summed per partition,
    equal the oracle's table of the plain FASTQ."""
    import kite_file as kf
    fq = _fastq(6000, 140, k, alphabet="ACGTACGTACGTACGTN")
    comp = bgzf.compress(fq, piece=7000)
    tab = ol.count_fastq(fq, k, canon, P, chunk_size=1 << 40, threads=4)
    with GpuContext(0, hbm_budget_bytes=6 << 20) as ctx:
        ctx.count_begin(k, canon, P)
        buf, pos, step = b"", 0, 50000
        while pos < len(comp):
            buf += comp[pos:pos + step]
            pos += step
            buf = buf[ctx.count_feed_bgzf(buf):]
        assert buf == b""
        st = ctx.count_finish()
        assert st.batches > 1
        emitted = ctx.count_emit()
    assert st.kmers == tab.n_kmers
    assert st.input_bytes == len(fq)
    W = ol.words_for(k)
    kw = {1: 64, 2: 128, 4: 256}[W]
    sums = [dict() for _ in range(P)]
    for part, codec, rows, raw, kmin, kmax, data in emitted:
        for key, c in kf.decode_run(codec, data, rows, kw):
            sums[part][key] = sums[part].get(key, 0) + c
    for p in range(P):
        a, b = int(tab.part_off[p]), int(tab.part_off[p + 1])
        exp = {ol.words_to_int(tab.keys[i]): int(tab.counts[i]) for i in range(a, b)}
        assert sums[p] == exp, f"partition {p}"


def test_cli_count_bgzf(tmp_path):
    fq = _fastq(5000, 150, 77, alphabet="ACGT")
    (tmp_path / "r.fq").write_bytes(fq)
    (tmp_path / "r.fq.gz").write_bytes(bgzf.compress(fq))
    env_dev = dict(os.environ)
    env_host = dict(os.environ, KITE_BGZF_ENGINE="host")
    outs = []
    for name, inp, env in (("plain", "r.fq", env_dev), ("bgzf", "r.fq.gz", env_dev), ("zlib", "r.fq.gz", env_host)):
        r = subprocess.run([str(KITE), "count", "-i", str(tmp_path / inp), "-o", str(tmp_path / f"{name}.kite"),
                            "-k", "25", "--canonical", "true", "--partitions", "32"], capture_output=True, text=True,
                           timeout=300, env=env)
        assert r.returncode == 0, r.stderr
        if name != "plain":
            assert "compression: bgzf" in r.stderr
        r = subprocess.run([str(KITE), "dump", "-i", str(tmp_path / f"{name}.kite"), "-o", "-"], capture_output=True,
                           timeout=300)
        assert r.returncode == 0
        outs.append(r.stdout)
    assert outs[0] == outs[1] == outs[2]
    ol.kite_dump(tmp_path / "bgzf.kite", tmp_path / "o.tsv")
    assert outs[1] == (tmp_path / "o.tsv").read_bytes()
