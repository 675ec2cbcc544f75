"""A BGZF writer for tests (SAM/BAM specification, section 4.1): gzip
members of at most 64 KiB of input, each with a BC extra field holding the
member size minus one, and the empty end-of-file member.  Raw deflate comes
from Python's zlib, so the inflater under test sees streams written by an
independent encoder (stored, fixed-Huffman and dynamic-Huffman blocks,
depending on level and strategy)."""
import struct
import zlib

EOF_MEMBER = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")


def member(piece: bytes, level: int = 6, strategy: int = zlib.Z_DEFAULT_STRATEGY) -> bytes:
    c = zlib.compressobj(level, zlib.DEFLATED, -15, 9, strategy)
    data = c.compress(piece) + c.flush()
    bsize = 12 + 6 + len(data) + 8
    head = struct.pack("<BBBBIBBH", 0x1F, 0x8B, 8, 4, 0, 0, 0xFF, 6) + b"BC" + struct.pack("<HH", 2, bsize - 1)
    tail = struct.pack("<II", zlib.crc32(piece) & 0xFFFFFFFF, len(piece))
    return head + data + tail


def compress(data: bytes, piece: int = 0xFF00, level: int = 6, strategy: int = zlib.Z_DEFAULT_STRATEGY,
             eof: bool = True) -> bytes:
    out = [member(data[i:i + piece], level, strategy) for i in range(0, len(data), piece)]
    if eof:
        out.append(EOF_MEMBER)
    return b"".join(out)
