"""The oracle against every known answer the reference's own unit tests hold
for this path (jiehua1995/Kite v0.7.5 #[test] functions, cited per test)."""
import numpy as np
import pytest

import brute
import kite_file as kf
import oracle_lib as ol


def rc(s: bytes) -> bytes:
    return bytes({65: 84, 84: 65, 67: 71, 71: 67}.get(b, b) for b in reversed(s))


# ---- src/count/kmer.rs:437-621 ---------------------------------------------
def test_kmer_simple():  # kmer.rs:438-448
    assert ol.extract(b"AACGT", 3, False, 1) == [1, 6, 27]


def test_kmer_n_reset():  # kmer.rs:450-458
    assert ol.extract(b"ACGNACT", 3, False, 1) == [6, 7]


def test_canonical_k1():  # kmer.rs:460-473
    assert ol.extract(b"ACT", 1, False, 1) == [0, 1, 3]
    assert ol.extract(b"ACT", 1, True, 1) == [0, 1, 0]


def test_canonical_k21():  # kmer.rs:475-501
    assert ol.extract(b"ACGTACGTACGTACGTACGT", 21, True, 1) == []
    s = b"ACGTACGTACGTACGTACGTA"
    k1 = ol.extract(s, 21, True, 1)
    assert len(k1) == 1
    assert k1 == ol.extract(rc(s), 21, True, 1)


def test_kmer128_simple():  # kmer.rs:503-512
    assert ol.extract(b"AACGT", 3, False, 2) == ol.extract(b"AACGT", 3, False, 1)


def test_kmer128_k51():  # kmer.rs:514-523
    s = b"ACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACG"
    assert len(s) == 51
    v = ol.extract(s, 51, False, 2)
    assert len(v) == 1 and v[0] < (1 << 102)


def test_kmer128_canonical():  # kmer.rs:525-548
    s = b"ACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACG"
    assert ol.extract(s, 51, True, 2) == ol.extract(rc(s), 51, True, 2)


def test_kmer128_n_reset():  # kmer.rs:550-558
    s = (b"ACGACGACGACGACGACGACGACGACGACGACGACGACGACGACGNACGACGACGACGACGACGACGACGACGACG"
         b"ACGACGACGACGACGACGACGACG")
    assert len(ol.extract(s, 50, False, 2)) == 5


def test_kmer128_agrees_with_u64_at_k32():  # kmer.rs:560-569
    s = b"ACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGT"
    assert ol.extract(s, 32, False, 2) == ol.extract(s, 32, False, 1)


def test_kmeru256_k65_basic():  # kmer.rs:573-583
    s = b"ACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTAG"
    assert len(s) == 66
    v = ol.extract(s, 65, False, 4)
    assert len(v) == 2
    assert (v[0] >> 128) < 4  # hi fits in 2 bits


def test_kmeru256_canonical():  # kmer.rs:585-607
    s = b"ACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTA"
    assert ol.extract(s, 65, True, 4) == ol.extract(rc(s), 65, True, 4)


def test_kmeru256_n_reset():  # kmer.rs:609-621
    s = bytearray(b"A" * 70)
    s[40] = ord("N")
    assert len(ol.extract(bytes(s), 65, False, 4)) == 0
    s2 = bytearray(b"A" * 150)
    s2[60] = ord("N")
    assert len(ol.extract(bytes(s2), 65, False, 4)) == 25


def test_key_width_dispatch():  # count/mod.rs:232-238
    assert [ol.words_for(k) for k in (1, 21, 31, 32, 63, 64, 65, 128)] == [1, 1, 1, 2, 2, 2, 4, 4]


# ---- src/count/sorter.rs:532-603 -------------------------------------------
def test_radix_sort_small():  # sorter.rs:536-543
    v = np.array([5, 3, 1, 4, 1, 5, 9, 2, 6, 5, 3, 5], dtype=np.uint64)
    keys, counts = ol.sort_rle(v, 1, 32)
    assert list(np.repeat(keys[:, 0], counts)) == sorted(v.tolist())


def test_radix_sort_large():  # sorter.rs:545-551
    v = np.arange(2000, dtype=np.uint64)[::-1]
    keys, counts = ol.sort_rle(v, 1, 32)
    assert keys[:, 0].tolist() == list(range(2000)) and set(counts.tolist()) == {1}


def test_radix_sort_kmer_k21():  # sorter.rs:572-590
    rng = np.random.default_rng(5)
    v = rng.integers(0, 1 << 42, size=5000, dtype=np.uint64)
    keys, counts = ol.sort_rle(v, 1, 21)
    assert np.repeat(keys[:, 0], counts).tolist() == sorted(v.tolist())


def test_rle_compress():  # sorter.rs:592-598
    keys, counts = ol.sort_rle(np.array([1, 1, 2, 3, 3, 3, 5], dtype=np.uint64), 1, 32)
    assert list(zip(keys[:, 0].tolist(), counts.tolist())) == [(1, 2), (2, 1), (3, 3), (5, 1)]


def test_rle_empty():  # sorter.rs:600-603
    keys, counts = ol.sort_rle(np.zeros(0, dtype=np.uint64), 1, 32)
    assert len(counts) == 0


# ---- src/count/parser.rs:79-118 --------------------------------------------
def test_parse_sequences():  # parser.rs:83-93
    data = b"@r1\nACGT\n+\nIIII\n@r2\nTTGCA\n+\nIIIII\n"
    assert ol.parse_sequences(data) == [b"ACGT", b"TTGCA"]


def test_parse_keeps_n():  # parser.rs:95-101
    assert ol.parse_sequences(b"@r\nACNGT\n+\nIIIII\n") == [b"ACNGT"]


def test_parse_empty():  # parser.rs:103-106
    assert ol.parse_sequences(b"") == []


def test_parse_bad_start():  # parser.rs:108-111
    assert ol.parse_sequences(b"ACGT\n+\nIIII\n") is None


def test_parse_records_and_bases():  # parser.rs:113-118
    seqs = ol.parse_sequences(b"@a\nACG\n+\nIII\n@b\nTTTT\n+\nIIII\n")
    assert (len(seqs), sum(map(len, seqs))) == (2, 7)


def test_parse_unterminated_last_sequence():  # parser.rs:59-62
    assert ol.parse_sequences(b"@a\nACG\n+\nIII\n@b\nTTGG") == [b"ACG", b"TTGG"]


# ---- src/count/decompress.rs:154-186, src/count/bgzf.rs:268-331 -------------
def test_split_4lines_quality_at():  # bgzf.rs: '@' at the start of a quality line
    data = b"@r1\nACGT\n+\n@III\n@r2\nAC"
    assert ol.lib().ok_fastq_split_4lines(data, len(data)) == len(b"@r1\nACGT\n+\n@III\n")


def test_split_4lines_partial_and_empty():
    assert ol.lib().ok_fastq_split_4lines(b"@r1\nACGT\n+", 10) == 0
    assert ol.lib().ok_fastq_split_4lines(b"", 0) == 0


def test_plain_chunks_record_aligned():  # decompress.rs:104-138
    rec = b"@r\nACGTACGT\n+\nIIIIIIII\n"  # 23 bytes
    data = rec * 10
    chunks = ol.plain_chunks(data, 50)
    assert sum(l for _, l in chunks) == len(data)
    for off, ln in chunks[:-1]:
        assert ln % len(rec) == 0 and off % len(rec) == 0


# ---- src/dump/decoder.rs:119-159 -------------------------------------------
def test_decode_acg():  # decoder.rs:122-127
    assert ol.decode_key(6, 1, 3) == "ACG"


def test_decode_k1():  # decoder.rs:129-137
    assert [ol.decode_key(i, 1, 1) for i in range(4)] == ["A", "C", "G", "T"]


def test_decode_k21_roundtrip():  # decoder.rs:139-152
    s = "ACGTTGCAACGTACGTTTGCA"
    assert ol.decode_key(brute.kmer_value(s), 1, 21) == s


def test_decode_u256_reference_layout():  # decoder.rs:57-107
    # the reference's u256 decode: 64 letters from all of hi, then the last k-64 of lo
    s = "C" * 3 + "G" * 64  # k = 67
    v = brute.kmer_value(s)
    assert ol.decode_key(v, 4, 67) == brute.decode(v, 67)
    assert ol.decode_key(v, 4, 67) == "A" * 61 + "CCC" + "G" * 3


# ---- src/kiteops/mod.rs:861-976, src/dump/loser_tree.rs:178-238 -----------
def _dump(path, tmp_path) -> list[tuple[str, int]]:
    out = tmp_path / "d.tsv"
    ol.kite_dump(path, out)
    return [(l.split("\t")[0], int(l.split("\t")[1])) for l in out.read_text().splitlines()]


def test_count_filter_bounds(tmp_path):  # kiteops/mod.rs:870-880
    f = tmp_path / "f.kite"
    rows = [(i, c) for i, c in enumerate([1, 2, 5, 10, 11], start=1)]
    kf.write(f, 3, False, 1, 64, [(0, rows)], min_count=2, max_count=10)
    assert [c for _, c in _dump(f, tmp_path)] == [2, 5, 10]
    kf.write(f, 3, False, 1, 64, [(0, rows)], min_count=2)
    assert [c for _, c in _dump(f, tmp_path)] == [2, 5, 10, 11]
    kf.write(f, 3, False, 1, 64, [(0, rows)], max_count=10)
    assert [c for _, c in _dump(f, tmp_path)] == [1, 2, 5, 10]


def test_merge_vectors_total(tmp_path):  # kiteops/mod.rs:883-894
    ins = [[(1, 2), (3, 4)], [(2, 5), (3, 6)], [(1, 7), (4, 8)]]
    paths = []
    for i, rows in enumerate(ins):
        p = tmp_path / f"in{i}.kite"
        kf.write(p, 3, False, 1, 64, [(0, rows)])
        paths.append(p)
    ol.kite_merge(paths, tmp_path / "m.kite")
    m = kf.read(tmp_path / "m.kite")
    assert m.runs(0) == [[(1, 9), (2, 5), (3, 10), (4, 8)]]
    assert m.codec_mask == 3


def test_union_rows_zero_filled(tmp_path):  # kiteops/mod.rs:933-945
    a, b = tmp_path / "a.kite", tmp_path / "b.kite"
    kf.write(a, 3, False, 1, 64, [(0, [(1, 2), (3, 4)])])
    kf.write(b, 3, False, 1, 64, [(0, [(2, 5)])])
    keys, counts = ol.kite_union([a, b])
    assert keys[:, 0].tolist() == [1, 2, 3]
    assert counts.tolist() == [[2, 0], [0, 5], [4, 0]]


def test_serialize_codec_and_bounds(tmp_path):  # kiteops/mod.rs:948-975
    t = ol.count_fastq(b"", 3, False, 1)
    t.part_off = np.array([0, 2], dtype=np.uint64)
    t.keys = np.array([[2], [5]], dtype=np.uint64)
    t.counts = np.array([3, 7], dtype=np.uint64)
    ol.write_table(tmp_path / "u.kite", t, codec_mask=3)
    c = kf.read(tmp_path / "u.kite").chunks[0]
    assert (c.codec, c.rows, c.key_min[0], c.key_max[0]) == (2, 2, 2, 5)
    w = ol.count_fastq(b"", 65, False, 1)
    w.part_off = np.array([0, 2], dtype=np.uint64)
    m = (1 << 128) - 1
    wide = [(1 << 128) | (m - 1), (2 << 128) | 1]
    w.keys = np.array([ol.key_to_words(v, 4) for v in wide], dtype=np.uint64)
    w.counts = np.array([9, 10], dtype=np.uint64)
    ol.write_table(tmp_path / "w.kite", w, codec_mask=3)
    c = kf.read(tmp_path / "w.kite").chunks[0]
    assert (c.codec, c.raw_bytes, c.key_min[1], c.key_max[1]) == (1, 72, 1, 2)


def test_loser_tree_basic(tmp_path):  # loser_tree.rs:186-201
    f = tmp_path / "lt.kite"
    kf.write(f, 3, False, 1, 64, [(0, [(1, 1), (4, 1)]), (0, [(2, 1), (3, 2)]), (0, [(1, 2), (5, 1)])])
    got = [(brute.kmer_value(s), c) for s, c in _dump(f, tmp_path)]
    assert got == [(1, 3), (2, 1), (3, 2), (4, 1), (5, 1)]


def test_loser_tree_single_and_empty(tmp_path):  # loser_tree.rs:203-218
    f = tmp_path / "one.kite"
    kf.write(f, 3, False, 1, 64, [(0, [(10, 5), (20, 3)])])
    assert [(brute.kmer_value(s), c) for s, c in _dump(f, tmp_path)] == [(10, 5), (20, 3)]
    kf.write(f, 3, False, 1, 64, [])
    assert _dump(f, tmp_path) == []


def test_loser_tree_many_streams(tmp_path):  # loser_tree.rs:220-238
    f = tmp_path / "many.kite"
    runs = [(0, [(i, 1)]) for i in reversed(range(100))] + [(0, [(50, 3)])]
    kf.write(f, 4, False, 1, 64, runs)
    got = [(brute.kmer_value(s), c) for s, c in _dump(f, tmp_path)]
    assert [k for k, _ in got] == list(range(100)) and got[50][1] == 4


def test_dump_u32_saturation(tmp_path):  # loser_tree.rs:166 saturating_add
    f = tmp_path / "sat.kite"
    kf.write(f, 3, False, 1, 64, [(0, [(7, 0xFFFFFFF0)]), (0, [(7, 0x20)])])
    assert _dump(f, tmp_path) == [("ACT", 0xFFFFFFFF)]


def test_histo_labels_and_union_of_bins(tmp_path):  # histo/mod.rs:132-145, kiteops:123-165
    a, b = tmp_path / "sample_a.kite", tmp_path / "x.fastq.gz.kite"
    kf.write(a, 3, False, 1, 64, [(0, [(1, 2), (2, 2), (3, 5)])])
    kf.write(b, 3, False, 1, 64, [(0, [(1, 1), (9, 5)])])
    ol.kite_histo([a, b], tmp_path / "h.tsv")
    assert (tmp_path / "h.tsv").read_text() == "count\tsample_a\tx.fastq.gz\n1\t0\t1\n2\t2\t0\n5\t1\t1\n"
