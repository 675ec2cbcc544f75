"""The oracle against brute force and the committed golden fixtures.

Brute force (tests/brute.py) derives every expected value from the
mathematics of the path -- string windows, string reverse complements,
Counter -- so a dropped term, a wrong shift or a swapped word in the oracle
fails here."""
import json
import random
from pathlib import Path

import numpy as np
import pytest

import brute
import kite_file as kf
import oracle_lib as ol

GOLD = Path(__file__).resolve().parent / "golden"


def rand_seq(rng, n, alphabet="ACGT"):
    return "".join(rng.choice(alphabet) for _ in range(n))


@pytest.mark.parametrize("k", [1, 2, 5, 15, 21, 31, 32, 33, 47, 63, 64, 65, 77, 100, 127, 128])
@pytest.mark.parametrize("canonical", [False, True])
def test_extract_matches_bruteforce(k, canonical):
    rng = random.Random(k * 7 + canonical)
    for _ in range(6):
        s = rand_seq(rng, rng.randint(0, 300), "ACGTACGTACGTacgtN\r")
        exp = brute.kmers(s, k, canonical)
        assert ol.extract(s.encode("latin-1"), k, canonical) == exp


def test_k32_both_iterators_agree():
    rng = random.Random(32)
    s = rand_seq(rng, 200).encode()
    assert ol.extract(s, 32, True, 1) == ol.extract(s, 32, True, 2) == brute.kmers(s.decode(), 32, True)


@pytest.mark.parametrize("k", [3, 21, 31, 32, 63, 64, 65, 128])
def test_bucket_matches_closed_form(k):
    rng = random.Random(k)
    w = brute.words_for(k)
    for bits in (0, 1, 8, 10, 13):
        for _ in range(50):
            v = rng.getrandbits(2 * k)
            assert ol.bucket(v, w, bits) == brute.partition(v, k, 1 << bits)


@pytest.mark.parametrize("words", [1, 2, 4])
def test_dvle_matches_independent_encoder(words):
    rng = random.Random(words)
    bits = {1: 62, 2: 126, 4: 250}[words]
    keys = sorted(set(rng.getrandbits(bits) for _ in range(300)))
    counts = [rng.choice([1, 2, 127, 128, 300, 2**32 - 1]) for _ in keys]
    kw = {1: 64, 2: 128, 4: 256}[words]
    codec, exp = kf.encode_run(list(zip(keys, counts)), kw)
    arr = np.array([ol.key_to_words(v, words) for v in keys], dtype=np.uint64)
    assert ol.dvle_encode(arr, np.array(counts, dtype=np.uint32), words) == exp


def _meta():
    return json.loads((GOLD / "small.json").read_text())


@pytest.mark.parametrize("name", list(json.loads((GOLD / "small.json").read_text())))
def test_golden_dump(name, tmp_path):
    m = _meta()[name]
    fq = (GOLD / "small.fastq").read_bytes()
    t = ol.count_fastq(fq, m["k"], m["canonical"], m["partitions"], chunk_size=997, threads=3)
    ol.write_table(tmp_path / "o.kite", t)
    ol.kite_dump(tmp_path / "o.kite", tmp_path / "o.tsv")
    assert (tmp_path / "o.tsv").read_bytes() == (GOLD / f"small_{name}.tsv").read_bytes()


def test_chunking_is_invisible_for_wellformed_fastq():
    fq = (GOLD / "small.fastq").read_bytes()
    a = ol.count_fastq(fq, 11, True, 8, chunk_size=1 << 20, threads=1)
    b = ol.count_fastq(fq, 11, True, 8, chunk_size=300, threads=4)
    assert np.array_equal(a.keys, b.keys) and np.array_equal(a.counts, b.counts)
    assert a.n_kmers == b.n_kmers == sum(int(c) for c in a.counts)


def test_chunk_not_starting_with_at_is_dropped():
    good = b"@a\nACGTACGTAC\n+\nIIIIIIIIII\n"
    bad = b"ja\nACGTACGTAC\n+\nIIIIIIIIII\n"
    assert len(good) == len(bad) == 27
    fq = good * 2 + bad + good
    # a 54-byte chunk size ends every read exactly on a record pair: chunks are
    # (good, good) and (bad, good); the second does not start with '@'
    assert ol.plain_chunks(fq, 54) == [(0, 54), (54, 54)]
    t = ol.count_fastq(fq, 5, False, 1, chunk_size=54, threads=1)
    assert t.n_reads == 2
    whole = ol.count_fastq(fq, 5, False, 1, chunk_size=1 << 20, threads=1)
    assert whole.n_reads == 4  # one chunk: '@' is only checked at the chunk start


def test_count_matches_bruteforce_random(tmp_path):
    rng = random.Random(99)
    recs = []
    for i in range(200):
        L = rng.randint(0, 60)
        s = rand_seq(rng, L, "AAACCGGTTTN")
        recs.append(f"@r{i}\n{s}\n+\n{'#' * L}\n")
    fq = "".join(recs).encode()
    for k, canon, P in [(4, False, 2), (7, True, 4), (31, True, 2), (40, False, 8), (66, True, 2)]:
        t = ol.count_fastq(fq, k, canon, P, chunk_size=512, threads=2)
        got = ol.dump_tsv_from_table(t)
        assert got == brute.dump(fq, k, canon, P)


def test_writer_reader_roundtrip_multirun(tmp_path):
    runs = [(1, [(3, 1), (9, 2)]), (0, [(4, 5)]), (1, [(3, 4)]), (3, [(10, 1), (20, 1)])]
    kf.write(tmp_path / "r.kite", 5, True, 4, 64, runs, min_count=1)
    t, marks = ol.kite_load(tmp_path / "r.kite")
    assert marks == {"min": 1, "max": None}
    assert t.part_off.tolist() == [0, 1, 3, 3, 5]
    assert t.keys[:, 0].tolist() == [4, 3, 9, 10, 20]
    assert t.counts.tolist() == [5, 5, 2, 1, 1]


def test_exact_windows_counts():
    """brute.exact_windows (the vectorised window count used at workload
    size) on hand-checked records and against brute.kmers on random
    records of several lengths with N, lower case and other bytes."""
    fq = np.frombuffer(b"@a\nACGTA\n+\nIIIII\n@b\nACNGTACG\n+\nIIIIIIII\n@c\nac\n+\nII\n", dtype=np.uint8)
    # k=2: a has 4; b has AC (1) + GTACG (4); c has 1 -> 10
    assert brute.exact_windows(fq, 2) == 10
    # k=5: a has 1; b: GTACG has 1; c none -> 2
    assert brute.exact_windows(fq, 5) == 2
    assert brute.exact_windows(fq, 6) == 0
    rng = random.Random(17)
    recs = []
    for i in range(600):
        L = rng.choice([0, 3, 31, 150, 151])
        s = "".join(rng.choice("ACGTACGTACGTacgtNRx") for _ in range(L))
        recs.append(f"@w{i}\n{s}\n+\n{'I' * L}\n")
    text = "".join(recs)
    arr = np.frombuffer(text.encode(), dtype=np.uint8)
    for k in (1, 5, 21, 31):
        exp = sum(len(brute.kmers(s, k, False)) for s in brute.fastq_sequences(text.encode()))
        assert brute.exact_windows(arr, k, reads_per_block=37) == exp, k


def test_reference_shaped_count_matches_checker():
    """ok_count_ref (the CPU baseline's restatement of the reference count
    loop: write-combined partition slots, MSD/LSD byte radix per worker run,
    RLE, DVLE) must give the checker's totals once its per-worker runs are
    merged: on the brute-force records, with chunks that do not start with
    '@', and on partitions large enough for the MSD path (>= 131072 keys in
    one worker's run)."""
    rng = random.Random(7)
    recs = []
    for i in range(300):
        L = rng.randint(0, 80)
        recs.append(f"@r{i}\n{rand_seq(rng, L, 'AAACCGGTTTN')}\n+\n{'#' * L}\n")
    fq = "".join(recs).encode()
    for k, canon, P, chunk, threads in [(4, False, 2, 512, 2), (11, True, 4, 300, 3), (21, False, 1, 1 << 20, 1),
                                        (31, True, 8, 700, 4)]:
        ref, st = ol.count_ref(fq, k, canon, P, chunk_size=chunk, threads=threads)
        chk = ol.count_fastq(fq, k, canon, P, chunk_size=chunk, threads=threads)
        assert np.array_equal(ref.part_off, chk.part_off), (k, P)
        assert np.array_equal(ref.keys, chk.keys) and np.array_equal(ref.counts, chk.counts), (k, P)
        assert st["kmers"] == chk.n_kmers and st["reads"] == chk.n_reads
        assert st["run_rows"] >= chk.n_unique and st["encoded_bytes"] > 0
    good = b"@a\nACGTACGTAC\n+\nIIIIIIIIII\n"
    bad = b"ja\nACGTACGTAC\n+\nIIIIIIIIII\n"
    fq2 = good * 2 + bad + good
    ref, st = ol.count_ref(fq2, 5, False, 1, chunk_size=54, threads=2)
    assert st["reads"] == 2 and ref.n_reads == 2
    # one partition of 400 k random 31-mers from a 300 k-base genome: the
    # worker's run takes the MSD path with leaves below 1024 keys
    g = "".join(rng.choice("ACGT") for _ in range(300_000))
    reads = []
    for i in range(4000):
        s = rng.randrange(0, len(g) - 150)
        reads.append(f"@m{i}\n{g[s:s + 150]}\n+\n{'I' * 150}\n")
    fq3 = "".join(reads).encode()
    ref, st = ol.count_ref(fq3, 31, True, 1, chunk_size=1 << 24, threads=1)
    chk = ol.count_fastq(fq3, 31, True, 1, chunk_size=1 << 24, threads=1)
    assert st["kmers"] == 4000 * 120 >= 131072
    assert np.array_equal(ref.keys, chk.keys) and np.array_equal(ref.counts, chk.counts)
    # the run holds the unique keys of the one worker, DVLE-encoded
    assert st["run_rows"] == chk.n_unique
    assert st["encoded_bytes"] == len(ol.dvle_encode(chk.keys, chk.counts.astype(np.uint32), 1))
    with pytest.raises(RuntimeError, match="one-word keys"):
        ol.count_ref(fq, 32, True, 2)
