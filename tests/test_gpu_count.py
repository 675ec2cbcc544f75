"""Parity of the gfx950 count pipeline (through the C ABI) with the oracle.

Bit-exact: per-partition unique keys, their order and their counts
(saturated at u32::MAX) must equal the oracle's on the same bytes."""
import json
import random
from pathlib import Path

import numpy as np
import pytest

import brute
import oracle_lib as ol

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def check_same(res, table, label=""):
    assert res.partitions == table.partitions, label
    assert np.array_equal(res.part_off, table.part_off), f"{label}: partition offsets differ"
    assert res.keys.shape == table.keys.shape, f"{label}: {res.keys.shape} vs {table.keys.shape}"
    assert np.array_equal(res.keys, table.keys), f"{label}: keys differ"
    exp = np.minimum(table.counts, np.uint64(0xFFFFFFFF)).astype(np.uint32)
    assert np.array_equal(res.counts, exp), f"{label}: counts differ"


def run_both(ctx, fq, k, canon, P, chunk_size=None):
    arr = np.frombuffer(fq, dtype=np.uint8) if isinstance(fq, (bytes, bytearray)) else fq
    chunks = ol.plain_chunks(arr, chunk_size) if chunk_size else None
    res, st = ctx.count(arr, k, canon, P, chunks=chunks)
    tab = ol.count_fastq(arr, k, canon, P, chunk_size=chunk_size or (1 << 40), threads=8)
    assert st.kmers == tab.n_kmers
    return res, tab, st


def dump_of(res) -> bytes:
    out = []
    for i in range(len(res.counts)):
        v = ol.words_to_int(res.keys[i])
        out.append(f"{ol.decode_key(v, res.words, res.k)}\t{int(res.counts[i])}\n")
    return "".join(out).encode()


@pytest.mark.parametrize("name", list(json.loads((GOLD / "small.json").read_text())))
def test_golden_fixture(gpu_ctx, name):
    m = json.loads((GOLD / "small.json").read_text())[name]
    fq = (GOLD / "small.fastq").read_bytes()
    res, st = gpu_ctx.count(fq, m["k"], m["canonical"], m["partitions"])
    assert dump_of(res) == (GOLD / f"small_{name}.tsv").read_bytes()


@pytest.mark.parametrize("k,canon,P", [(1, False, 1), (5, True, 2), (21, False, 256), (31, True, 256),
                                       (32, True, 64), (33, False, 16), (63, True, 1024),
                                       (64, True, 8), (65, True, 4), (96, False, 32), (128, True, 2)])
def test_random_reads_all_key_widths(gpu_ctx, k, canon, P):
    rng = random.Random(k * 131 + P)
    recs = []
    for i in range(3000):
        L = rng.randint(0, 300)
        s = "".join(rng.choice("ACGTACGTACGTACGTNacgt") for _ in range(L))
        recs.append(f"@r{i} x\n{s}\n+\n{'I' * L}\n")
    fq = "".join(recs).encode()
    res, tab, _ = run_both(gpu_ctx, fq, k, canon, P)
    check_same(res, tab, f"k={k}")


def test_every_byte_value_in_sequences(gpu_ctx):
    """Sequence lines drawn from all byte values but the newline: only
    A, C, G, T in either case are bases, every other byte breaks the k-mer
    (BASE_ENCODE, src/count/kmer.rs:2-13)."""
    rng = random.Random(99)
    pool = bytes(b for b in range(256) if b != 0x0A) + b"ACGTacgt" * 40
    recs = []
    for i in range(4000):
        L = rng.randint(0, 200)
        s = bytes(rng.choice(pool) for _ in range(L))
        recs.append(b"@e%d\n" % i + s + b"\n+\n" + b"I" * L + b"\n")
    fq = b"".join(recs)
    for k, canon, P in [(3, True, 4), (7, False, 16), (21, True, 64), (40, True, 8)]:
        res, tab, _ = run_both(gpu_ctx, fq, k, canon, P)
        check_same(res, tab, f"bytes k={k}")


def test_long_reads_cross_subtiles(gpu_ctx):
    rng = random.Random(7)
    recs = []
    for i in range(40):
        L = rng.randint(5000, 40000)
        s = "".join(rng.choice("ACGT") for _ in range(L))
        if i % 3 == 0:
            s = s[:L // 2] + "N" + s[L // 2 + 1:]
        recs.append(f"@long{i}\n{s}\n+\n{'I' * L}\n")
    fq = "".join(recs).encode()
    for k, canon in [(21, True), (47, False), (101, True)]:
        res, tab, _ = run_both(gpu_ctx, fq, k, canon, 64)
        check_same(res, tab, f"long k={k}")


def test_line_endings_and_unterminated(gpu_ctx):
    fq = (b"@a\r\nACGTACGTACGTTTGCA\r\n+\r\nIIIIIIIIIIIIIIIII\r\n"
          b"@b\nacgtacgtAAAACCCCGGGG\n+\nIIIIIIIIIIIIIIIIIIII\n"
          b"@c\nTTTTGGGGCCCCAAAATTTT")
    for k in (3, 9, 15):
        res, tab, _ = run_both(gpu_ctx, fq, k, True, 4)
        check_same(res, tab)
        assert dump_of(res) == brute.dump(fq, k, True, 4)


def test_empty_and_degenerate_inputs(gpu_ctx):
    for fq in (b"", b"@", b"@r\n", b"ACGT\nACGT\n+\nIIII\n", b"@r\nACG\n+\nIII\n", b"\n\n\n\n"):
        res, tab, st = run_both(gpu_ctx, fq, 4, False, 8)
        check_same(res, tab, repr(fq))


def test_identical_reads_force_deep_splits(gpu_ctx):
    # one 150 bp read repeated: few unique keys with huge counts; poly-A: one key
    rng = random.Random(3)
    s = "".join(rng.choice("ACGT") for _ in range(150))
    fq = ("".join(f"@r{i}\n{s}\n+\n{'I' * 150}\n" for i in range(20000))).encode()
    res, tab, st = run_both(gpu_ctx, fq, 31, True, 4)
    check_same(res, tab, "identical reads")
    fq = ("".join(f"@r{i}\n{'A' * 150}\n+\n{'I' * 150}\n" for i in range(20000))).encode()
    for k in (21, 64, 99):
        res, tab, st = run_both(gpu_ctx, fq, k, False, 2)
        check_same(res, tab, f"poly-A k={k}")
        assert len(res.counts) == 1 and int(res.counts[0]) == 20000 * (150 - k + 1)


def test_host_chunks_with_dropped_chunk(gpu_ctx):
    good = b"@a\nACGTACGTACGGTTCA\n+\nIIIIIIIIIIIIIIII\n"
    bad = b"xa\nACGTACGTACGGTTCA\n+\nIIIIIIIIIIIIIIII\n"
    assert len(good) == len(bad)
    fq = good * 4 + bad + good * 3 + b"@z\nTTTTTTTTCCCC"
    res, tab, st = run_both(gpu_ctx, fq, 5, True, 2, chunk_size=2 * len(good))
    check_same(res, tab, "chunked")
    assert st.skipped_chunks == 1


@pytest.mark.parametrize("P", [1, 2, 4096, 8192])
def test_partition_counts(gpu_ctx, P):
    from kite_amd.workloads import K21_1MBP
    fq = K21_1MBP.fastq(n_reads=20000)
    res, tab, _ = run_both(gpu_ctx, fq, 21, False, P)
    check_same(res, tab, f"P={P}")


@pytest.mark.parametrize("wname,n_reads", [("k21-1mbp-200k", 200000), ("k31-canon-50mbp-5m", 300000),
                                           ("k63-canon-100mbp-10m", 100000)])
def test_workload_prefixes(gpu_ctx, wname, n_reads):
    from kite_amd.workloads import WORKLOADS
    w = WORKLOADS[wname]
    fq = w.fastq(n_reads=n_reads)
    res, tab, st = run_both(gpu_ctx, fq, w.k, w.canonical, w.partitions)
    check_same(res, tab, wname)


def test_full_size_properties(gpu_ctx):
    """Size-independent checks on a 2M-read k=31 canonical workload prefix:
    totals, order, partition membership and agreement on a key sample."""
    from kite_amd.workloads import K31_CANON_50MBP as w
    fq = w.fastq(n_reads=2_000_000)
    res, st = gpu_ctx.count(fq, w.k, w.canonical, w.partitions)
    assert int(res.counts.astype(np.uint64).sum()) == st.kmers
    k = res.keys[:, 0]
    P = res.partitions
    bits = P.bit_length() - 1
    part = ((k * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(64 - bits)).astype(np.int64)
    for p in range(P):
        a, b = int(res.part_off[p]), int(res.part_off[p + 1])
        assert np.all(part[a:b] == p)
        assert np.all(np.diff(k[a:b].astype(np.int64)) > 0) if b - a > 1 else True
    # every valid window of the FASTQ bytes is one input k-mer
    assert st.kmers == brute.exact_windows(fq, w.k)


@pytest.mark.parametrize("dedup", ["0", "1"])
@pytest.mark.parametrize("split_div", ["1", "4", "16"])
def test_sort_variants_agree(gpu_ctx, monkeypatch, dedup, split_div):
    """Both LDS bucket sorts (plain LSD, hash-dedup first) and several split
    widths give the oracle's table."""
    monkeypatch.setenv("KITE_SORT_DEDUP", dedup)
    monkeypatch.setenv("KITE_SPLIT_DIV", split_div)
    from kite_amd.workloads import K31_CANON_50MBP, K63_CANON_100MBP
    for w, n in ((K31_CANON_50MBP, 150000), (K63_CANON_100MBP, 40000)):
        fq = w.fastq(n_reads=n)
        res, tab, _ = run_both(gpu_ctx, fq, w.k, w.canonical, 64)
        check_same(res, tab, f"{w.name} dedup={dedup} div={split_div}")
    rng = random.Random(11)
    s = "".join(rng.choice("ACGT") for _ in range(150))
    fq = ("".join(f"@r{i}\n{s}\n+\n{'I' * 150}\n" for i in range(5000))).encode()
    for k in (15, 40, 90):
        res, tab, _ = run_both(gpu_ctx, fq, k, True, 2)
        check_same(res, tab, f"repeats k={k} dedup={dedup}")


@pytest.mark.parametrize("dedup", ["0", "1"])
def test_low_complexity_keys(gpu_ctx, monkeypatch, dedup):
    """Reads of a long A run followed by random bases: most distinct keys
    share their top key bits, so whole partitions land in one LDS bin (the
    hash-then-bin sort then takes its LSD fallback)."""
    monkeypatch.setenv("KITE_SORT_DEDUP", dedup)
    rng = random.Random(21)
    recs = []
    for i in range(400):
        tail = "".join(rng.choice("ACGT") for _ in range(40))
        s = "A" * 100 + tail
        recs.append(f"@lc{i}\n{s}\n+\n{'I' * len(s)}\n")
    fq = "".join(recs).encode()
    for k, canon, P in ((31, False, 8), (31, True, 4), (45, False, 8), (21, False, 1)):
        res, tab, _ = run_both(gpu_ctx, fq, k, canon, P)
        check_same(res, tab, f"low complexity k={k} canon={canon} P={P} dedup={dedup}")


@pytest.mark.parametrize("chunk,subs", [("3", "3"), ("1", "7"), ("8", "16")])
def test_workgroup_chunking_tunables(gpu_ctx, monkeypatch, chunk, subs):
    """Split chunks of several tiles and extract workgroups of several
    sub-tiles (including counts that do not divide the work) give the
    oracle's table."""
    monkeypatch.setenv("KITE_SPLIT_CHUNK", chunk)
    monkeypatch.setenv("KITE_EXTRACT_SUBS", subs)
    from kite_amd.workloads import K31_CANON_50MBP, K63_CANON_100MBP
    for w, n, P in ((K31_CANON_50MBP, 200000, 16), (K63_CANON_100MBP, 30000, 64)):
        fq = w.fastq(n_reads=n)
        res, tab, _ = run_both(gpu_ctx, fq, w.k, w.canonical, P)
        check_same(res, tab, f"{w.name} chunk={chunk} subs={subs}")


@pytest.mark.parametrize("staged", ["0", "1"])
def test_extract_scatter_modes(gpu_ctx, monkeypatch, staged):
    """The extract scatter with keys grouped by partition in LDS (chunks of
    1024 / W keys) and with direct per-lane stores give the oracle's table,
    across key widths, partition counts at the staging limit, and long reads
    whose sub-tiles hold more k-mers than one chunk."""
    monkeypatch.setenv("KITE_EXTRACT_STAGED", staged)
    rng = random.Random(5)
    recs = []
    for i in range(1500):
        L = rng.choice([0, 40, 150, 151, 2500])
        s = "".join(rng.choice("ACGTACGTACGTN") for _ in range(L))
        recs.append(f"@s{i}\n{s}\n+\n{'I' * L}\n")
    fq = "".join(recs).encode()
    for k, canon, P in [(31, True, 512), (31, True, 256), (25, False, 1), (63, True, 256), (45, False, 64),
                        (101, True, 128), (128, False, 4)]:
        res, tab, _ = run_both(gpu_ctx, fq, k, canon, P)
        check_same(res, tab, f"staged={staged} k={k} P={P}")


@pytest.mark.parametrize("staged", ["1", "0"])
@pytest.mark.parametrize("max_bits", ["12", "5"])
def test_split_scatter_modes(gpu_ctx, monkeypatch, staged, max_bits):
    """The MSD split scatter with each tile regrouped by child in LDS and
    with direct per-lane stores give the oracle's table, for 1-, 2- and
    4-word keys, with partitions large enough for splits of 11 and 12 bits
    and, with the digit capped at 5 bits, several split levels."""
    monkeypatch.setenv("KITE_SPLIT_STAGED", staged)
    monkeypatch.setenv("KITE_MAX_DBITS", max_bits)
    from kite_amd.workloads import K31_CANON_50MBP, K63_CANON_100MBP
    rng = random.Random(11)
    long_reads = "".join(f"@l{i}\n{s}\n+\n{'I' * len(s)}\n" for i, s in enumerate(
        "".join(rng.choice("ACGT") for _ in range(400)) for _ in range(3000))).encode()
    for fq, k, canon, P in ((K31_CANON_50MBP.fastq(n_reads=150000), 31, True, 4),
                            (K31_CANON_50MBP.fastq(n_reads=150000), 31, True, 1),
                            (K63_CANON_100MBP.fastq(n_reads=30000), 63, True, 2),
                            (long_reads, 101, False, 1)):
        res, tab, _ = run_both(gpu_ctx, fq, k, canon, P)
        check_same(res, tab, f"split staged={staged} max_bits={max_bits} k={k} P={P}")


@pytest.mark.parametrize("classes,div", [("3", "4"), ("4", "1"), ("4", "2")])
def test_hash_kernel_on_large_classes(gpu_ctx, monkeypatch, classes, div):
    """The hash kernel with 1024 threads on the 4096-key class and, with
    four hash classes, on the largest class (8 / W keys per thread), where
    a split target of the whole largest capacity fills that class."""
    monkeypatch.setenv("KITE_HASH_CLASSES", classes)
    monkeypatch.setenv("KITE_SPLIT_DIV", div)
    from kite_amd.workloads import K31_CANON_50MBP, K63_CANON_100MBP
    for w, n, P in ((K31_CANON_50MBP, 300000, 8), (K63_CANON_100MBP, 40000, 4)):
        fq = w.fastq(n_reads=n)
        res, tab, _ = run_both(gpu_ctx, fq, w.k, w.canonical, P)
        check_same(res, tab, f"{w.name} hash classes {classes} split div {div}")


def test_autotune_choice_keeps_parity(gpu_ctx):
    """kite_gpu_autotune times its candidate settings on a sample; later
    counts of that size class on the context use the choice, and must still
    give the oracle's table (one-word and two-word keys)."""
    from kite_amd.workloads import K31_CANON_50MBP, K63_CANON_100MBP
    from kite_amd import GpuContext
    with GpuContext(0) as ctx:  # its own context: the tuned settings stay with it
        for w, n in ((K31_CANON_50MBP, 150000), (K63_CANON_100MBP, 40000)):
            fq = w.fastq(n_reads=n)
            rep = ctx.autotune(fq, w.k, w.canonical, w.partitions)
            assert rep["words"] == (1 if w.k <= 31 else 2) and rep["size_class"] == 0
            assert rep["trials"] >= 12 and rep["tuned_ms"] <= rep["default_ms"]
            assert rep["chosen"]["max_digit_bits"] in (8, 10, 11, 12)
            res, tab, _ = run_both(ctx, fq, w.k, w.canonical, w.partitions)
            check_same(res, tab, f"{w.name} after autotune {rep['chosen']}")


def test_set_tuned_settings_keep_parity():
    """kite_gpu_set_tuned applies fixed settings to one size class without
    timing: counts of that class must still give the oracle's table, here
    with every setting away from the built-in one; values outside the
    autotuner's candidates are rejected."""
    from kite_amd.workloads import K31_CANON_50MBP
    from kite_amd import GpuContext, KiteError
    w = K31_CANON_50MBP
    fq = w.fastq(n_reads=120000)
    with GpuContext(0) as ctx:
        chosen = {"split_target_div": 8, "max_digit_bits": 8, "hash_classes": 2, "extract_subs": 1,
                  "split_above_div": 2}
        ctx.set_tuned(w.k, w.partitions, 0, chosen)
        res, tab, _ = run_both(ctx, fq, w.k, w.canonical, w.partitions)
        check_same(res, tab, f"{w.name} with fixed settings {chosen}")
        for bad in ({**chosen, "split_target_div": 3}, {**chosen, "max_digit_bits": 13},
                    {**chosen, "extract_subs": 0}):
            with pytest.raises(KiteError, match="invalid count settings"):
                ctx.set_tuned(w.k, w.partitions, 0, bad)
        with pytest.raises(KiteError, match="invalid count settings"):
            ctx.set_tuned(w.k, w.partitions, 4, chosen)
