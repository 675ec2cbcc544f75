"""`kite profile` on the GPU union against the oracle's union.

The expected table is built from the oracle's union (per-input merged,
header-filtered u64 counts, zero where absent), formatted the way
render_partition_rows does (src/profile/mod.rs:140-169): "kmer" then one
column per label, a row per union k-mer in partition then key order.
"""
import os
import random
import subprocess
from pathlib import Path

import numpy as np
import pytest

import kite_file as kf
import oracle_lib as ol

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
KITE = ROOT / "kite_amd" / "lib" / "kite"
GOLD = ROOT / "tests" / "golden"


def kite(*args, check=True, env=None):
    r = subprocess.run([str(KITE), *map(str, args)], capture_output=True, text=True, timeout=600, env=env)
    if check and r.returncode != 0:
        raise AssertionError(f"kite {args} failed ({r.returncode}):\n{r.stderr}")
    return r


def expected(paths, labels, words, k):
    keys, matrix = ol.kite_union(paths)
    out = ["kmer\t" + "\t".join(labels)]
    for r in range(len(keys)):
        kmer = ol.decode_key(ol.words_to_int(keys[r]), words, k)
        out.append(kmer + "".join(f"\t{int(c)}" for c in matrix[r]))
    return "\n".join(out) + "\n", matrix


def saturating_runs(rng, P, n_runs, key_bits, shared):
    runs = []
    for _ in range(n_runs):
        keys = sorted(set(rng.sample(range(1 << key_bits), 50)) | set(rng.sample(shared, 8)))
        runs.append((rng.randrange(P), [(x, rng.choice([1, 2, 3, 9, 0x7FFFFFFF, 0xFFFFFFF0])) for x in keys]))
    return runs


def test_profile_saturated_sums_and_filters(tmp_path):
    """Multi-run partitions whose sums pass u32::MAX (exact u64 in the table),
    header filters on two of the inputs, labels inferred from file names."""
    rng = random.Random(5)
    shared = list(range(0, 1 << 22, 1 << 16))
    filters = [(None, None), (2, None), (None, 0xFFFFFFFF)]
    names = ["alpha.kite", "beta.fastq.gz", "gamma"]
    paths = []
    for i, (mn, mx) in enumerate(filters):
        p = tmp_path / names[i]
        kf.write(p, 11, False, 8, 64, saturating_runs(rng, 8, 30, 22, shared), min_count=mn, max_count=mx)
        paths.append(p)
    want, matrix = expected(paths, ["alpha", "beta", "gamma"], 1, 11)
    assert matrix.max() > 0xFFFFFFFF, "the case must hold sums above u32::MAX"
    r = kite("profile", "-i", *paths, "-o", tmp_path / "p.tsv", "--test-mode")
    assert (tmp_path / "p.tsv").read_text() == want
    assert f"rows={len(matrix)}" in r.stderr and "labels=alpha,beta,gamma (inferred from file names)" in r.stderr
    # the same table when the partitions go through the GPU in groups of
    # one and of three
    for rows in ("1", "2000"):
        env = dict(os.environ, KITE_PROFILE_GROUP_ROWS=rows)
        kite("profile", "-i", *paths, "-o", tmp_path / f"g{rows}.tsv", env=env)
        assert (tmp_path / f"g{rows}.tsv").read_text() == want, f"group rows {rows}"


@pytest.mark.parametrize("kw,k,canon", [(128, 33, True), (256, 65, False)])
def test_profile_wide_keys_user_labels(tmp_path, kw, k, canon):
    rng = random.Random(kw + k)
    paths = []
    pool = [rng.randrange(1 << (2 * k)) for _ in range(300)]
    for i in range(3):
        runs = []
        for _ in range(6):
            keys = sorted(set(rng.sample(pool, 60)))
            runs.append((rng.randrange(4), [(v, rng.randrange(1, 9)) for v in keys]))
        p = tmp_path / f"in{i}.kite"
        kf.write(p, k, canon, 4, kw, runs)
        paths.append(p)
    labels = ["x", "y", "z"]
    want, _ = expected(paths, labels, kw // 64, k)
    out = tmp_path / "nested" / "dir" / "p.tsv"
    kite("profile", "-i", *paths, "-o", out, "--label", *labels, "-t", 3)
    assert out.read_text() == want
    r = kite("profile", "-i", *paths, "-o", "-", "--label", *labels)
    assert r.stdout == want


def test_profile_of_counted_samples(tmp_path):
    """Samples counted by the oracle from the golden reads (overlapping
    subsets), so the union has shared, private and missing k-mers; the last
    sample's min_count=2 filter removes all of its rows (a zero column)."""
    lines = (GOLD / "small.fastq").read_bytes().split(b"\n")
    paths = []
    for i in range(4):
        recs = lines[4 * 10 * i: 4 * (120 - 10 * i)]
        t = ol.count_fastq(b"\n".join(recs) + b"\n", 21, True, 16)
        p = tmp_path / f"s{i}.kite"
        ol.write_table(p, t, min_count=2 if i == 3 else None)
        paths.append(p)
    want, matrix = expected(paths, [f"s{i}" for i in range(4)], 1, 21)
    assert (matrix[:, :3] == 0).any() and (matrix[:, :3] > 0).all(axis=1).any() and not matrix[:, 3].any()
    kite("profile", "-i", *paths, "-o", tmp_path / "p.tsv")
    assert (tmp_path / "p.tsv").read_text() == want


def test_profile_errors(tmp_path):
    t = ol.count_fastq((GOLD / "small.fastq").read_bytes(), 15, True, 8)
    a, b = tmp_path / "a.kite", tmp_path / "b.kite"
    ol.write_table(a, t)
    t2 = ol.count_fastq((GOLD / "small.fastq").read_bytes(), 17, True, 8)
    ol.write_table(b, t2)
    r = kite("profile", "-i", a, a, "-o", tmp_path / "p.tsv", "--label", "one", check=False)
    assert r.returncode == 1 and "--label count (1) must match input count (2)" in r.stderr
    r = kite("profile", "-i", a, b, "-o", tmp_path / "p.tsv", check=False)
    assert r.returncode == 1 and "incompatible" in r.stderr
