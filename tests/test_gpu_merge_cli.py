"""GPU merge path and the drop-in CLI on the GPU, against the oracle.

* kite_gpu_merge_tables / kite_gpu_union_tables: sorted union with u64
  sums and the zero-filled union matrix, bit-exact.
* `kite count` -> `kite dump` byte-identical to the oracle's dump;
  `kite merge` output .kite byte-identical to the oracle's merge (same
  staging writer layout); `kite histo` identical.
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


def kite(*args, check=True):
    r = subprocess.run([str(KITE), *map(str, args)], capture_output=True, text=True, timeout=600)
    if check and r.returncode != 0:
        raise AssertionError(f"kite {args} failed ({r.returncode}):\n{r.stderr}")
    return r


def lanes():
    return min(6, max(2, ((os.cpu_count() or 1) + 7) // 8))


def random_table(rng, P, words, n_per, bits):
    keys, counts, off = [], [], [0]
    for p in range(P):
        ks = sorted(set(rng.getrandbits(bits) for _ in range(rng.randrange(n_per))))
        keys += ks
        counts += [rng.choice([1, 2, 5, 2**32 - 1, 2**40]) for _ in ks]
        off.append(len(keys))
    K = np.array([ol.key_to_words(v, words) for v in keys], dtype=np.uint64).reshape(-1, words)
    return K, np.array(counts, dtype=np.uint64), np.array(off, dtype=np.uint64), keys, counts


@pytest.mark.parametrize("words,n_in", [(1, 2), (1, 5), (2, 3), (4, 2)])
def test_merge_tables_matches_python(gpu_ctx, words, n_in):
    rng = random.Random(words * 10 + n_in)
    P = 8
    bits = {1: 40, 2: 100, 4: 200}[words]
    tabs, raw = [], []
    for i in range(n_in):
        # shared key pool so the inputs overlap
        K, C, O, keys, counts = random_table(rng, P, words, 600, 12 if i % 2 else bits)
        tabs.append((K, C, O))
        raw.append((keys, counts, O))
    ok, oc, oo = gpu_ctx.merge_tables(tabs, words)
    for p in range(P):
        acc = {}
        for keys, counts, O in raw:
            for j in range(int(O[p]), int(O[p + 1])):
                acc[keys[j]] = min(acc.get(keys[j], 0) + counts[j], 2**64 - 1)
        exp = sorted(acc.items())
        got = [(ol.words_to_int(ok[j]), int(oc[j])) for j in range(int(oo[p]), int(oo[p + 1]))]
        assert got == exp


def test_union_tables_matches_oracle(gpu_ctx, tmp_path):
    lines = (GOLD / "small.fastq").read_bytes().split(b"\n")
    paths, tabs = [], []
    for i in range(4):
        recs = lines[4 * 10 * i: 4 * (120 - 10 * i)]
        t = ol.count_fastq(b"\n".join(recs) + b"\n", 9, True, 8)
        assert t.n_unique > 0
        p = tmp_path / f"u{i}.kite"
        ol.write_table(p, t)
        paths.append(p)
        tabs.append((t.keys, t.counts, t.part_off))
    keys, matrix = ol.kite_union(paths)
    uk, um, uo = gpu_ctx.union_tables(tabs, 1)
    assert np.array_equal(uk, keys)
    assert np.array_equal(um.astype(np.uint64), matrix)


def test_cli_count_dump_matches_oracle(tmp_path):
    from kite_amd.workloads import K31_CANON_50MBP
    w = K31_CANON_50MBP.scaled(60000, genome_len=3_000_000)
    fq = tmp_path / "s.fastq"
    w.fastq().tofile(fq)
    kite("count", "-i", fq, "-o", tmp_path / "g.kite", "-k", 31, "--canonical", "true", "--partitions", 256,
         "--min", 2)
    kite("dump", "-i", tmp_path / "g.kite", "-o", tmp_path / "g.tsv")
    t = ol.count_fastq(np.fromfile(fq, dtype=np.uint8), 31, True, 256)
    ol.write_table(tmp_path / "o.kite", t, min_count=2)
    ol.kite_dump(tmp_path / "o.kite", tmp_path / "o.tsv")
    assert (tmp_path / "g.tsv").read_bytes() == (tmp_path / "o.tsv").read_bytes()
    # histo over the GPU file equals the oracle's over its own file
    kite("histo", "-i", tmp_path / "g.kite", "-o", tmp_path / "gh.tsv")
    ol.kite_histo([tmp_path / "o.kite"], tmp_path / "oh.tsv", labels=["g"])
    assert (tmp_path / "gh.tsv").read_bytes() == (tmp_path / "oh.tsv").read_bytes()


@pytest.mark.parametrize("k,canon,P", [(21, False, 256), (63, True, 64), (77, True, 4)])
def test_cli_merge_bytes_match_oracle(tmp_path, k, canon, P):
    rng = random.Random(k)
    paths = []
    for i in range(2):
        recs = []
        for r in range(3000):
            L = rng.randint(60, 200)
            s = "".join(rng.choice("ACGT") for _ in range(L))
            recs.append(f"@r{r}\n{s}\n+\n{'I' * L}\n")
        fq = tmp_path / f"in{i}.fq"
        fq.write_text("".join(recs) * 2)
        out = tmp_path / f"in{i}.kite"
        kite("count", "-i", fq, "-o", out, "-k", k, "--canonical", str(canon).lower(), "--partitions", P)
        paths.append(out)
    kite("info", "-i", paths[0], "--set-min", 2)
    kite("merge", "-i", *paths, "-o", tmp_path / "g.kite")
    ol.kite_merge(paths, tmp_path / "o.kite", lanes=lanes())
    assert (tmp_path / "g.kite").read_bytes() == (tmp_path / "o.kite").read_bytes()
    assert (tmp_path / "g.merge_report.md").exists()


def test_cli_analyze_tables(tmp_path):
    import sys
    sys.path.insert(0, str(ROOT / "oracle"))
    import corr_oracle as co
    from kite_amd.workloads import K31_CANON_50MBP
    paths = []
    for i in range(4):
        w = K31_CANON_50MBP.scaled(4000 + 2000 * i, genome_len=400_000)
        t = ol.count_fastq(w.fastq(), 21, True, 16)
        p = tmp_path / f"sample{i}.kite"
        ol.write_table(p, t, min_count=2 if i == 3 else None)
        paths.append(p)
    out = tmp_path / "an"
    kite("analyze", "-i", *paths, "-o", out, "--correlation", "all", "--group", "a", "a", "b", "b")
    keys, matrix = ol.kite_union(paths)
    rows = (out / "exact_count_matrix.tsv").read_text().splitlines()
    assert rows[0] == "kmer\tsample0\tsample1\tsample2\tsample3"
    assert len(rows) == 1 + len(keys)
    for r in range(0, len(keys), max(1, len(keys) // 500)):
        f = rows[1 + r].split("\t")
        assert f[0] == ol.decode_key(ol.words_to_int(keys[r]), 1, 21)
        assert [int(x) for x in f[1:]] == matrix[r].tolist()
    V = co.transform(matrix)

    def table(name):
        lines = (out / name).read_text().splitlines()
        assert lines[0].split("\t")[1:] == [f"sample{i}" for i in range(4)]
        return np.array([[float(x) for x in l.split("\t")[1:]] for l in lines[1:]])

    assert np.max(np.abs(table("correlation_pearson.tsv") - co.pearson(V))) <= 1e-5 + 1e-8
    assert np.max(np.abs(table("correlation_spearman.tsv") - co.spearman(V))) <= 1e-5 + 1e-8
    emb = [l.split("\t") for l in (out / "embedding_pca.tsv").read_text().splitlines()]
    assert emb[0] == ["sample", "group", "method", "dim1", "dim2"]
    assert [e[1] for e in emb[1:]] == ["a", "a", "b", "b"] and {e[2] for e in emb[1:]} == {"pca"}
    S, _ = co.pca_scores(co.covariance(V), 2)
    got = np.array([[float(e[3]), float(e[4])] for e in emb[1:]])
    assert np.max(np.abs(co.sign_normalize(got) - co.sign_normalize(S))) <= 1e-4 * max(1, np.abs(S).max())
