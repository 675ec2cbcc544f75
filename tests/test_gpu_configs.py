"""The workload configurations end to end on the GPU, against the oracle.

* k=31 canonical, 5,000,000 reads (the whole workload): the GPU table equals
  the oracle's row for row, and both equal an exact window count taken
  straight from the FASTQ bytes.
* k=63 canonical, two samples of one genome (different read seeds), P=1024:
  `kite count` on each, `kite merge` byte-identical to the oracle's merge
  of the same files, and `kite dump` of the merge identical to the dump of
  the oracle's own count -> merge chain (run_merge, src/merge/mod.rs:22-214;
  merge_vectors_total, src/kiteops/mod.rs:587-724).
* The 24-sample cohort (20 Mbp base + 2% SNVs per sample) at 2% of its read
  counts: counted through the pipelined batch, then the union at --min 2,
  Pearson, Spearman, covariance and PCA (kite_gpu_analyze_tables and the
  `kite analyze` CLI) against the oracle's union and the FP64 numpy oracle
  (run_analyze, src/analyze/mod.rs:847-1326).
* `kite count` without --partitions: the auto plan's P and a dump equal to
  the oracle's at that P.
"""
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as ol
from brute import exact_windows

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
KITE = ROOT / "kite_amd" / "lib" / "kite"
sys.path.insert(0, str(ROOT / "oracle"))
import corr_oracle as co  # noqa: E402

THREADS = max(1, min(16, os.cpu_count() or 1))


def kite(*args, env=None):
    r = subprocess.run([str(KITE), *map(str, args)], capture_output=True, text=True, timeout=600,
                       env=None if env is None else dict(os.environ, **env))
    assert r.returncode == 0, f"kite {args} failed ({r.returncode}):\n{r.stderr}"
    return r


def test_config2_full_workload_matches_oracle(gpu_ctx):
    """k=31 canonical, all 5,000,000 reads of the 50 Mbp repeat-rich genome."""
    from kite_amd.workloads import K31_CANON_50MBP as w
    fq = w.fastq()
    res, st = gpu_ctx.count(fq, w.k, w.canonical, w.partitions)
    exp_windows = exact_windows(fq, w.k)
    assert st.kmers == exp_windows
    tab = ol.count_fastq(fq, w.k, w.canonical, w.partitions, chunk_size=64 << 20, threads=THREADS)
    del fq
    assert tab.n_kmers == exp_windows
    assert np.array_equal(res.part_off, tab.part_off)
    assert np.array_equal(res.keys, tab.keys)
    assert np.array_equal(res.counts, np.minimum(tab.counts, 0xFFFFFFFF).astype(np.uint32))
    assert int(res.counts.astype(np.uint64).sum()) == exp_windows


def _dump_bytes(path) -> bytes:
    return subprocess.run([str(KITE), "dump", "-i", str(path), "-o", "-"], capture_output=True,
                          timeout=600, check=True).stdout


@pytest.mark.parametrize("genome_len", [100_000_000, 1_000_000])
def test_config3_merge_leg(tmp_path, genome_len):
    """Two k=63 canonical samples of one genome with different read seeds,
    P=1024: on the workload's 100 Mbp genome (little overlap at this read
    count) and on a 1 Mbp genome (7x coverage each: most keys in both)."""
    from kite_amd.workloads import K63_CANON_100MBP
    import dataclasses
    base = K63_CANON_100MBP.scaled(30_000, genome_len=genome_len)
    samples = [dataclasses.replace(base, seed=3003, genome_seed=3003),
               dataclasses.replace(base, seed=3004, genome_seed=3003, name="k63-second")]
    genome = samples[0].genome()
    gk, ok_ = [], []
    for i, w in enumerate(samples):
        fq = w.fastq(genome=genome)
        p = tmp_path / f"s{i}.fastq"
        fq.tofile(p)
        g = tmp_path / f"g{i}.kite"
        kite("count", "-i", p, "-o", g, "-k", 63, "--canonical", "true", "--partitions", 1024)
        gk.append(g)
        t = ol.count_fastq(fq, 63, True, 1024, threads=THREADS)
        o = tmp_path / f"o{i}.kite"
        ol.write_table(o, t)
        ok_.append(o)
    # the merge of the GPU-written files, byte for byte against the oracle's merge
    kite("merge", "-i", *gk, "-o", tmp_path / "gm.kite")
    ol.kite_merge(gk, tmp_path / "om.kite", lanes=2)
    assert (tmp_path / "gm.kite").read_bytes() == (tmp_path / "om.kite").read_bytes()
    # the dump of that merge against the oracle's count -> merge -> dump chain
    ol.kite_merge(ok_, tmp_path / "oo.kite", lanes=2)
    ol.kite_dump(tmp_path / "oo.kite", tmp_path / "oo.tsv")
    text = _dump_bytes(tmp_path / "gm.kite")
    assert text == (tmp_path / "oo.tsv").read_bytes()
    rows = text.count(b"\n")
    t0 = ol.kite_load(tmp_path / "o0.kite")[0]
    t1 = ol.kite_load(tmp_path / "o1.kite")[0]
    assert max(t0.n_unique, t1.n_unique) <= rows <= t0.n_unique + t1.n_unique
    if genome_len == 1_000_000:
        # the genome's ~1 M canonical 63-mers are in both inputs; read errors
        # (0.8% a base: about 40% of windows) give each input ~2 M of its own
        assert t0.n_unique + t1.n_unique - rows > 500_000


def _filtered(keys, counts, off, min_count):
    m = counts >= min_count
    cs = np.concatenate([np.zeros(1, dtype=np.uint64), np.cumsum(m, dtype=np.uint64)])
    return keys[m], counts[m].astype(np.uint64), cs[off.astype(np.int64)]


def test_config5_cohort_analyze(gpu_ctx, tmp_path):
    """The 24-sample cohort at 2% of its read counts, --min 2."""
    from kite_amd import BATCH_PIPELINED
    from kite_amd._native import CORR_FP64_BASELINE
    from kite_amd.workloads import batch24
    ws = batch24(24, scale=0.02)
    fqs = [w.fastq() for w in ws]
    k, P = ws[0].k, ws[0].partitions
    res = gpu_ctx.count_batch(fqs, k, True, P, mode=BATCH_PIPELINED)
    paths, tabs = [], []
    for i, (w, fq) in enumerate(zip(ws, fqs)):
        t = ol.count_fastq(fq, k, True, P, threads=THREADS)
        r, st = res[i]
        assert np.array_equal(r.keys, t.keys) and np.array_equal(r.counts.astype(np.uint64), t.counts)
        p = tmp_path / f"sample{i:02d}.kite"
        ol.write_table(p, t, min_count=2)
        paths.append(p)
        tabs.append(_filtered(r.keys, r.counts.astype(np.uint64), r.part_off, 2))
    ukeys, umat = ol.kite_union(paths)
    assert umat.shape[0] > 10_000
    g = gpu_ctx.analyze_tables(tabs, 1, methods=3, flags=CORR_FP64_BASELINE, want_union=True)
    assert np.array_equal(g["keys"], ukeys)
    assert np.array_equal(g["matrix"].astype(np.uint64), umat)
    V = co.transform(umat)
    p_o, s_o, c_o = co.pearson(V), co.spearman(umat.astype(np.float64)), co.covariance(V)
    assert np.max(np.abs(g["pearson"] - p_o)) <= 1e-5
    assert np.max(np.abs(g["spearman"] - s_o)) <= 1e-5
    assert np.max(np.abs(g["pearson_fp64"] - p_o)) <= 1e-9
    # The reference's one-pass Sxx - Sx^2/N (pearson_slice, src/analyze/mod.rs:
    # 559-578), which the oracle keeps, cancels on rank columns that are
    # mostly tied zeros (variance tiny next to the squared mean): its
    # diagonal comes out near 1 - 5e-7 here.  The FP64 GPU path is held to
    # 1e-9 of a centered two-pass FP64 computation (np.corrcoef of the
    # average ranks) and, like the split-bf16 path, to 1e-5 of the oracle.
    ranks = np.stack([co.average_ranks(umat[:, j].astype(np.float64)) for j in range(umat.shape[1])], axis=1)
    s_c = np.corrcoef(ranks, rowvar=False)
    assert np.max(np.abs(g["spearman_fp64"] - s_c)) <= 1e-9
    assert np.max(np.abs(g["spearman_fp64"] - s_o)) <= 1e-5
    assert np.max(np.abs(g["cov"] - c_o)) <= 1e-5 * max(1.0, np.abs(c_o).max())
    # PCA (top 10 components): the device Jacobi on the oracle's covariance,
    # then end to end from the device covariance.  Eigenvectors of nearly
    # equal eigenvalues are not unique, so scores are compared (sign
    # normalised) only for components whose eigenvalue stands apart.
    S, w_o = co.pca_scores(c_o, 10)
    scale = max(1.0, float(np.abs(w_o).max()))
    sc, ev = gpu_ctx.pca(c_o, 10)
    assert np.max(np.abs(ev - w_o)) <= 1e-9 * scale
    sc2, ev2 = gpu_ctx.pca(g["cov"], 10)
    assert np.max(np.abs(ev2 - w_o)) <= 1e-5 * scale
    for d in range(10):
        gap = min([abs(w_o[d] - w_o[j]) for j in (d - 1, d + 1) if 0 <= j < len(w_o)])
        if gap > 1e-3 * scale:
            assert np.max(np.abs(co.sign_normalize(sc[:, d:d + 1]) - co.sign_normalize(S[:, d:d + 1]))) <= 1e-7 * scale
        if gap > 1e-2 * scale:
            assert np.max(np.abs(co.sign_normalize(sc2[:, d:d + 1]) - co.sign_normalize(S[:, d:d + 1]))) <= 1e-3 * scale
    # the drop-in CLI on the same files
    out = tmp_path / "an"
    kite("analyze", "-i", *paths, "-o", out, "--correlation", "all")

    def table(name):
        lines = (out / name).read_text().splitlines()
        assert lines[0].split("\t")[1:] == [f"sample{i:02d}" for i in range(24)]
        return np.array([[float(x) for x in ln.split("\t")[1:]] for ln in lines[1:]])

    assert np.max(np.abs(table("correlation_pearson.tsv") - p_o)) <= 1e-5 + 1e-8
    assert np.max(np.abs(table("correlation_spearman.tsv") - s_o)) <= 1e-5 + 1e-8
    rows = (out / "exact_count_matrix.tsv").read_text().splitlines()
    assert len(rows) == 1 + umat.shape[0]
    for r in range(0, umat.shape[0], max(1, umat.shape[0] // 300)):
        f = rows[1 + r].split("\t")
        assert f[0] == ol.decode_key(ol.words_to_int(ukeys[r]), 1, k)
        assert [int(x) for x in f[1:]] == umat[r].tolist()


def test_auto_partitions_count_dump(tmp_path):
    """`kite count` with no --partitions: a 60k-read k=31 file plans P=256
    (the auto minimum, src/count/mod.rs:40, 923-943), and the dump equals
    the oracle's at P=256."""
    from kite_amd.workloads import K31_CANON_50MBP
    w = K31_CANON_50MBP.scaled(60_000, genome_len=3_000_000)
    fq = w.fastq()
    p = tmp_path / "s.fastq"
    fq.tofile(p)
    r = kite("count", "-i", p, "-o", tmp_path / "g.kite", "-k", 31, "--canonical", "true", "--memory", "8G",
             "-t", 8, "--test-mode")
    assert "[auto] P=256," in r.stderr
    t = ol.count_fastq(fq, 31, True, 256, threads=THREADS)
    ol.write_table(tmp_path / "o.kite", t)
    ol.kite_dump(tmp_path / "o.kite", tmp_path / "o.tsv")
    assert _dump_bytes(tmp_path / "g.kite") == (tmp_path / "o.tsv").read_bytes()


def test_config3_full_size_properties(gpu_ctx):
    """k=63 canonical, all 10,000,000 reads of 250 bp (5.15 GB of FASTQ: an
    input above 4 GiB), P=1024: the counts add up to the exact window count
    of the FASTQ bytes, every partition's 128-bit keys are strictly
    ascending, and every key sits in its Fibonacci partition of lo ^ hi
    (src/count/kmer.rs:111-116)."""
    from kite_amd.workloads import K63_CANON_100MBP as w
    fq = w.fastq()
    assert fq.size > 1 << 32
    res, st = gpu_ctx.count(fq, w.k, w.canonical, w.partitions)
    assert st.kmers == exact_windows(fq, w.k)
    del fq
    assert int(res.counts.astype(np.uint64).sum()) == st.kmers
    lo, hi = res.keys[:, 0], res.keys[:, 1]
    part = ((lo ^ hi) * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(64 - 10)
    off = res.part_off.astype(np.int64)
    assert np.array_equal(np.searchsorted(part, np.arange(w.partitions + 1), side="left"), off)
    assert np.all(np.diff(part.astype(np.int64)) >= 0)
    same_part = part[1:] == part[:-1]
    up = (hi[1:] > hi[:-1]) | ((hi[1:] == hi[:-1]) & (lo[1:] > lo[:-1]))
    assert np.all(up[same_part])
