"""Split-bf16 MFMA correlation and FP64 Jacobi PCA against the FP64 oracle.

Tolerance: max |r_gpu - r_fp64_oracle| <= 1e-5 for the split-bf16 Gram
(the stated target); <= 1e-9 for the FP64 MFMA Gram baseline."""
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as ol

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import corr_oracle as co  # noqa: E402

pytestmark = pytest.mark.gpu
TOL_BF16 = 1e-5
TOL_FP64 = 1e-9


def synth_counts(rng, N, n, scale=30.0, zero_frac=0.3):
    base = rng.gamma(2.0, scale / 2.0, size=(N, 1))
    m = rng.poisson(base * rng.uniform(0.3, 1.7, size=(1, n)))
    m[rng.random((N, n)) < zero_frac] = 0
    return m.astype(np.uint32)


def check_corr(got, ref, const, tol):
    """|got - ref| <= tol outside the constant columns.  A constant column
    has zero variance: the exact (centred) result is 0 in its row and column.
    The reference's one-pass formula Sxx - Sx^2/N leaves rounding noise of
    either sign there (so r is noise, or 1 on the diagonal), and that noise
    depends on the summation order: it is not a parity target."""
    keep = ~const
    if keep.any():
        assert np.max(np.abs(got[np.ix_(keep, keep)] - ref[np.ix_(keep, keep)])) <= tol
    assert np.all(got[const, :] == 0) and np.all(got[:, const] == 0)


@pytest.mark.parametrize("N,n", [(1, 3), (17, 5), (5000, 24), (200000, 24), (100003, 32)])
def test_corr_counts_matches_fp64_oracle(gpu_ctx, N, n):
    rng = np.random.default_rng(N + n)
    M = synth_counts(rng, N, n)
    M[:, 1 % n] = 7 if n > 2 else M[:, 1 % n]  # a constant column -> zero correlations
    out = gpu_ctx.corr_counts(M, 3, 1)
    V = co.transform(M)
    P, S, Cv = co.pearson(V), co.spearman(V), co.covariance(V)
    const = np.all(M == M[:1], axis=0)
    check_corr(out["pearson"], P, const, TOL_BF16)
    check_corr(out["spearman"], S, const, TOL_BF16)
    check_corr(out["pearson_fp64"], P, const, TOL_FP64)
    check_corr(out["spearman_fp64"], S, const, TOL_FP64)
    assert np.max(np.abs(out["cov"] - Cv)) <= 1e-5 * max(1.0, np.abs(Cv).max())
    if N > 1:
        assert np.all(np.diag(out["pearson"])[~const] == 1.0)


def test_large_counts_use_rank_codes(gpu_ctx):
    """Codes are ranks among a column's distinct counts: their width follows
    the number of distinct counts (700 here: two bytes), not their size."""
    rng = np.random.default_rng(9)
    M = synth_counts(rng, 20000, 6)
    M[::97, 2] = 1_000_000
    M[:, 4] = np.arange(20000, dtype=np.uint32) % 700  # 700 distinct counts
    out = gpu_ctx.corr_counts(M, 3, 1)
    assert out["stats"]["code_bytes"] == 2
    V = co.transform(M)
    assert np.max(np.abs(out["pearson"] - co.pearson(V))) <= TOL_BF16
    assert np.max(np.abs(out["spearman"] - co.spearman(V))) <= TOL_BF16


def test_counts_far_above_2_26(gpu_ctx):
    """Counts up to u32::MAX (the engine once refused counts above 2^26), a
    column with 70,000 distinct counts (four-byte codes) and columns whose
    large counts tie."""
    rng = np.random.default_rng(26)
    N, n = 150_000, 9
    M = synth_counts(rng, N, n)
    M[:, 0] = rng.permutation(np.arange(N, dtype=np.uint64) % 70_000 * 61_357 + 3).astype(np.uint32)
    M[::5, 1] = np.uint32(0xFFFFFFFF)
    M[::7, 1] = np.uint32(1 << 30)
    M[:, 2] = rng.integers(1 << 26, 1 << 32, size=N, dtype=np.uint64).astype(np.uint32)
    out = gpu_ctx.corr_counts(M, 3, 1)
    assert out["stats"]["code_bytes"] == 4
    V = co.transform(M)
    P, S = co.pearson(V), co.spearman(V)
    const = np.all(M == M[:1], axis=0)
    check_corr(out["pearson"], P, const, TOL_BF16)
    check_corr(out["spearman"], S, const, TOL_BF16)
    check_corr(out["pearson_fp64"], P, const, TOL_FP64)
    check_corr(out["spearman_fp64"], S, const, TOL_FP64)


@pytest.mark.parametrize("n", [33, 40, 100])
def test_more_than_32_samples(gpu_ctx, n):
    """More samples than one 32-column tile: the Gram of every pair of
    16-column blocks, assembled into the n x n matrices."""
    rng = np.random.default_rng(1000 + n)
    M = synth_counts(rng, 30_000, n)
    M[:, 5] = 4  # a constant column
    out = gpu_ctx.corr_counts(M, 3, 1)
    assert out["stats"]["passes"] == (((n + 15) // 16) * ((n + 15) // 16 - 1)) // 2
    V = co.transform(M)
    P, S, Cv = co.pearson(V), co.spearman(V), co.covariance(V)
    const = np.all(M == M[:1], axis=0)
    check_corr(out["pearson"], P, const, TOL_BF16)
    check_corr(out["spearman"], S, const, TOL_BF16)
    check_corr(out["pearson_fp64"], P, const, TOL_FP64)
    check_corr(out["spearman_fp64"], S, const, TOL_FP64)
    assert np.max(np.abs(out["cov"] - Cv)) <= 1e-5 * max(1.0, np.abs(Cv).max())


def test_analyze_tables_counts_above_u32(gpu_ctx):
    """Input counts above u32::MAX (u64 sums of several runs): the union the
    correlation reads keeps the exact u64 counts."""
    rng = np.random.default_rng(64)
    P, n = 4, 5
    pool = rng.choice(1 << 40, size=3000, replace=False).astype(np.uint64)
    tabs, dicts = [], []
    for i in range(n):
        ks = np.sort(rng.choice(pool, size=1500, replace=False))
        cs = rng.integers(1, 50, size=ks.size).astype(np.uint64)
        cs[::11] = np.uint64(1 << 33) + np.uint64(i)
        cs[::13] = np.uint64(0xFFFFFFFF) + np.uint64(7)
        part = ((ks * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(62)).astype(np.int64)
        order = np.lexsort((ks, part))
        ks, cs, part = ks[order], cs[order], part[order]
        off = np.searchsorted(part, np.arange(P + 1)).astype(np.uint64)
        tabs.append((ks.reshape(-1, 1), cs, off))
        dicts.append(dict(zip(ks.tolist(), cs.tolist())))
    out = gpu_ctx.analyze_tables(tabs, 1, 3, 1, want_union=True)
    keys = out["keys"][:, 0]
    exact = np.array([[d.get(int(k), 0) for d in dicts] for k in keys], dtype=np.uint64)
    assert np.array_equal(out["matrix"].astype(np.uint64), np.minimum(exact, np.uint64(0xFFFFFFFF)))
    V = co.transform(exact)
    assert np.max(np.abs(out["pearson"] - co.pearson(V))) <= TOL_BF16
    assert np.max(np.abs(out["spearman"] - co.spearman(V))) <= TOL_BF16
    assert np.max(np.abs(out["pearson_fp64"] - co.pearson(V))) <= TOL_FP64
    assert np.max(np.abs(out["spearman_fp64"] - co.spearman(V))) <= TOL_FP64


@pytest.mark.parametrize("n", [2, 6, 24, 40, 64, 65, 100, 300])
def test_pca_matches_eigh(gpu_ctx, n):
    """The one-wave Jacobi (n <= 64) and the round-robin parallel Jacobi
    (n > 64) against numpy's eigh: eigenvalues, and sign-normalised scores
    of the leading components (well separated for these matrices)."""
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n + 5, n)) * np.linspace(3.0, 0.5, n)
    C = np.cov(A, rowvar=False)
    scores, ev = gpu_ctx.pca(C, min(n, 10))
    S, w = co.pca_scores(C, min(n, 10))
    assert np.max(np.abs(ev - w)) <= 1e-9 * max(1.0, np.abs(w).max())
    assert np.max(np.abs(co.sign_normalize(scores) - co.sign_normalize(S))) <= 1e-7 * max(1.0, np.abs(S).max())


def test_analyze_tables_union_and_correlation(gpu_ctx, tmp_path):
    from kite_amd.workloads import K31_CANON_50MBP
    paths, tabs = [], []
    for i in range(5):
        w = K31_CANON_50MBP.scaled(3000 + 1500 * i, genome_len=300_000)
        fq = w.fastq()
        t = ol.count_fastq(fq, 21, True, 16)
        p = tmp_path / f"s{i}.kite"
        ol.write_table(p, t)
        paths.append(p)
        tabs.append((t.keys, t.counts, t.part_off))
    out = gpu_ctx.analyze_tables(tabs, 1, 3, 1, want_union=True)
    keys, matrix = ol.kite_union(paths)
    assert np.array_equal(out["keys"], keys)
    assert np.array_equal(out["matrix"].astype(np.uint64), matrix)
    V = co.transform(matrix)
    assert np.max(np.abs(out["pearson"] - co.pearson(V))) <= TOL_BF16
    assert np.max(np.abs(out["spearman"] - co.spearman(V))) <= TOL_BF16


@pytest.mark.parametrize("distinct,expect", [(256, 1), (257, 2)])
def test_dense_eight_bit_codes(gpu_ctx, distinct, expect):
    """Columns with at most 256 distinct counts are read as exact 8-bit
    ranks of their values; one more distinct count falls back to u16."""
    rng = np.random.default_rng(distinct)
    M = synth_counts(rng, 50000, 7)
    M[:, 3] = (np.arange(50000, dtype=np.uint32) * 7919) % distinct * 3  # exactly `distinct` values, up to 768
    out = gpu_ctx.corr_counts(M, 3, 1)
    assert out["stats"]["code_bytes"] == expect
    V = co.transform(M)
    P, S = co.pearson(V), co.spearman(V)
    const = np.all(M == M[:1], axis=0)
    check_corr(out["pearson"], P, const, TOL_BF16)
    check_corr(out["spearman"], S, const, TOL_BF16)
    check_corr(out["pearson_fp64"], P, const, TOL_FP64)
    check_corr(out["spearman_fp64"], S, const, TOL_FP64)


def test_cli_analyze_40_inputs_and_u64_cells(tmp_path):
    """`kite analyze` on 40 inputs; one input holds a key in two runs whose
    sum exceeds u32::MAX: exact_count_matrix.tsv and the sample totals print
    the u64 sum, and the correlation uses it."""
    import subprocess

    import kite_file as kf
    root = Path(__file__).resolve().parents[1]
    rng = np.random.default_rng(40)
    paths = []
    pool = sorted(set(int(x) for x in rng.integers(0, 1 << 40, size=4000)))
    for i in range(40):
        keys = sorted(rng.choice(pool, size=600, replace=False).tolist())
        runs = {}
        for key in keys:
            p = ol.bucket(key, 1, 2)
            runs.setdefault(p, []).append((key, int(rng.integers(1, 40))))
        rl = [(p, rows) for p, rows in sorted(runs.items())]
        if i == 7:  # a second run of the first partition repeating its first key near u32::MAX
            p0, k0 = rl[0][0], rl[0][1][0][0]
            rl[0] = (p0, [(k0, 0xFFFFFFF0)] + rl[0][1][1:])
            rl.insert(1, (p0, [(k0, 0xFFFFFFF0)]))
        f = tmp_path / f"s{i:02d}.kite"
        kf.write(f, 21, True, 4, 64, rl)
        paths.append(f)
    out = tmp_path / "an"
    r = subprocess.run([str(root / "kite_amd" / "lib" / "kite"), "analyze", "-i", *map(str, paths), "-o", str(out),
                        "--correlation", "all"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    keys, matrix = ol.kite_union(paths)
    assert int(matrix.max()) == 2 * 0xFFFFFFF0
    rows = (out / "exact_count_matrix.tsv").read_text().splitlines()
    assert len(rows) == 1 + len(keys)
    for ri in range(len(keys)):
        f = rows[1 + ri].split("\t")
        assert f[0] == ol.decode_key(int(keys[ri, 0]), 1, 21)
        assert [int(x) for x in f[1:]] == matrix[ri].tolist()
    summ = (out / "sample_summary.tsv").read_text().splitlines()
    assert int(summ[1 + 7].split("\t")[1]) == int(matrix[:, 7].sum())
    V = co.transform(matrix)

    def table(name):
        lines = (out / name).read_text().splitlines()
        return np.array([[float(x) for x in ln.split("\t")[1:]] for ln in lines[1:]])

    assert np.max(np.abs(table("correlation_pearson.tsv") - co.pearson(V))) <= 1e-5 + 1e-8
    assert np.max(np.abs(table("correlation_spearman.tsv") - co.spearman(V))) <= 1e-5 + 1e-8
    emb = (out / "embedding_pca.tsv").read_text().splitlines()
    assert len(emb) == 41
