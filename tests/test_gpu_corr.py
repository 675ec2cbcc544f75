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


def test_large_counts_use_wide_codes(gpu_ctx):
    rng = np.random.default_rng(9)
    M = synth_counts(rng, 20000, 6)
    M[::97, 2] = 1_000_000
    M[:, 4] = np.arange(20000, dtype=np.uint32) % 700  # 700 distinct counts
    out = gpu_ctx.corr_counts(M, 3, 1)
    assert out["stats"]["code_bytes"] == 4
    V = co.transform(M)
    assert np.max(np.abs(out["pearson"] - co.pearson(V))) <= TOL_BF16
    assert np.max(np.abs(out["spearman"] - co.spearman(V))) <= TOL_BF16


@pytest.mark.parametrize("n", [2, 6, 24, 40])
def test_pca_matches_eigh(gpu_ctx, n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n + 5, n))
    C = np.cov(A, rowvar=False)
    scores, ev = gpu_ctx.pca(C, min(n, 10))
    S, w = co.pca_scores(C, min(n, 10))
    assert np.max(np.abs(ev - w)) <= 1e-9 * max(1.0, np.abs(w).max())
    assert np.max(np.abs(co.sign_normalize(scores) - co.sign_normalize(S))) <= 1e-7


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
