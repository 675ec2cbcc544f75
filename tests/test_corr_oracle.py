"""The FP64 correlation oracle against the reference's analyze unit tests
(src/analyze/mod.rs:2858-2990), scipy, and closed forms."""
import sys
from pathlib import Path

import numpy as np
import pytest
import scipy.stats

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import corr_oracle as co  # noqa: E402


def test_average_ranks_handles_ties():  # analyze/mod.rs:2866-2876
    assert co.average_ranks(np.array([10.0, 20.0, 20.0, 30.0])).tolist() == [1.0, 2.5, 2.5, 4.0]
    assert co.average_ranks(np.array([3.0, 1.0, 3.0, 2.0])).tolist() == [3.5, 1.0, 3.5, 2.0]


def test_pearson_slice_expected():  # analyze/mod.rs:2878-2883
    assert abs(co.pearson_slice(np.array([1.0, 2, 3]), np.array([2.0, 4, 6])) - 1.0) < 1e-8
    assert abs(co.pearson_slice(np.array([1.0, 2, 3]), np.array([6.0, 4, 2])) + 1.0) < 1e-8
    assert co.pearson_slice(np.array([1.0, 1, 1]), np.array([2.0, 3, 4])) == 0.0


def test_accumulator_rows():  # analyze/mod.rs:2893-2911 (Pearson part)
    rows = np.array([[1.0, 2.0], [2.0, 4.0], [3.0, 6.0]])
    assert abs(co.pearson(rows)[0, 1] - 1.0) < 1e-8


def test_ranks_match_scipy_and_spearman_matches_scipy():
    rng = np.random.default_rng(1)
    X = rng.poisson(2.0, size=(500, 5)).astype(np.float64)
    for j in range(5):
        assert np.array_equal(co.average_ranks(X[:, j]), scipy.stats.rankdata(X[:, j], method="average"))
    rho = scipy.stats.spearmanr(X).statistic
    assert np.max(np.abs(co.spearman(X) - rho)) < 1e-12


def test_pearson_matches_corrcoef_and_covariance_matches_cov():
    rng = np.random.default_rng(2)
    V = co.transform(rng.poisson(5.0, size=(1000, 6)))
    assert np.max(np.abs(co.pearson(V) - np.corrcoef(V, rowvar=False))) < 1e-12
    assert np.max(np.abs(co.covariance(V) - np.cov(V, rowvar=False))) < 1e-12


def test_pca_reconstructs_covariance():
    rng = np.random.default_rng(3)
    V = rng.normal(size=(200, 6)) @ rng.normal(size=(6, 6))
    C = co.covariance(V)
    S, w = co.pca_scores(C, 6)
    assert np.all(np.diff(w) <= 1e-12)
    assert np.max(np.abs(S @ S.T - C)) < 1e-9  # v diag(w) v^T with w >= 0
