"""FP64 restatement of the correlation part of jiehua1995/Kite v0.7.5
`kite analyze` (src/analyze/mod.rs).

TEST INFRASTRUCTURE ONLY: the checker for the GPU correlation engine.
Nothing in the shipped engine imports or calls it.

* values: ln(1 + count) in FP64 (fill_transformed_values, :780-784)
* Pearson: one-pass sums like PearsonAccumulator (:194-236): r = (Sxy -
  SxSy/N) / sqrt(max(0, Sxx - Sx^2/N) * max(0, Syy - Sy^2/N)), 0 when the
  denominator is <= 0
* Spearman: 1-based average-tie ranks per column over all rows
  (average_ranks, :406-427), then pearson_slice (:559-578)
* covariance: (Sxy - SxSy/N) / (N - 1), zeros for N <= 1 (:238-252)
* PCA: symmetric eigendecomposition, eigenvalues descending, scores
  v * sqrt(max(lambda, 0)) (pca_scores, :1605-1625).  nalgebra's
  eigenvector signs are not reproducible here: compare sign-normalized.
"""
from __future__ import annotations

import numpy as np


def transform(counts: np.ndarray) -> np.ndarray:
    return np.log1p(counts.astype(np.float64))


def _corr_from_sums(n: int, s: np.ndarray, ss: np.ndarray, cross: np.ndarray) -> np.ndarray:
    num = cross - np.outer(s, s) / n
    d = ss - s * s / n
    den = np.sqrt(np.outer(np.maximum(d, 0.0), np.maximum(d, 0.0)))
    out = np.zeros_like(num)
    ok = den > 0
    out[ok] = num[ok] / den[ok]
    return out


def pearson(values: np.ndarray) -> np.ndarray:
    """PearsonAccumulator.correlation over rows of `values` (N x n)."""
    n = values.shape[0]
    if n == 0:
        return np.zeros((values.shape[1], values.shape[1]))
    s = values.sum(axis=0)
    ss = (values * values).sum(axis=0)
    cross = values.T @ values
    return _corr_from_sums(n, s, ss, cross)


def pearson_slice(x: np.ndarray, y: np.ndarray) -> float:
    if len(x) != len(y) or len(x) == 0:
        return 0.0
    return float(pearson(np.stack([x, y], axis=1))[0, 1])


def average_ranks(col: np.ndarray) -> np.ndarray:
    """1-based ranks, ties get the average of their positions."""
    order = np.argsort(col, kind="stable")
    sv = col[order]
    ranks = np.empty(len(col), dtype=np.float64)
    start = 0
    n = len(col)
    # boundaries of equal runs
    bounds = np.flatnonzero(np.diff(sv) != 0) + 1 if n else np.array([], dtype=np.int64)
    edges = np.concatenate([[0], bounds, [n]]) if n else np.array([0])
    for a, b in zip(edges[:-1], edges[1:]):
        ranks[order[a:b]] = (a + 1 + b) / 2.0
    return ranks


def spearman(values: np.ndarray) -> np.ndarray:
    if values.shape[0] == 0:
        return np.zeros((values.shape[1], values.shape[1]))
    r = np.stack([average_ranks(values[:, j]) for j in range(values.shape[1])], axis=1)
    return pearson(r)


def covariance(values: np.ndarray) -> np.ndarray:
    n = values.shape[0]
    k = values.shape[1]
    if n <= 1:
        return np.zeros((k, k))
    s = values.sum(axis=0)
    return (values.T @ values - np.outer(s, s) / n) / (n - 1)


def pca_scores(cov: np.ndarray, dims: int) -> tuple[np.ndarray, np.ndarray]:
    """(scores n x dims, eigenvalues descending)."""
    w, v = np.linalg.eigh(cov)
    order = np.argsort(-w, kind="stable")
    w, v = w[order], v[:, order]
    out = np.zeros((cov.shape[0], dims))
    m = min(dims, cov.shape[0])
    out[:, :m] = v[:, :m] * np.sqrt(np.maximum(w[:m], 0.0))
    return out, w


def sign_normalize(scores: np.ndarray) -> np.ndarray:
    """Flip each column so its largest-magnitude entry is positive."""
    out = scores.copy()
    for d in range(out.shape[1]):
        i = np.argmax(np.abs(out[:, d]))
        if out[i, d] < 0:
            out[:, d] = -out[:, d]
    return out
