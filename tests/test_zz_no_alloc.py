"""kite_gpu_pca_noalloc (fixed device working memory, no allocation) against
kite_gpu_pca through the C ABI: the same kernel on the same input, so the
results must be bit-identical, on consecutive calls of shrinking size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spd(rng, n):
    m = rng.standard_normal((n + 5, n))
    return m.T @ m / (n + 5)


def test_noalloc_pca_equals_pca_on_consecutive_calls(gpu_ctx):
    rng = np.random.default_rng(12)
    for n, dims in ((40, 10), (7, 3), (64, 64), (1, 2), (23, 2)):
        cov = _spd(rng, n)
        s_ref, e_ref = gpu_ctx.pca(cov, dims)
        s_na, e_na = gpu_ctx.pca_noalloc(cov, dims)
        assert np.array_equal(e_na, e_ref), n
        assert np.array_equal(s_na, s_ref), n
        # and both against numpy's eigenvalues (descending)
        w = np.sort(np.linalg.eigvalsh(cov))[::-1]
        assert np.max(np.abs(e_na - w)) <= 1e-9 * max(1.0, np.abs(w).max())


def test_noalloc_pca_rejects_sizes_beyond_its_memory(gpu_ctx):
    from kite_amd import KiteError
    with pytest.raises(KiteError) as ei:
        gpu_ctx.pca_noalloc(np.eye(65), 2)
    assert ei.value.code == -1 and "1..64" in str(ei.value)
    # the entry point still works after the rejected call
    s, e = gpu_ctx.pca_noalloc(np.diag([3.0, 1.0, 2.0]), 2)
    assert e.tolist() == [3.0, 2.0, 1.0]
