"""Pins the oracle's CSR product (orc_csr_mv, restating
src/kernels/cpu/matrix/matrix_cpu_kernels.cpp:92-131) against
scipy.sparse and the dense product, including repeated entries (summed),
empty rows, unsorted column indices and the beta == 0 rule (y unread)."""
import numpy as np
import pytest
import scipy.sparse as sp

from tests.test_gpu_csr import random_csr


@pytest.mark.parametrize("m,n,mean", [(50, 7, 3), (7, 50, 12), (1, 1, 1), (30, 4, 0.2)])
@pytest.mark.parametrize("trans", [0, 1])
def test_matches_scipy_and_dense(orc, m, n, mean, trans):
    rng = np.random.default_rng(m * 100 + n)
    rp, ci, v = random_csr(rng, m, n, mean)
    S = sp.csr_matrix((v, ci, rp), shape=(m, n))
    D = np.zeros((m, n))
    for i in range(m):
        for k in range(rp[i], rp[i + 1]):
            D[i, ci[k]] += v[k]
    x = rng.standard_normal(m if trans else n)
    y0 = rng.standard_normal(n if trans else m)
    ref = (S.T @ x if trans else S @ x) * 1.5 - 0.25 * y0
    got = orc.csr_mv(trans, 1.5, m, n, v, rp, ci, x, beta=-0.25, y=y0)
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-13)
    assert np.allclose(got, (D.T if trans else D) @ x * 1.5 - 0.25 * y0, rtol=1e-13, atol=1e-13)
    nan = np.full_like(y0, np.nan)
    got0 = orc.csr_mv(trans, 2.0, m, n, v, rp, ci, x, beta=0.0, y=nan)
    assert np.all(np.isfinite(got0)) and np.allclose(got0, 2.0 * (S.T @ x if trans else S @ x), rtol=1e-13,
                                                       atol=1e-13)


def test_planted_transpose_mistake_is_caught(orc):
    """A^T x computed as A x on a square non-symmetric matrix differs."""
    rp, ci, v = np.array([0, 2, 3]), np.array([0, 1, 1]), np.array([1.0, 2.0, 3.0])
    x = np.array([1.0, 1.0])
    assert not np.allclose(orc.csr_mv(1, 1.0, 2, 2, v, rp, ci, x), orc.csr_mv(0, 1.0, 2, 2, v, rp, ci, x))
