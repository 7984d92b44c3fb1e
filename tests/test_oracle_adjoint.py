"""Pins the least-squares gradient used by the adjoint against central
finite differences of the FP64 least-squares map (numpy lstsq), so a
dropped term or sign in  w = (A^T A)^-1 g,  grad_b = A w,
grad_A = r w^T - (A w) x^T  fails here.  CPU only."""
import numpy as np

from tests.adjoint_ref import dense_grads


def loss(A, b, g):
    return g @ np.linalg.lstsq(A, b, rcond=None)[0]


def test_gradient_formula_vs_central_differences():
    rng = np.random.default_rng(11)
    m, n = 60, 7
    A = rng.standard_normal((m, n)) * np.logspace(0, -1, n)
    b = rng.standard_normal(m)  # inconsistent: r != 0, so the r w^T term matters
    g = rng.standard_normal(n)
    w, gb, gA, x, r = dense_grads(A, b, g)
    assert np.linalg.norm(r) > 0.1 * np.linalg.norm(b)
    h = 1e-3
    fd_b = np.array([(loss(A, b + h * e, g) - loss(A, b - h * e, g)) / (2 * h) for e in np.eye(m)])
    assert np.abs(fd_b - gb).max() < 1e-6 * np.abs(gb).max()
    for (i, j) in [(0, 0), (5, 3), (59, 6), (17, 2), (33, 5)]:
        E = np.zeros_like(A)
        E[i, j] = h
        fd = (loss(A + E, b, g) - loss(A - E, b, g)) / (2 * h)
        assert abs(fd - gA[i, j]) < 1e-5 * np.abs(gA).max(), (i, j, fd, gA[i, j])


def test_gradient_special_cases():
    rng = np.random.default_rng(12)
    A = rng.standard_normal((40, 5))
    g = rng.standard_normal(5)
    # square, invertible A: x = A^-1 b, grad_b = A^-T g, r = 0
    Q = rng.standard_normal((5, 5)) + 5 * np.eye(5)
    b5 = rng.standard_normal(5)
    w, gb, gA, x, r = dense_grads(Q, b5, g)
    assert np.allclose(gb, np.linalg.solve(Q.T, g))
    assert np.allclose(gA, -np.outer(gb, x))
    # consistent system (b in range(A)): r = 0 and grad_A = -(A w) x^T
    b = A @ rng.standard_normal(5)
    w, gb, gA, x, r = dense_grads(A, b, g)
    assert np.abs(r).max() < 1e-12 * np.abs(b).max()
    assert np.allclose(gA, -np.outer(gb, x), atol=1e-12)
    # A^T grad_b = g (grad_b is the min-norm solution of A^T z = g)
    assert np.allclose(A.T @ gb, g)
