"""FP64 dense reference for the least-squares gradients (test helper).

x = argmin |A x - b| via Householder QR (numpy), and for an upstream g:
w = R^-1 R^-T g = (A^T A)^-1 g, grad_b = A w, grad_A = r w^T - (A w) x^T."""
import numpy as np


def dense_grads(A, b, g):
    A = np.asarray(A, np.float64)
    b = np.asarray(b, np.float64)
    g = np.asarray(g, np.float64)
    Q, R = np.linalg.qr(A)
    x = np.linalg.solve(R, Q.T @ b)
    w = np.linalg.solve(R, np.linalg.solve(R.T, g))
    gb = A @ w
    r = b - A @ x
    gA = np.outer(r, w) - np.outer(gb, x)
    return w, gb, gA, x, r
