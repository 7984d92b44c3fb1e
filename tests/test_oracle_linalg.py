"""Pins the oracle's GEMV tiers, CountSketch, QR/triangular inverse and LSQR
(oracle/tomos_oracle.c) to library routines, closed forms and the
reference's known-answer systems."""
import json
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "lsqr_known_answers.json")))
M64 = (1 << 64) - 1


def splitmix64_py(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def test_splitmix64_and_indices(orc):
    L = orc.lib()
    rng = np.random.default_rng(0)
    for x in list(rng.integers(0, 1 << 63, 200, dtype=np.int64)) + [0, 1, M64]:
        assert L.orc_splitmix64(int(x)) == splitmix64_py(int(x))
    # the first outputs of splitmix64 seeded with 0 (published test vector of
    # the generator: state += golden gamma, then mix)
    assert splitmix64_py(0) == 0xE220A8397B1DCDAF
    m, s = 5000, 37
    b, g = orc.countsketch_indices(m, s)
    for i in range(0, m, 97):
        assert b[i] == splitmix64_py(i ^ 0x12345678ABCDEF00) % s
        assert g[i] == (1 if splitmix64_py(i ^ 0xF00DD00D42BABE55) & 1 else -1)


def test_countsketch_apply_is_the_sparse_product(orc):
    rng = np.random.default_rng(42)
    m, n, s = 3000, 7, 50
    X = rng.standard_normal((m, n))
    b, g = orc.countsketch_indices(m, s)
    S = np.zeros((s, m))
    S[b, np.arange(m)] = g
    assert np.allclose(orc.countsketch_apply(X, s), S @ X, rtol=1e-13, atol=1e-12)
    # E|Sx|^2/|x|^2 ~ 1 over 200 trials (test_countsketch_cpu_kernels.cpp:93-179)
    T = rng.standard_normal((m, 200))
    Y = orc.countsketch_apply(T, s)
    ratio = (np.linalg.norm(Y, axis=0) ** 2 / np.linalg.norm(T, axis=0) ** 2).mean()
    assert abs(ratio - 1) < 0.05


@pytest.mark.parametrize("dec,variant", [(0, 0), (1, 0), (3, 0), (4, 3), (5, 3), (6, 4)])
@pytest.mark.parametrize("trans", [0, 1])
def test_gemv_tiers_vs_numpy(orc, dec, variant, trans):
    rng = np.random.default_rng(dec * 2 + trans)
    m, n = 300, 41
    A = (rng.standard_normal((m, n)) * 2.0 ** rng.integers(-8, 9, (m, n))).astype(np.float32)
    wh, wr = orc.pack(variant, A)
    Ad = orc.unpack(dec, wh, wr).astype(np.float64)
    x = rng.standard_normal(n if not trans else m).astype(np.float32)
    y0 = rng.standard_normal(m if not trans else n).astype(np.float32)
    got = orc.mv_packed(dec, trans, -0.5, wh, wr, m, x, 0.25, y0)
    ref = -0.5 * ((Ad.T if trans else Ad) @ x.astype(np.float64)) + 0.25 * y0
    assert np.allclose(got, ref.astype(np.float32), rtol=1e-6, atol=1e-6 * np.abs(ref).max())


@pytest.mark.parametrize("k,n", [(400, 100), (30, 5), (64, 64)])
def test_householder_qr_vs_numpy(orc, k, n):
    rng = np.random.default_rng(k)
    A = rng.standard_normal((k, n)) * np.logspace(0, -5, n)
    R = orc.householder_r(A)
    Rn = np.linalg.qr(A, mode="r")
    sg = np.sign(np.diag(R)) * np.sign(np.diag(Rn))
    assert np.allclose(R, sg[:, None] * Rn, rtol=1e-10, atol=1e-13 * np.abs(Rn).max())
    Ri = orc.trtri_upper(R)
    assert np.allclose(Ri, np.linalg.inv(R), rtol=1e-9, atol=1e-12 * np.abs(Ri).max())
    with pytest.raises(ValueError):
        orc.householder_r(np.ones((3, 5)))
    with pytest.raises(ValueError):
        orc.trtri_upper(np.zeros((3, 3)))


@pytest.mark.parametrize("sysd", GOLD["systems"], ids=[s["name"] for s in GOLD["systems"]])
def test_lsqr_known_answers(orc, sysd):
    m, n = sysd["m"], sysd["n"]
    A = np.array(sysd["a_rows"]).reshape(m, n)
    b = np.array(sysd["b"])
    x, it, hist = orc.lsqr_f64(A, b, sysd["tol"], sysd["max_iters"], sysd["tol"])
    assert np.allclose(x, sysd["x_expect"], atol=sysd["x_tol"], rtol=0)


def test_lsqr_sketch_precond_known_answer(orc):
    sp = GOLD["sketch_precond"]
    m, n, k = sp["m"], sp["n"], sp["k"]
    A = np.array(sp["a_colmajor"]).reshape(n, m).T
    xt = np.array(sp["x_true"])
    Y = orc.countsketch_apply(A, k)
    Rinv = orc.trtri_upper(orc.householder_r(Y))
    x, it, _ = orc.lsqr_f64(A, A @ xt, sp["tol"], sp["max_iters"], sp["tol"], Rinv=Rinv)
    assert np.allclose(x, xt, atol=sp["x_tol"], rtol=0)
    assert np.linalg.norm(A @ x - A @ xt) / np.linalg.norm(A @ xt) < sp["relres_max"]


def test_lsqr_vs_lstsq_and_scipy(orc):
    """The oracle's iterates equal scipy's textbook LSQR at the same step
    count (before rounding differences grow), its |phi_bar| history is
    monotone, and on a kappa=10 problem it stops at 1e-6 next to the FP64
    lstsq solution.  (On the C1 shape, kappa=1e3 with a log-spaced
    spectrum, 400 plain iterations do not reach 1e-6 -- neither here nor
    in scipy.)"""
    from scipy.sparse.linalg import lsqr as sp_lsqr
    rng = np.random.default_rng(0)
    m, n = 5000, 80
    A = rng.standard_normal((m, n)) * np.logspace(0, -3, n)
    b = A @ rng.standard_normal(n) + 1e-6 * rng.standard_normal(m)
    for k in (1, 3, 8):
        x, it, hist = orc.lsqr_f64(A, b, 1e-30, k, 1e-30)
        xs = sp_lsqr(A, b, atol=0, btol=0, conlim=0, iter_lim=k)[0]
        assert it == k and np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-12
    x, it, hist = orc.lsqr_f64(A, b, 1e-6, 400, 1e-6)
    assert it == 400 and np.all(np.diff(hist) <= 0)
    A2 = rng.standard_normal((m, n)) * np.logspace(0, -1, n)
    b2 = A2 @ rng.standard_normal(n) + 1e-7 * rng.standard_normal(m)
    x2, it2, hist2 = orc.lsqr_f64(A2, b2, 1e-6, 400, 1e-6)
    xls = np.linalg.lstsq(A2, b2, rcond=None)[0]
    assert it2 < 100 and hist2[-1] < 1e-6
    assert np.linalg.norm(b2 - A2 @ x2) / np.linalg.norm(b2) < 1.01e-6
    assert np.linalg.norm(x2 - xls) / np.linalg.norm(xls) < 1e-5


def test_lsqr_restart_history_contract(orc):
    rng = np.random.default_rng(2)
    A = rng.standard_normal((200, 10))
    b = rng.standard_normal(200)
    x, it, hist = orc.lsqr_f64(A, b, 1e-30, 3, 1e-30, restarts=4)
    # 4 restarts x 3 iterations, plus one check entry at each of the 3 later
    # restarts (check_stopping_criteria runs at the top of every restart)
    assert it == 12 and len(hist) == 15
    assert hist[3] == hist[2] and hist[7] == hist[6] and hist[11] == hist[10]
    xls = np.linalg.lstsq(A, b, rcond=None)[0]
    assert np.linalg.norm(x - xls) / np.linalg.norm(xls) < 1e-6


def test_lsqr_f32_with_openblas(orc):
    rc = orc.blas_load(2)
    rng = np.random.default_rng(4)
    A = (rng.standard_normal((3000, 50)) * np.logspace(0, -2, 50)).astype(np.float32)
    xt = rng.standard_normal(50)
    b = (A.astype(np.float64) @ xt).astype(np.float32)
    x, it, _ = orc.lsqr_f32(A, b, 1e-5, 400, 1e-5)
    assert rc in (0, -1)
    assert np.linalg.norm(A.astype(np.float64) @ x - b) / np.linalg.norm(b) < 2e-5


def test_lsqr_f32_preconditioned_with_openblas(orc):
    """Right-preconditioned FP32 LSQR (the CPU baseline's solver): with Rinv
    from the QR of a CountSketch of A it converges in a few iterations at
    kappa 1e4, where the plain FP32 run needs far more, and its solution
    matches numpy's FP64 least squares and the FP64 oracle's."""
    assert orc.blas_load(2) == 0 and orc.lib().orc_blas_loaded() == 1  # sgemv + BLAS-1 bound
    rng = np.random.default_rng(9)
    m, n = 4000, 40
    A = (rng.standard_normal((m, n)) * np.logspace(0, -4, n)).astype(np.float32, order="F")
    xt = rng.standard_normal(n)
    b = A.astype(np.float64) @ xt
    sk = orc.countsketch_apply_f32(A, 4 * n)
    R = np.linalg.qr(sk, mode="r")
    Rinv = np.linalg.inv(R)
    x, it, hist = orc.lsqr_f32(A, b.astype(np.float32), 1e-6, 8, 1e-6, restarts=20, Rinv=Rinv)
    rr = np.linalg.norm(A.astype(np.float64) @ x - b) / np.linalg.norm(b)
    assert rr < 2e-6, rr
    assert it < 40, it
    x64, it64, _ = orc.lsqr_f64(A.astype(np.float64), b, 1e-13, 8, 1e-13, restarts=40, Rinv=Rinv)
    xls = np.linalg.lstsq(A.astype(np.float64), b, rcond=None)[0]
    # kappa 1e4 amplifies FP32 rounding in the solution, not in the residual
    assert np.linalg.norm(x - xls) / np.linalg.norm(xls) < 1e-2
    assert np.linalg.norm(x64 - xls) / np.linalg.norm(xls) < 1e-8
    _, it_plain, _ = orc.lsqr_f32(A, b.astype(np.float32), 1e-6, 8, 1e-6, restarts=20)
    assert it_plain > 2 * it


def test_countsketch_f32_omp_matches_fp64(orc):
    """The OpenMP FP32-input CountSketch sums each bucket in the same row
    order as the FP64 oracle, so the two agree bit for bit; both match the
    explicit sketch matrix product."""
    rng = np.random.default_rng(10)
    m, n, s = 5003, 37, 101
    A = rng.standard_normal((m, n)).astype(np.float32, order="F")
    Y = orc.countsketch_apply_f32(A, s)
    assert np.array_equal(Y, orc.countsketch_apply(A.astype(np.float64), s))
    bk, sg = orc.countsketch_indices(m, s)
    S = np.zeros((s, m))
    S[bk, np.arange(m)] = sg
    assert np.allclose(Y, S @ A.astype(np.float64), rtol=1e-12, atol=1e-12)


def test_stream_copy_probe(orc):
    assert orc.stream_copy_gbs(1 << 24, 2) > 0.1
