"""GPU CountSketch (bit-exact indices), Householder QR and triangular inverse."""
import numpy as np
import pytest

from tests import hipcall as H
from tomos_amd import _capi as C
from tomos_amd._capi import DeviceBuffer

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,s", [(1000000, 4000), (40, 30), (12345, 7)])
def test_indices_bit_exact(capi, orc, m, s):
    b, g = H.countsketch_indices(m, s)
    eb, eg = orc.countsketch_indices(m, s)
    assert np.array_equal(b, eb) and np.array_equal(g, eg)


@pytest.mark.parametrize("a_type,tier", [(C.F64, 0), (C.F32, 0), (C.FP32_BF16, 2), (C.FP32_BF16, 0),
                                         (C.FP32_16, 2), (C.FP32_16, 0)])
@pytest.mark.parametrize("m,lda,n,s", [
    (20011, 20011, 37, 300),     # odd lda: element-wise kernel
    (40000, 40000, 37, 300),     # 16-byte row vectors, full column groups + a partial one
    (40003, 40008, 13, 4000),    # ragged row tail inside the last split, padded lda
    (100000, 100000, 9, 20000),  # large s: one column per group
])
def test_apply_vs_oracle(capi, orc, a_type, tier, m, lda, n, s):
    rng = np.random.default_rng(9)
    A = rng.standard_normal((m, n)).astype(np.float32, order="F")
    if a_type in (C.FP32_BF16, C.FP32_16):
        variant = C.PACK_FP32_BF16 if a_type == C.FP32_BF16 else C.PACK_FP32_16
        wh, wr = H.pack(variant, A, ld=lda)
        dec = {(C.FP32_BF16, 2): 4, (C.FP32_BF16, 0): 5, (C.FP32_16, 2): 0, (C.FP32_16, 0): 1}[(a_type, tier)]
        eh, er = orc.pack(0 if a_type == C.FP32_16 else 3, A)
        Aref = orc.unpack(dec, eh, er).astype(np.float64)
        Y = H.countsketch_apply(a_type, tier, m, n, wh, wr, lda, s)
    else:
        dt = np.float64 if a_type == C.F64 else np.float32
        Ap = np.zeros((lda, n), dt, order="F")
        Ap[:m] = A
        Aref = A.astype(np.float64)
        Y = H.countsketch_apply(a_type, tier, m, n, H.dev(Ap), None, lda, s)
    ref = orc.countsketch_apply(Aref, s)
    err = np.abs(Y - ref).max() / np.abs(ref).max()
    assert err < 2e-6, err


def test_norm_preservation(capi):
    """E |Sx|^2 / |x|^2 ~ 1 (test_countsketch_cpu_kernels.cpp:93-179, 200 trials)."""
    m, s, trials = 5000, 500, 200
    rng = np.random.default_rng(42)
    X = rng.standard_normal((m, trials))
    Y = H.countsketch_apply(C.F64, 0, m, trials, H.dev(X), None, m, s)
    ratio = (np.linalg.norm(Y, axis=0) ** 2 / np.linalg.norm(X, axis=0) ** 2).mean()
    assert abs(ratio - 1.0) < 0.05, ratio
    Y2 = H.countsketch_apply(C.F64, 0, m, trials, H.dev(2 * X + 1), None, m, s)
    Y1 = H.countsketch_apply(C.F64, 0, m, trials, H.dev(np.ones_like(X)), None, m, s)
    # values are FP32 sums within a row split, FP64 across splits
    assert np.allclose(Y2, 2 * Y + Y1, rtol=0, atol=2e-5 * np.abs(Y2).max())


@pytest.mark.parametrize("k,n", [(4000, 1000), (30, 5), (257, 256), (64, 64), (500, 77), (33, 33), (1000, 320)])
def test_qr_r_matches_numpy(capi, k, n):
    rng = np.random.default_rng(k + n)
    A = rng.standard_normal((k, n)) * np.logspace(0, -6, n)
    R = H.geqrf_r(A)
    Rn = np.linalg.qr(A, mode="r")
    sgn = np.sign(np.diag(R)) * np.sign(np.diag(Rn))
    assert np.allclose(np.tril(R, -1), 0)
    err = np.abs(R - sgn[:, None] * Rn).max() / np.abs(Rn).max()
    assert err < 1e-12, err


def test_qr_zero_column(capi):
    """A zero column gives an inactive reflector (tau = 0, R(c,c) = 0), as dlarfg."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((100, 40))
    A[:, 5] = 0.0
    A[:, 33] = 0.0
    R = H.geqrf_r(A)
    Rn = np.linalg.qr(A, mode="r")
    sgn = np.sign(np.diag(R)) * np.sign(np.diag(Rn))
    sgn[sgn == 0] = 1.0
    assert R[5, 5] == 0.0 and R[33, 33] == 0.0
    assert np.abs(R - sgn[:, None] * Rn).max() < 1e-12 * np.abs(Rn).max()


def test_trtri_and_precond(capi):
    rng = np.random.default_rng(3)
    n = 1000
    R = np.triu(rng.standard_normal((n, n))) + np.diag(np.full(n, 30.0))
    Ri = H.trtri(R)
    assert np.allclose(np.tril(Ri, -1), 0)
    assert np.abs(R @ Ri - np.eye(n)).max() < 1e-10
    A = rng.standard_normal((2000, 300))
    rinv, _ = H.sketch_rinv(C.F64, 0, 2000, 300, H.dev(A), None, 2000, 900)
    Rs = np.linalg.inv(rinv)
    # R^-1 preconditions A: the condition number of A R^-1 is small
    s = np.linalg.svd(A @ rinv, compute_uv=False)
    assert s[0] / s[-1] < 10, s[0] / s[-1]
    assert np.allclose(np.tril(Rs, -1), 0, atol=1e-9 * np.abs(Rs).max())


def test_errors(capi):
    import ctypes
    A = H.dev(np.ones((3, 5)))
    rc = C.lib().tomos_sketch_qr_precond_f64(3, 5, A.ptr, 3, None)
    assert rc == 1  # k < n (sketch_qr.cpp:41-44)
    Z = H.dev(np.zeros((4, 4)))
    rc = C.lib().tomos_trtri_upper_f64(4, Z.ptr, 4, None)
    assert rc == 6  # singular R (sketch_qr.cpp:59-63)


def test_batched_precond_matches_single(capi):
    """Uneven sketches factored together give bit-identical R^-1 to one-at-a-time calls."""
    import ctypes
    rng = np.random.default_rng(17)
    shapes = [(256, 64), (900, 200), (40, 33), (2048, 512), (130, 1)]
    mats = [rng.standard_normal((k, n)) * np.logspace(0, -4, n) for k, n in shapes]
    bufs = [H.dev(np.asfortranarray(A)) for A in mats]
    probs = (C.QrProblem * len(shapes))(*[C.QrProblem(k, n, k, b.ptr) for (k, n), b in zip(shapes, bufs)])
    sb = C.lib().tomos_sketch_qr_precond_batched_scratch_size(len(shapes), probs)
    scr = DeviceBuffer(sb)
    C.call("tomos_sketch_qr_precond_batched_f64", len(shapes), probs, scr.ptr, sb, None)
    for (k, n), A, b in zip(shapes, mats, bufs):
        single = H.dev(np.asfortranarray(A))
        C.call("tomos_sketch_qr_precond_f64", k, n, single.ptr, k, None)
        got = b.to_numpy(np.float64, (k, n))[:n, :n]
        ref = single.to_numpy(np.float64, (k, n))[:n, :n]
        assert np.array_equal(got, ref), (k, n)
        Rn = np.linalg.qr(A, mode="r")
        prod = np.abs(Rn) @ np.abs(got)
        assert np.abs(np.abs(Rn @ got) - np.eye(n)).max() < 1e-9 * max(1.0, prod.max())


def test_apply_deterministic_and_zero_column(capi, orc):
    """The fixed-point accumulation is exact: repeated calls agree bit for bit."""
    m, n, s = 100000, 20, 300
    rng = np.random.default_rng(21)
    A = (rng.standard_normal((m, n)) * np.logspace(0, -8, n)).astype(np.float32, order="F")
    A[:, 7] = 0.0
    wh, wr = H.pack(C.PACK_FP32_BF16, A)
    Y1 = H.countsketch_apply(C.FP32_BF16, C.TIER_EXACT, m, n, wh, wr, m, s)
    Y2 = H.countsketch_apply(C.FP32_BF16, C.TIER_EXACT, m, n, wh, wr, m, s)
    assert np.array_equal(Y1, Y2)
    assert not Y1[:, 7].any()
    ref = orc.countsketch_apply(A.astype(np.float64), s)
    # each column to its own scale (d spans 1..1e-8)
    for j in range(n):
        if j == 7:
            continue
        assert np.abs(Y1[:, j] - ref[:, j]).max() <= 1e-9 * np.abs(ref[:, j]).max()


def test_apply_saturated_fp32_16_falls_back(capi, orc):
    """A saturated FP32_16 word decodes to Inf: the float path handles that column set."""
    m, n, s = 20000, 5, 100
    rng = np.random.default_rng(22)
    A = rng.standard_normal((m, n)).astype(np.float32, order="F")
    A[123, 2] = 3.0e38  # beyond FP32_16's range: stored as +Inf
    wh, wr = H.pack(C.PACK_FP32_16, A)
    Y = H.countsketch_apply(C.FP32_16, C.TIER_EXACT, m, n, wh, wr, m, s)
    eh, er = orc.pack(C.PACK_FP32_16, A)
    ref = orc.countsketch_apply(orc.unpack(C.DECODE_FP32_16_EXACT, eh, er).astype(np.float64), s)
    assert np.array_equal(np.isinf(Y), np.isinf(ref)) and np.isinf(Y).sum() == 1
    fin = np.isfinite(ref)
    assert np.abs(Y[fin] - ref[fin]).max() <= 2e-6 * np.abs(ref[fin]).max()


@pytest.mark.parametrize("k,n", [(12000, 300), (9000, 70)])
def test_qr_tall_wavefront_path(capi, k, n):
    """k above 8192 takes the panel wavefront factorization (a block per
    panel column, per-column flags between blocks) instead of TSQR: R
    against numpy's QR up to row signs, and R^-1 from the sketch-QR entry
    point inverts it."""
    print("wavefront qr k=%d n=%d" % (k, n), flush=True)
    rng = np.random.default_rng(k + n)
    A = rng.standard_normal((k, n)) * np.logspace(0, -5, n)
    R = H.geqrf_r(A)
    Rn = np.linalg.qr(A, mode="r")
    sgn = np.sign(np.diag(R)) * np.sign(np.diag(Rn))
    assert np.allclose(np.tril(R, -1), 0)
    err = np.abs(R - sgn[:, None] * Rn).max() / np.abs(Rn).max()
    assert err < 1e-12, err
    ad = H.dev(np.asfortranarray(A))
    C.call("tomos_sketch_qr_precond_f64", k, n, ad.ptr, k, None)
    Ri = ad.to_numpy(np.float64, (k, n))[:n]
    assert np.abs(R @ Ri - np.eye(n)).max() < 1e-8


def test_qr_tall_wavefront_batched_matches_single(capi):
    """Two problems above the TSQR row limit in one batched call give the
    single calls' R^-1 bit for bit."""
    rng = np.random.default_rng(77)
    shapes = [(12000, 96), (8500, 40)]
    mats = [rng.standard_normal(s) * np.logspace(0, -4, s[1]) for s in shapes]
    singles = []
    for A in mats:
        ad = H.dev(np.asfortranarray(A))
        C.call("tomos_sketch_qr_precond_f64", A.shape[0], A.shape[1], ad.ptr, A.shape[0], None)
        singles.append(ad.to_numpy(np.float64, A.shape)[:A.shape[1]])
    bufs = [H.dev(np.asfortranarray(A)) for A in mats]
    qp = (C.QrProblem * 2)(*[C.QrProblem(A.shape[0], A.shape[1], A.shape[0], b.ptr) for A, b in zip(mats, bufs)])
    C.call("tomos_sketch_qr_precond_batched_f64", 2, qp, None, 0, None)
    for A, b, s in zip(mats, bufs, singles):
        assert np.array_equal(b.to_numpy(np.float64, A.shape)[:A.shape[1]], s)
