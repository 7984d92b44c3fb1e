"""GPU CountSketch (bit-exact indices), Householder QR and triangular inverse.

The QR tests run under both panel factorizations (tomos_qr_panels_set):
0 = CholeskyQR2 panels with reconstructed Householder vectors (falling
back to the column-by-column panels for an ill-conditioned panel), 1 = the
column-by-column Householder panels (TSQR for k <= 8192, the wavefront
kernel above).  R is checked against numpy's (LAPACK) QR up to row signs."""
import contextlib

import numpy as np
import pytest

from tests import hipcall as H
from tomos_amd import _capi as C
from tomos_amd._capi import DeviceBuffer

pytestmark = pytest.mark.gpu

PANELS = [0, 1]


@contextlib.contextmanager
def panels(mode):
    C.call("tomos_qr_panels_set", mode)
    try:
        yield
    finally:
        C.call("tomos_qr_panels_set", 0)


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


@pytest.mark.parametrize("a_type,tier", [(C.F32, 0), (C.FP32_BF16, 2), (C.FP32_16, 2)])
@pytest.mark.parametrize("spike", [False, True])
def test_apply_sampled_exponents(capi, orc, a_type, tier, spike):
    """A tall enough A takes the fixed-point exponents from a row sample
    (plus a margin); with `spike` one value far above the sampled range
    sits in a row no sample covers, so the bound check sends the call back
    to exponents over every row.  Both match the oracle."""
    m, n, s = 300000, 9, 2000
    rng = np.random.default_rng(21)
    A = rng.standard_normal((m, n)).astype(np.float32, order="F")
    if spike:
        A[5000, 3] = 3.0e7  # rows 1024..9367 are between the first two sampled segments
    if a_type == C.F32:
        Aref = A.astype(np.float64)
        Y = H.countsketch_apply(a_type, tier, m, n, H.dev(A), None, m, s)
    else:
        variant = C.PACK_FP32_BF16 if a_type == C.FP32_BF16 else C.PACK_FP32_16
        wh, wr = H.pack(variant, A)
        eh, er = orc.pack(0 if a_type == C.FP32_16 else 3, A)
        Aref = orc.unpack(4 if a_type == C.FP32_BF16 else 0, eh, er).astype(np.float64)
        Y = H.countsketch_apply(a_type, tier, m, n, wh, wr, m, s)
    ref = orc.countsketch_apply(Aref, s)
    for j in range(n):  # column by column: the spike dominates its column only
        err = np.abs(Y[:, j] - ref[:, j]).max() / np.abs(ref[:, j]).max()
        assert err < 2e-6, (j, err)


@pytest.mark.parametrize("a_type,tier", [(C.F32, 0), (C.F64, 0), (C.FP32_BF16, 2), (C.FP32_16, 2)])
@pytest.mark.parametrize("own_scratch", [False, True])
def test_apply_batched_matches_single(capi, a_type, tier, own_scratch):
    """tomos_countsketch_apply_batched over uneven problems equals the
    one-problem call bit for bit on the fixed-point path, grouped (4- and
    8-column groups, several splits, one split) or not (s > 16384): sampled
    exponents that hold, a spike that sends its problem back to full-row
    exponents, a problem below the sampling threshold, m = 0 (zeros).  An
    lda that is not a multiple of the 16-byte row vector (FP32_BF16/FP32_16:
    8 rows) takes the float path, whose LDS float atomics add in no fixed
    order: that one agrees to FP32 rounding.  n = 0 writes nothing."""
    rng = np.random.default_rng(33)
    specs = [(300000, 9, 2000, 0, True), (300000, 9, 2000, 0, False), (50000, 40, 160, 0, False),
             (5000, 20, 80, 0, False), (1001, 7, 28, 3, False), (0, 5, 20, 0, False), (100, 0, 10, 0, False),
             (20000, 3, 17000, 0, False), (70001, 13, 700, 7, False)]
    probs, keep, singles = [], [], []
    for m, n, s_rows, pad, spike in specs:
        lda = max(m + pad, 1)
        A = np.zeros((lda, n), dtype=np.float32, order="F")
        A[:m] = rng.standard_normal((m, n))
        if spike:
            A[5000, 3] = 3.0e7  # between the first two sampled segments
        if a_type == C.F32:
            a, wr = H.dev(A), None
        elif a_type == C.F64:
            a, wr = H.dev(A.astype(np.float64)), None
        else:
            variant = C.PACK_FP32_BF16 if a_type == C.FP32_BF16 else C.PACK_FP32_16
            a, wr = H.pack(variant, A)
        y = H.dev(np.full((s_rows, n), np.nan))
        keep += [a, wr, y]
        probs.append(C.CsProblem(m, n, lda, s_rows, s_rows, a.ptr, wr.ptr if wr else None, y.ptr))
        singles.append(H.countsketch_apply(a_type, tier, m, n, a, wr, lda, s_rows) if n else None)
    table = (C.CsProblem * len(probs))(*probs)
    if own_scratch:
        nb = C.lib().tomos_countsketch_apply_batched_scratch_size(len(probs), table)
        assert nb > 0
        scr = DeviceBuffer(nb)
        C.call("tomos_countsketch_apply_batched", a_type, tier, len(probs), table, scr.ptr, nb, None)
    else:
        C.call("tomos_countsketch_apply_batched", a_type, tier, len(probs), table, None, 0, None)
    for k, ((m, n, s_rows, pad, _), ref) in enumerate(zip(specs, singles)):
        got = keep[3 * k + 2].to_numpy(np.float64, (s_rows, n))
        if n == 0:
            continue
        if pad and a_type not in (C.F32, C.F64):  # float path
            assert np.abs(got - ref).max() <= 1e-6 * np.abs(ref).max(), (k, m, n)
        else:
            assert np.array_equal(got, ref), (k, m, n)
    rc = C.lib().tomos_countsketch_apply_batched(a_type, tier, len(probs), table, None, 1, None)
    assert rc == 0  # NULL scratch ignores the size
    if own_scratch:
        rc = C.lib().tomos_countsketch_apply_batched(a_type, tier, len(probs), table, scr.ptr, nb - 1, None)
        assert rc != 0  # scratch too small


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
@pytest.mark.parametrize("mode", PANELS)
def test_qr_r_matches_numpy(capi, k, n, mode):
    rng = np.random.default_rng(k + n)
    A = rng.standard_normal((k, n)) * np.logspace(0, -6, n)
    with panels(mode):
        R = H.geqrf_r(A)
    Rn = np.linalg.qr(A, mode="r")
    sgn = np.sign(np.diag(R)) * np.sign(np.diag(Rn))
    assert np.allclose(np.tril(R, -1), 0)
    err = np.abs(R - sgn[:, None] * Rn).max() / np.abs(Rn).max()
    assert err < 1e-12, err


@pytest.mark.parametrize("mode", PANELS)
def test_qr_zero_column(capi, mode):
    """A zero column gives an inactive reflector (tau = 0, R(c,c) = 0), as
    dlarfg; under the CholeskyQR2 panels its Gram matrix is singular and the
    factorization is redone with the Householder panels (same bits)."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((100, 40))
    A[:, 5] = 0.0
    A[:, 33] = 0.0
    with panels(mode):
        R = H.geqrf_r(A)
    with panels(1):
        assert np.array_equal(R, H.geqrf_r(A))
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


@pytest.mark.parametrize("mode", PANELS)
def test_batched_precond_matches_single(capi, mode):
    """Uneven sketches factored together give bit-identical R^-1 to one-at-a-time calls."""
    C.call("tomos_qr_panels_set", mode)
    try:
        _batched_precond_matches_single()
    finally:
        C.call("tomos_qr_panels_set", 0)


def _batched_precond_matches_single():
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


@pytest.mark.parametrize("mode", PANELS)
@pytest.mark.parametrize("k,n", [(12000, 300), (9000, 70)])
def test_qr_tall_wavefront_path(capi, k, n, mode):
    """k above 8192: with the Householder panels the panel wavefront
    factorization (a block per panel column, per-column flags between
    blocks) instead of TSQR; the CholeskyQR2 panels take any k.  R against
    numpy's QR up to row signs, and R^-1 from the sketch-QR entry point
    inverts it."""
    print("tall qr k=%d n=%d panels=%d" % (k, n, mode), flush=True)
    C.call("tomos_qr_panels_set", mode)
    try:
        _tall_qr(k, n)
    finally:
        C.call("tomos_qr_panels_set", 0)


def _tall_qr(k, n):
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


@pytest.mark.parametrize("mode", PANELS)
def test_qr_tall_wavefront_batched_matches_single(capi, mode):
    """Two problems above the TSQR row limit in one batched call give the
    single calls' R^-1 bit for bit."""
    C.call("tomos_qr_panels_set", mode)
    try:
        _tall_batched()
    finally:
        C.call("tomos_qr_panels_set", 0)


def _tall_batched():
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


def test_cholqr_panels_taken_and_fallback(capi):
    """Which panels ran: a well-conditioned sketch is factored by the
    CholeskyQR2 panels (its R differs in the last bits from the Householder
    panels' R, both within 1e-13 of LAPACK's); a sketch with two columns of
    one panel equal to 1e-9 (cond ~1e9, beyond two Gram passes) is redone
    with the Householder panels, bit for bit their result."""
    rng = np.random.default_rng(31)
    A = rng.standard_normal((3000, 200)) * np.logspace(0, -3, 200)
    Rn = np.linalg.qr(A, mode="r")
    R0 = H.geqrf_r(A)
    with panels(1):
        R1 = H.geqrf_r(A)
    assert not np.array_equal(R0, R1)
    for R in (R0, R1):
        sgn = np.sign(np.diag(R)) * np.sign(np.diag(Rn))
        assert np.abs(R - sgn[:, None] * Rn).max() < 1e-13 * np.abs(Rn).max()
    B = A.copy()
    B[:, 40] = B[:, 41] + 1e-9 * rng.standard_normal(3000) * np.abs(B[:, 41]).max()
    Rb = H.geqrf_r(B)
    with panels(1):
        assert np.array_equal(Rb, H.geqrf_r(B))
    # cond(B) ~ 1e9 determines R only to ~1e-7 (LAPACK's differs by that
    # much); what a backward-stable QR guarantees is R^T R = B^T B to rounding
    gram = B.T @ B
    assert np.abs(Rb.T @ Rb - gram).max() < 1e-13 * np.abs(gram).max()


@pytest.mark.parametrize("k,n,cond", [(4000, 1000, 1e6), (2000, 512, 1e10), (700, 64, 1e14)])
def test_cholqr_graded_columns(capi, k, n, cond):
    """Column scales spread over `cond` (the C4 / C5 sketches are of this
    kind): Cholesky is invariant to column scaling, so each panel stays
    well-conditioned and the CholeskyQR2 panels keep R within 1e-12 of
    LAPACK's, relative to each row's diagonal."""
    rng = np.random.default_rng(int(np.log10(cond)) + n)
    A = rng.standard_normal((k, n)) * np.logspace(0, -np.log10(cond), n)
    R = H.geqrf_r(A)
    with panels(1):
        R1 = H.geqrf_r(A)
    assert not np.array_equal(R, R1)  # the CholeskyQR2 panels ran
    Rn = np.linalg.qr(A, mode="r")
    sgn = np.sign(np.diag(R)) * np.sign(np.diag(Rn))
    rowerr = np.abs(R - sgn[:, None] * Rn).max(axis=1) / np.abs(np.diag(Rn))
    assert rowerr.max() < 1e-12, rowerr.max()
