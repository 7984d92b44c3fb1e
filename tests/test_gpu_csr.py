"""GPU CSR products, transpose and LSQR over CSR matrices against the
oracle (orc_csr_mv, matrix_cpu_kernels.cpp:92-131; orc_lsqr_f64 on the
same matrix densified: the reference's sparse LSQR runs the dense
iteration with CSR products, lsqr_registries.cpp:164-196) and the
reference's own sparse known-answer tests (test_lsqr.cpp:375-440, :880-925).

Tolerances: products accumulate in FP64 on both sides, in different
orders: |gpu - oracle| <= 1e-13 * sum|a_k x_k| per entry, plus FP32
rounding of y when y is FP32.  The transpose is integer/byte work and is
compared exactly."""
import numpy as np
import pytest

from tests import hipcall as H
from tomos_amd import _capi as C

pytestmark = pytest.mark.gpu  # the helpers below are shared with the CPU tests


def random_csr(rng, m, n, mean, empty_frac=0.1):
    """Row lengths ~ Poisson(mean) (some rows empty), unsorted columns with
    repeats, N(0,1) values."""
    lens = rng.poisson(mean, m) if m else np.zeros(0, int)
    lens[rng.random(m) < empty_frac] = 0
    if n == 0:
        lens[:] = 0
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ci = rng.integers(0, max(n, 1), int(rp[-1])).astype(np.int64)
    v = rng.standard_normal(int(rp[-1]))
    return rp, ci, v


SHAPES = [
    (1000, 37, 1.0),     # 1 lane per row
    (3000, 300, 8.0),    # 4 lanes
    (500, 2000, 30.0),   # 16 lanes
    (64, 5000, 300.0),   # 64 lanes
    (200000, 100, 5.0),  # more rows than one pass of the grid
    (8, 100000, 20000.0),  # few long rows: rows split into segments
    (300, 50000, 2000.0),  # split, fewer segments per row
    (3, 4, 0.0),         # no entries
    (0, 5, 0.0),         # no rows
]


def bound(A, x, trans):
    rp, ci, v = A.host
    S = np.zeros(A.n if trans else A.m)
    rows = np.repeat(np.arange(A.m), np.diff(rp))
    if trans:
        np.add.at(S, ci, np.abs(v) * np.abs(x[rows]))
    else:
        np.add.at(S, rows, np.abs(v) * np.abs(x[ci]))
    return S


@pytest.mark.parametrize("vt", [C.F64, C.F32, C.F16, C.BF16])
@pytest.mark.parametrize("it", [C.INDEX_INT32, C.INDEX_INT64])
@pytest.mark.parametrize("m,n,mean", SHAPES)
@pytest.mark.parametrize("trans", [0, 1])
def test_mv_matches_oracle(capi, orc, vt, it, m, n, mean, trans):
    rng = np.random.default_rng(int(m * 7 + n + mean * 13 + vt))
    rp, ci, v = random_csr(rng, m, n, mean)
    A = H.Csr(m, n, rp, ci, v, vt, it)
    x = rng.standard_normal(m if trans else n)
    y0 = rng.standard_normal(n if trans else m)
    got = H.csr_mv(A, trans, 1.5, x, beta=-0.5, y0=y0)
    ref = orc.csr_mv(trans, 1.5, m, n, A.host[2], rp, ci, x, beta=-0.5, y=y0)
    tol = 1e-13 * (1.5 * bound(A, x, trans) + 0.5 * np.abs(y0)) + 1e-300
    assert np.all(np.abs(got - ref) <= tol), np.max(np.abs(got - ref) / tol)


@pytest.mark.parametrize("trans", [0, 1])
def test_f32_vectors_and_beta_zero_ignores_y(capi, orc, trans):
    rng = np.random.default_rng(5)
    m, n = 5000, 80
    rp, ci, v = random_csr(rng, m, n, 6.0)
    A = H.Csr(m, n, rp, ci, v, C.F32)
    x = rng.standard_normal(m if trans else n).astype(np.float32)
    y0 = np.full(n if trans else m, np.nan, np.float32)
    got = H.csr_mv(A, trans, 2.0, x, beta=0.0, y0=y0, xt=C.F32, yt=C.F32)
    ref = orc.csr_mv(trans, 2.0, m, n, A.host[2], rp, ci, x.astype(np.float64))
    assert np.all(np.isfinite(got))
    assert np.all(np.abs(got - ref) <= 2.0 ** -23 * np.abs(ref) + 1e-12 * 2 * bound(A, x.astype(np.float64), trans))


def test_repeatable_bit_for_bit(capi):
    rng = np.random.default_rng(6)
    rp, ci, v = random_csr(rng, 40000, 700, 20.0)
    A = H.Csr(40000, 700, rp, ci, v, C.F64)
    x, u = rng.standard_normal(700), rng.standard_normal(40000)
    assert np.array_equal(H.csr_mv(A, 0, 1.0, x), H.csr_mv(A, 0, 1.0, x))
    assert np.array_equal(H.csr_mv(A, 1, 1.0, u), H.csr_mv(A, 1, 1.0, u))


@pytest.mark.parametrize("vt", [C.F64, C.F32, C.BF16])
@pytest.mark.parametrize("it", [C.INDEX_INT32, C.INDEX_INT64])
def test_transpose_exact(capi, vt, it):
    """Row c of A^T holds A's entries of column c in A's order (a stable sort
    of the column indices)."""
    rng = np.random.default_rng(7)
    m, n = 30000, 513
    rp, ci, v = random_csr(rng, m, n, 9.0)
    A = H.Csr(m, n, rp, ci, v, vt, it)
    pt, idx, vals = H.csr_transpose(A)
    perm = np.argsort(ci, kind="stable")
    rows = np.repeat(np.arange(m), np.diff(rp))
    assert np.array_equal(pt, np.searchsorted(ci[perm], np.arange(n + 1)))
    assert np.array_equal(idx, rows[perm])
    assert np.array_equal(vals, H.csr_vals_bits(vt, v)[perm])


@pytest.mark.parametrize("rp,ci,msg", [
    ([1, 2, 3], [0, 1], "row_ptrs\\[0\\] must be 0"),
    ([0, 2, 1], [0, 1], "non-decreasing"),
    ([0, 1, 3], [0, 1], "must equal nnz"),
    ([0, 1, 2], [0, 9], "column index out of range"),
    ([0, 1, 2], [0, -1], "column index out of range"),
])
def test_invalid_csr_rejected(capi, rp, ci, msg):
    A = H.Csr(2, 3, rp, ci, [1.0, 2.0])
    with pytest.raises(C.TomosError, match=msg):
        C.call("tomos_csr_check", *A.args(), A.rp.ptr, A.ci.ptr, None)
    with pytest.raises(C.TomosError, match=msg):
        H.csr_mv(A, 1, 1.0, np.ones(2))
    with pytest.raises(C.TomosError, match=msg):
        H.lsqr_csr(A, np.ones(2))


# ---- LSQR over CSR ------------------------------------------------------------
def diag3():
    return H.Csr(3, 3, [0, 1, 2, 3], [0, 1, 2], [1.0, 2.0, 4.0])


def csr_4x2(vt=C.F64):
    # rows [1,0],[0,1],[1,1],[1,-1] (test_lsqr.cpp:407-440)
    return H.Csr(4, 2, [0, 1, 2, 4, 6], [0, 1, 0, 1, 0, 1], [1, 1, 1, 1, 1, -1], vt)


def test_known_answers(capi):
    x, it, _ = H.lsqr_csr(diag3(), [1.0, 4.0, 8.0], tol=1e-10, max_iters=50)
    assert np.allclose(x, [1, 2, 2], atol=1e-8, rtol=0)
    x, it, _ = H.lsqr_csr(csr_4x2(), [1.0, 1.0, 2.0, 0.0], tol=1e-10, max_iters=50)
    assert np.allclose(x, [1, 1], atol=1e-8, rtol=0)


@pytest.mark.parametrize("mv", [C.F64, C.F32, C.F16])
@pytest.mark.parametrize("val", [C.F64, C.F32])
@pytest.mark.parametrize("ref", [C.F64, C.F32])
@pytest.mark.parametrize("pre", [C.F64, C.F32])
def test_precision_combos_preconditioned(capi, mv, val, ref, pre):
    """LsqrAllPrecisionCombosTest, sparse half (test_lsqr.cpp:880-925):
    approximate preconditioner diag(0.9, 0.9), tol 1e-8, |x - 1| <= 1e-3
    (0.2 when a type is FP16)."""
    x, it, _ = H.lsqr_csr(csr_4x2(mv), [1.0, 1.0, 2.0, 0.0], value_type=val, refine_type=ref,
                          rinv=np.diag([0.9, 0.9]), precond_type=pre, tol=1e-8, max_iters=100)
    assert np.allclose(x, [1, 1], atol=0.2 if mv == C.F16 else 1e-3, rtol=0), (x, it)


def sparse_problem(seed, m=4000, n=80, mean=6.0, noise=1e-3):
    rng = np.random.default_rng(seed)
    rp, ci, v = random_csr(rng, m, n, mean, empty_frac=0.0)
    rows = np.repeat(np.arange(m), np.diff(rp))
    v = v * np.logspace(0, -2, n)[ci]  # condition ~1e2
    D = np.zeros((m, n))
    np.add.at(D, (rows, ci), v)
    xt = rng.standard_normal(n)
    b = D @ xt + noise * rng.standard_normal(m)
    return rp, ci, v, D, xt, b


def test_matches_oracle_iteration_by_iteration(capi, orc):
    rp, ci, v, D, xt, b = sparse_problem(11)
    A = H.Csr(D.shape[0], D.shape[1], rp, ci, v, C.F64)
    x, it, hist = H.lsqr_csr(A, b, tol=1e-9, max_iters=400)
    xo, ito, histo = orc.lsqr_f64(D, b, 1e-9, 400, 1e-9)
    assert abs(it - ito) <= 1
    assert np.allclose(hist[:15], histo[:15], rtol=1e-9, atol=0)
    r = lambda z: np.linalg.norm(D @ z - b) / np.linalg.norm(b)
    assert abs(r(x) - r(xo)) <= 1e-3 * r(xo)
    x2, it2, hist2 = H.lsqr_csr(A, b, tol=1e-9, max_iters=400)
    assert it2 == it and np.array_equal(x2, x) and np.array_equal(hist2, hist)


def test_tall_matrix_uses_split_transpose_and_matches_oracle(capi, orc):
    """A^T of a 100000 x 50 matrix has 50 rows of ~16000 entries: the
    segmented kernel forms A^T u; iterates still follow the oracle."""
    rp, ci, v, D, xt, b = sparse_problem(14, m=100000, n=50, mean=8.0)
    A = H.Csr(D.shape[0], D.shape[1], rp, ci, v, C.F64)
    x, it, hist = H.lsqr_csr(A, b, tol=1e-9, max_iters=300)
    xo, ito, histo = orc.lsqr_f64(D, b, 1e-9, 300, 1e-9)
    assert abs(it - ito) <= 1
    assert np.allclose(hist[:15], histo[:15], rtol=1e-9, atol=0)
    x2, it2, _ = H.lsqr_csr(A, b, tol=1e-9, max_iters=300)
    assert it2 == it and np.array_equal(x2, x)


@pytest.mark.parametrize("it_type", [C.INDEX_INT32, C.INDEX_INT64])
def test_preconditioned_and_restarted(capi, it_type):
    rp, ci, v, D, xt, b = sparse_problem(12, noise=0.0)
    R = np.linalg.qr(D, mode="r")
    rinv = np.linalg.inv(R)
    A = H.Csr(D.shape[0], D.shape[1], rp, ci, v, C.F32, it_type)
    D32 = H.Csr(D.shape[0], D.shape[1], rp, ci, v, C.F32).host[2]
    x, it, hist = H.lsqr_csr(A, b, value_type=C.F32, refine_type=C.F64, rinv=rinv, precond_type=C.F64,
                             tol=1e-10, max_iters=10, restarts=10)
    Dr = np.zeros_like(D)
    rows = np.repeat(np.arange(D.shape[0]), np.diff(rp))
    np.add.at(Dr, (rows, ci), D32)
    assert np.linalg.norm(Dr @ x - b) / np.linalg.norm(b) < 2e-7 and it >= 1


def test_explicit_duplicate_restarts(capi):
    """F16 values per iteration, FP64 values of the same pattern for the
    restart residuals (mtx_refine_precision, lsqr_registries.cpp:180-183)."""
    rp, ci, v, D, xt, b = sparse_problem(13, noise=0.0)
    A16 = H.Csr(D.shape[0], D.shape[1], rp, ci, v, C.F16)
    R = np.linalg.qr(D, mode="r")
    rinv = np.linalg.inv(R)
    r = lambda z: np.linalg.norm(D @ z - b) / np.linalg.norm(b)
    x0, _, _ = H.lsqr_csr(A16, b, rinv=rinv, precond_type=C.F64, tol=1e-9, max_iters=8, restarts=30)
    x1, it1, _ = H.lsqr_csr(A16, b, rinv=rinv, precond_type=C.F64, tol=1e-9, max_iters=8, restarts=30,
                            ref=(C.F64, v))
    assert r(x1) < 2e-9 and r(x0) > 10 * r(x1), (r(x0), r(x1))


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("it", [C.INDEX_INT32, C.INDEX_INT64])
def test_csr_sketch_pattern(capi, orc, kind, it):
    """The reference's CSR sketch form (one entry per row,
    countsketch_kernels.cpp:34-53): row i holds column z_i % cols, z_i the
    i-th SplitMix64 output from seed ^ 0xc2b2ae3d27d4eb4f; values +-1 from
    z_i's top bit, or the Gaussian fill of a rows x 1 column."""
    from tests.test_oracle_gaussian import stream
    rows, cols, seed = 1001, 70000, 1234
    idt = np.int64 if it == C.INDEX_INT64 else np.int32
    rp, ci, v = (H.DeviceBuffer((rows + 1) * 8), H.DeviceBuffer(rows * 8), H.DeviceBuffer(rows * 8))
    C.call("tomos_csr_sketch", kind, C.F64, it, rows, cols, v.ptr, rp.ptr, ci.ptr, seed, None)
    z = [stream(seed ^ 0xC2B2AE3D27D4EB4F, i) for i in range(rows)]
    assert np.array_equal(rp.to_numpy(idt, (rows + 1,)), np.arange(rows + 1))
    assert np.array_equal(ci.to_numpy(idt, (rows,)), [q % cols for q in z])
    vals = v.to_numpy(np.float64, (rows,))
    if kind == 0:
        assert np.array_equal(vals, [1.0 if q >> 63 else -1.0 for q in z])
    else:
        ref = orc.gaussian_fill(rows, 1, seed)[:, 0]
        assert np.all(np.abs(vals - ref) <= 1e-6 * (1 + np.abs(ref)))


@pytest.mark.parametrize("it", [C.INDEX_INT32, C.INDEX_INT64])
@pytest.mark.parametrize("xt,yt,beta", [(C.F64, C.F64, 0.0), (C.F32, C.F64, 0.5), (C.F64, C.F32, -1.0)])
def test_csr_mm_vs_oracle_columns(capi, orc, it, xt, yt, beta):
    """Y = alpha A X + beta Y (tomos_csr_mm, the apply_count_sketch kernel) =
    the oracle's CSR product column by column; ragged rows, repeated and
    empty rows, ldx / ldy above the row counts."""
    import ctypes  # noqa: F401
    rng = np.random.default_rng(31 + it)
    m, n, ncols = 700, 900, 37
    rp, ci, v = random_csr(rng, m, n, 3.0)
    idt = np.int64 if it == C.INDEX_INT64 else np.int32
    xdt = np.float64 if xt == C.F64 else np.float32
    ydt = np.float64 if yt == C.F64 else np.float32
    ldx, ldy = n + 3, m + 5
    X = np.zeros((ldx, ncols), xdt, order="F")
    X[:n] = rng.standard_normal((n, ncols))
    Y0 = np.zeros((ldy, ncols), ydt, order="F")
    Y0[:m] = rng.standard_normal((m, ncols))
    vals, rpd, cid = H.dev(v), H.dev(rp.astype(idt)), H.dev(ci.astype(idt))
    Xd, Yd = H.dev(X), H.dev(Y0)
    C.call("tomos_csr_mm", C.F64, it, 2.0, m, n, len(v), vals.ptr, rpd.ptr, cid.ptr, ncols, xt, Xd.ptr, ldx, beta,
           yt, Yd.ptr, ldy, None)
    Y = Yd.to_numpy(ydt, (ldy, ncols))
    for j in range(ncols):
        ref = orc.csr_mv(0, 2.0, m, n, v, rp, ci, X[:n, j].astype(np.float64), beta, Y0[:m, j].astype(np.float64))
        tol = 1e-13 if yt == C.F64 else 1e-6
        scale = np.abs(ref).max() + 1.0
        assert np.max(np.abs(Y[:m, j] - ref)) <= tol * scale * 4, j
    # rows between m and ldy are untouched
    assert np.array_equal(Y[m:], Y0[m:])


def test_convert_roundtrip(capi):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((50, 7))
    src = np.zeros((53, 7), np.float64, order="F")
    src[:50] = A
    s, d32, d64 = H.dev(src), H.DeviceBuffer(52 * 7 * 4), H.DeviceBuffer(50 * 7 * 8)
    C.call("tomos_convert", C.F64, C.F32, 50, 7, s.ptr, 53, d32.ptr, 52, None)
    C.call("tomos_convert", C.F32, C.F64, 50, 7, d32.ptr, 52, d64.ptr, 50, None)
    got32 = d32.to_numpy(np.float32, (52, 7))[:50]
    assert np.array_equal(got32, A.astype(np.float32))
    assert np.array_equal(d64.to_numpy(np.float64, (50, 7)), A.astype(np.float32).astype(np.float64))
    assert C.lib().tomos_convert(C.F16, C.F32, 1, 1, s.ptr, 1, d32.ptr, 1, None) == 1
