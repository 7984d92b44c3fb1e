"""CPU oracle loader (test infrastructure only).

Thin ctypes/numpy wrapper over ``libtomos_oracle.so`` built from
``tomos_oracle.c`` in this folder.  Only tests, ``__graft_entry__.smoke()``
and bench.py's CPU-baseline leg import this module, as the checker or the
timed CPU baseline.  The product (``tomos_amd``) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TOMOS_ORACLE_LIB selects another build of the same source (the sanitizer
# build of tests/run_sanitizers.sh)
_LIB_PATH = os.environ.get("TOMOS_ORACLE_LIB") or os.path.join(_HERE, "libtomos_oracle.so")
_lib = None

i64 = ctypes.c_int64
u16p = ctypes.POINTER(ctypes.c_uint16)
f32p = ctypes.POINTER(ctypes.c_float)
f64p = ctypes.POINTER(ctypes.c_double)

# encoder / decoder ids (mirror tomos_oracle.c enums)
SPLIT16, SPLIT16_CLAMPED, SPLIT16_TRUNC, SPLITBF, SPLITBF_ROUNDED = range(5)
(DEC16_EXACT, DEC16_HALF, DEC16_HALF_FAST, DEC16_BF16, DECBF_EXACT, DECBF_HI,
 DECBF_EXACT_ROUNDED, DEC16_EXACT_TRUNC) = range(8)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_half_to_float.restype = ctypes.c_float
        L.orc_half_to_float.argtypes = [ctypes.c_uint16]
        L.orc_float_to_half.restype = ctypes.c_uint16
        L.orc_float_to_half.argtypes = [ctypes.c_float]
        L.orc_float_to_bf16.restype = ctypes.c_uint16
        L.orc_float_to_bf16.argtypes = [ctypes.c_float]
        for name in ("orc_fp32_16_split", "orc_fp32_16_split_clamped",
                     "orc_fp32_16_split_truncated", "orc_fp32_bf16_split",
                     "orc_fp32_bf16_split_rounded"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [ctypes.c_float, u16p, u16p]
        for name in ("orc_fp32_16_combine", "orc_fp32_16_combine_fast_normal",
                     "orc_fp32_16_combine_fast_normal_wrapborrow",
                     "orc_fp32_16_combine_branchless_nohw",
                     "orc_fp32_16_combine_truncated", "orc_fp32_bf16_combine",
                     "orc_fp32_bf16_combine_rounded"):
            f = getattr(L, name)
            f.restype = ctypes.c_float
            f.argtypes = [ctypes.c_uint16, ctypes.c_uint16]
        for name in ("orc_fp32_16_half_only", "orc_fp32_16_half_only_fast",
                     "orc_fp32_16_half_only_fast2", "orc_fp32_16_half_only_branchless",
                     "orc_fp32_16_half_only_no_overflow", "orc_fp32_16_bf16_only"):
            f = getattr(L, name)
            f.restype = ctypes.c_uint16
            f.argtypes = [ctypes.c_uint16]
        L.orc_pack.restype = None
        L.orc_pack.argtypes = [ctypes.c_int, i64, i64, f32p, i64, u16p, u16p, i64]
        L.orc_unpack.restype = None
        L.orc_unpack.argtypes = [ctypes.c_int, i64, i64, u16p, u16p, i64, f32p, i64]
        L.orc_mv_packed.restype = None
        L.orc_mv_packed.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, i64, i64,
                                    u16p, u16p, i64, f32p, ctypes.c_double, f32p]
        L.orc_mv_fp32_16_pure_half.restype = None
        L.orc_mv_fp32_16_pure_half.argtypes = [ctypes.c_int, ctypes.c_double, i64, i64, u16p,
                                               i64, f32p, ctypes.c_double, f32p]
        L.orc_mm_packed.restype = None
        L.orc_mm_packed.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    i64, i64, i64, u16p, u16p, i64, f32p, i64,
                                    ctypes.c_double, f32p, i64]
        L.orc_splitmix64.restype = ctypes.c_uint64
        L.orc_splitmix64.argtypes = [ctypes.c_uint64]
        L.orc_countsketch_indices.restype = None
        L.orc_countsketch_indices.argtypes = [i64, i64, ctypes.POINTER(ctypes.c_int32),
                                              ctypes.POINTER(ctypes.c_int8)]
        L.orc_countsketch_apply_f64.restype = None
        L.orc_countsketch_apply_f64.argtypes = [i64, i64, f64p, i64, i64, f64p, i64]
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.orc_csr_mv.restype = None
        L.orc_csr_mv.argtypes = [ctypes.c_int, ctypes.c_double, i64, i64, f64p, i64p, i64p, f64p,
                                 ctypes.c_double, f64p]
        L.orc_gaussian_fill_f64.restype = None
        L.orc_gaussian_fill_f64.argtypes = [i64, i64, f64p, i64, ctypes.c_uint64]
        L.orc_householder_qr.restype = ctypes.c_int
        L.orc_householder_qr.argtypes = [i64, i64, f64p, i64]
        L.orc_trtri_upper.restype = ctypes.c_int
        L.orc_trtri_upper.argtypes = [i64, f64p, i64]
        L.orc_lsqr_f64.restype = ctypes.c_int
        L.orc_lsqr_f64.argtypes = [i64, i64, f64p, i64, f64p, f64p, f64p, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), f64p, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int)]
        L.orc_lsqr_f32.restype = ctypes.c_int
        L.orc_lsqr_f32.argtypes = [i64, i64, f32p, i64, f32p, f32p, f32p, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), f64p, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int)]
        L.orc_blas_load.restype = ctypes.c_int
        L.orc_blas_load.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.orc_blas_loaded.restype = ctypes.c_int
        L.orc_blas_loaded.argtypes = []
        L.orc_countsketch_apply_f32_omp.restype = None
        L.orc_countsketch_apply_f32_omp.argtypes = [i64, i64, f32p, i64, i64, f64p, i64]
        L.orc_pack_compare_range.restype = ctypes.c_int64
        L.orc_pack_compare_range.argtypes = [ctypes.c_int, ctypes.c_uint64, i64, u16p, u16p,
                                             ctypes.POINTER(ctypes.c_int64)]
        L.orc_sgemv.restype = None
        L.orc_sgemv.argtypes = [ctypes.c_int, i64, i64, ctypes.c_float, f32p, i64, f32p, ctypes.c_float, f32p]
        L.orc_stream_copy.restype = None
        L.orc_stream_copy.argtypes = [i64, f64p, f64p]
        L.orc_exhaustive_split_combine.restype = ctypes.c_int64
        L.orc_exhaustive_split_combine.argtypes = [ctypes.c_int, ctypes.c_uint64,
                                                   ctypes.c_uint64,
                                                   ctypes.POINTER(ctypes.c_int64)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def pack(variant, A, ld_dst=None):
    """A: float32 (m, n) Fortran-ordered or any array; returns (wh, wr) as
    uint16 Fortran (ld_dst, n) arrays."""
    A = np.asfortranarray(A, dtype=np.float32)
    m, n = A.shape
    ld = ld_dst or m
    wh = np.zeros((ld, n), dtype=np.uint16, order="F")
    wr = np.zeros((ld, n), dtype=np.uint16, order="F")
    lib().orc_pack(variant, m, n, _p(A, f32p), max(m, 1), _p(wh, u16p), _p(wr, u16p), ld)
    return wh, wr


def unpack(dec, wh, wr, m=None):
    wh = np.asfortranarray(wh, dtype=np.uint16)
    wr = np.asfortranarray(wr if wr is not None else np.zeros_like(wh), dtype=np.uint16)
    ld, n = wh.shape
    m = m or ld
    out = np.zeros((m, n), dtype=np.float32, order="F")
    lib().orc_unpack(dec, m, n, _p(wh, u16p), _p(wr, u16p), ld, _p(out, f32p), max(m, 1))
    return out


def mv_packed(dec, trans, alpha, wh, wr, m, x, beta=0.0, y=None):
    wh = np.asfortranarray(wh, dtype=np.uint16)
    wr = np.asfortranarray(wr if wr is not None else np.zeros_like(wh), dtype=np.uint16)
    lda, n = wh.shape
    rows = n if trans else m
    y = np.zeros(rows, np.float32) if y is None else np.array(y, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    lib().orc_mv_packed(dec, int(trans), alpha, m, n, _p(wh, u16p), _p(wr, u16p), lda,
                        _p(x, f32p), beta, _p(y, f32p))
    return y


def mv_pure_half(trans, alpha, wh, m, x, beta=0.0, y=None):
    wh = np.asfortranarray(wh, dtype=np.uint16)
    lda, n = wh.shape
    rows = n if trans else m
    y = np.zeros(rows, np.float32) if y is None else np.array(y, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    lib().orc_mv_fp32_16_pure_half(int(trans), alpha, m, n, _p(wh, u16p), lda, _p(x, f32p),
                                   beta, _p(y, f32p))
    return y


def countsketch_indices(m, s):
    b = np.zeros(m, np.int32)
    sg = np.zeros(m, np.int8)
    lib().orc_countsketch_indices(m, s, _p(b, ctypes.POINTER(ctypes.c_int32)),
                                  _p(sg, ctypes.POINTER(ctypes.c_int8)))
    return b, sg


def countsketch_apply(X, s):
    X = np.asfortranarray(X, dtype=np.float64)
    m, n = X.shape
    Y = np.zeros((s, n), np.float64, order="F")
    lib().orc_countsketch_apply_f64(m, n, _p(X, f64p), max(m, 1), s, _p(Y, f64p), s)
    return Y


def csr_mv(trans, alpha, m, n, vals, row_ptrs, col_idxs, x, beta=0.0, y=None):
    """y = alpha op(A) x + beta y for CSR A (FP64 arithmetic)."""
    vals = np.ascontiguousarray(vals, np.float64)
    rp = np.ascontiguousarray(row_ptrs, np.int64)
    ci = np.ascontiguousarray(col_idxs, np.int64)
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(n if trans else m, np.float64) if y is None else np.array(y, np.float64)
    i64p = ctypes.POINTER(ctypes.c_int64)
    lib().orc_csr_mv(int(trans), alpha, m, n, _p(vals, f64p), _p(rp, i64p), _p(ci, i64p), _p(x, f64p),
                     beta, _p(out, f64p))
    return out


def gaussian_fill(rows, cols, seed):
    """(rows x cols) FP64 N(0,1) matrix of the seeded counter layout."""
    Y = np.zeros((rows, cols), np.float64, order="F")
    if rows and cols:
        lib().orc_gaussian_fill_f64(rows, cols, _p(Y, f64p), rows, seed & (2**64 - 1))
    return Y


def householder_r(A):
    """R (n x n upper) of the Householder QR of A (k x n), FP64."""
    A = np.array(A, dtype=np.float64, order="F")
    k, n = A.shape
    rc = lib().orc_householder_qr(k, n, _p(A, f64p), k)
    if rc != 0:
        raise ValueError("householder_qr: k < n")
    return np.triu(A[:n, :n])


def trtri_upper(R):
    R = np.array(R, dtype=np.float64, order="F")
    n = R.shape[0]
    rc = lib().orc_trtri_upper(n, _p(R, f64p), n)
    if rc != 0:
        raise ValueError("trtri: singular R (info=%d)" % rc)
    return R


def lsqr_f64(A, b, tol, max_iters, max_relres, restarts=1, Rinv=None, history_cap=100000):
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n, np.float64)
    hist = np.zeros(history_cap, np.float64)
    it = ctypes.c_int(0)
    hl = ctypes.c_int(0)
    Rp = None
    if Rinv is not None:
        Rinv = np.asfortranarray(Rinv, dtype=np.float64)
        Rp = _p(Rinv, f64p)
    lib().orc_lsqr_f64(m, n, _p(A, f64p), max(m, 1), _p(b, f64p), _p(x, f64p), Rp, tol,
                       max_iters, max_relres, restarts, ctypes.byref(it), _p(hist, f64p),
                       history_cap, ctypes.byref(hl))
    return x, it.value, hist[:min(hl.value, history_cap)].copy()


def openblas_path():
    import numpy
    d = os.path.join(os.path.dirname(os.path.dirname(numpy.__file__)), "numpy.libs")
    for f in sorted(os.listdir(d)) if os.path.isdir(d) else []:
        if f.startswith("libscipy_openblas64_"):
            return os.path.join(d, f)
    return None


def blas_load(threads):
    p = openblas_path()
    if p is None:
        return -1
    return lib().orc_blas_load(p.encode(), int(threads))


def lsqr_f32(A, b, tol, max_iters, max_relres, restarts=1, history_cap=100000, Rinv=None):
    """FP32 LSQR (OpenBLAS sgemv / BLAS-1 once blas_load() has bound them),
    optionally right-preconditioned by the n x n FP32 matrix Rinv."""
    A = np.asfortranarray(A, dtype=np.float32)
    m, n = A.shape
    b = np.ascontiguousarray(b, np.float32)
    x = np.zeros(n, np.float32)
    hist = np.zeros(history_cap, np.float64)
    it = ctypes.c_int(0)
    hl = ctypes.c_int(0)
    Rp = None
    if Rinv is not None:
        Rinv = np.asfortranarray(Rinv, dtype=np.float32)
        assert Rinv.shape == (n, n)
        Rp = _p(Rinv, f32p)
    lib().orc_lsqr_f32(m, n, _p(A, f32p), max(m, 1), _p(b, f32p), _p(x, f32p), Rp, tol, max_iters,
                       max_relres, restarts, ctypes.byref(it), _p(hist, f64p), history_cap,
                       ctypes.byref(hl))
    return x, it.value, hist[:min(hl.value, history_cap)].copy()


def countsketch_apply_f32(A, s):
    """Hashed CountSketch S A of an FP32 matrix, FP64 result (s x n), over
    the host cores (OpenMP across columns)."""
    A = np.asfortranarray(A, dtype=np.float32)
    m, n = A.shape
    Y = np.zeros((s, n), np.float64, order="F")
    lib().orc_countsketch_apply_f32_omp(m, n, _p(A, f32p), max(m, 1), s, _p(Y, f64p), s)
    return Y


def sgemv_ms(trans, A, reps=3):
    """Best time (ms) of y = A x (trans 0) or y = A^T x (trans 1) through
    the bound cblas_sgemv, on an FP32 Fortran-ordered A."""
    import time
    A = np.asfortranarray(A, dtype=np.float32)
    m, n = A.shape
    x = np.ones(m if trans else n, np.float32)
    y = np.zeros(n if trans else m, np.float32)
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        lib().orc_sgemv(int(trans), m, n, 1.0, _p(A, f32p), max(m, 1), _p(x, f32p), 0.0, _p(y, f32p))
        best = min(best, time.perf_counter() - t0)
    return best * 1e3


def stream_copy_gbs(nbytes=1 << 31, reps=5):
    """Host copy bandwidth (STREAM copy, OpenMP threads), GB/s counting the
    read and the write; best of reps."""
    import time
    n = nbytes // 8
    a = np.ones(n, np.float64)
    d = np.empty(n, np.float64)
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        lib().orc_stream_copy(n, _p(a, f64p), _p(d, f64p))
        best = min(best, time.perf_counter() - t0)
    return 2.0 * 8 * n / best / 1e9
