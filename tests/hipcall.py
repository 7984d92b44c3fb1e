"""Small helpers for GPU tests: numpy <-> device and typed C-ABI calls."""
import ctypes

import numpy as np

from tomos_amd import _capi as C
from tomos_amd._capi import DeviceBuffer


def dev(arr):
    return DeviceBuffer.from_numpy(np.asfortranarray(arr))


def pack(variant, A, ld=None):
    A = np.asfortranarray(A, dtype=np.float32)
    m, n = A.shape
    ld = ld or m
    src = dev(A)
    wh = DeviceBuffer(max(ld * n * 2, 2))
    wr = DeviceBuffer(max(ld * n * 2, 2))
    wh.zero()
    wr.zero()
    C.call("tomos_pack", variant, m, n, src.ptr, max(m, 1), wh.ptr, wr.ptr, ld, None)
    return wh, wr


def from_dev16(buf, ld, n):
    return buf.to_numpy(np.uint16, (ld, n))


def gemv(a_type, tier, trans, alpha, m, n, a_buf, wr_buf, lda, x, beta=0.0, y0=None,
         x_type=C.F32, y_type=C.F32, scratch=None):
    xdt = np.float64 if x_type == C.F64 else np.float32
    ydt = np.float64 if y_type == C.F64 else np.float32
    rows = n if trans else m
    xd = dev(np.asarray(x, dtype=xdt))
    y_init = np.zeros(rows, ydt) if y0 is None else np.asarray(y0, dtype=ydt)
    yd = dev(y_init)
    sc_ptr, sc_bytes = None, 0
    if scratch == "explicit":
        need = C.lib().tomos_gemv_scratch_size(a_type, tier, trans, m, n)
        sb = DeviceBuffer(max(need, 1))
        sc_ptr, sc_bytes = sb.ptr, need
    C.call("tomos_gemv", a_type, tier, trans, alpha, m, n, a_buf.ptr if a_buf else None,
           wr_buf.ptr if wr_buf else None, lda, x_type, xd.ptr, beta, y_type, yd.ptr, sc_ptr,
           sc_bytes, None)
    return yd.to_numpy(ydt, (rows,))


def lsqr(A_or_planes, b, a_type=C.F32, mv_tier=C.TIER_DEFAULT, restart_tier=None,
         value_type=C.F32, refine_type=C.F32, rinv=None, precond_type=-1, tol=1e-6,
         max_iters=400, max_relres=None, restarts=1, check_every=1, lda=None, m=None, n=None,
         fused=0, history_cap=100000, refine_dup=None, confirm=0):
    """Runs tomos_lsqr_solve; A_or_planes is a numpy matrix (F32/F64) or a
    (wh_buf, wr_buf) pair of packed device planes (then m, n, lda needed)."""
    if restart_tier is None:
        restart_tier = mv_tier
    if max_relres is None:
        max_relres = tol
    if isinstance(A_or_planes, tuple):
        a_buf, wr_buf = A_or_planes
        assert m is not None and n is not None
        lda = lda or m
    else:
        A = np.asfortranarray(A_or_planes)
        m, n = A.shape
        lda = m
        a_buf, wr_buf = dev(A), None
    rdt = np.float64 if refine_type == C.F64 else np.float32
    bd = dev(np.asarray(b, dtype=rdt))
    xd = DeviceBuffer(max(n, 1) * np.dtype(rdt).itemsize)
    rd = None
    ldr = max(n, 1)
    if rinv is not None:
        pdt = np.float64 if precond_type == C.F64 else np.float32
        rd = dev(np.asarray(rinv, dtype=pdt))
    opt = C.LsqrOptions(a_type, mv_tier, restart_tier, value_type, refine_type, precond_type,
                        tol, max_relres, max_iters, restarts, check_every, fused, confirm)
    ab = C.lib().tomos_lsqr_arena_size(ctypes.byref(opt), m, n)
    sb = C.lib().tomos_lsqr_scratch_size(ctypes.byref(opt), m, n)
    assert ab > 0 and sb > 0, C.lib().tomos_last_error()
    arena, scr = DeviceBuffer(ab), DeviceBuffer(sb)
    it = ctypes.c_int(0)
    hl = ctypes.c_int(0)
    hist = np.zeros(history_cap, np.float64)
    if refine_dup is None:
        C.call("tomos_lsqr_solve", ctypes.byref(opt), m, n, a_buf.ptr, wr_buf.ptr if wr_buf else None,
               lda, bd.ptr, xd.ptr, rd.ptr if rd is not None else None, ldr, arena.ptr, ab, scr.ptr,
               sb, ctypes.byref(it), hist.ctypes.data, history_cap, ctypes.byref(hl), None)
    else:
        # (ref_type, device buffer, lda): restarts read this duplicate
        ref_type, ref_buf, lda_ref = refine_dup
        C.call("tomos_lsqr_solve_refined", ctypes.byref(opt), m, n, a_buf.ptr, wr_buf.ptr if wr_buf else None,
               lda, ref_type, ref_buf.ptr, lda_ref, bd.ptr, xd.ptr, rd.ptr if rd is not None else None, ldr,
               arena.ptr, ab, scr.ptr, sb, ctypes.byref(it), hist.ctypes.data, history_cap, ctypes.byref(hl),
               None)
    x = xd.to_numpy(rdt, (n,))
    return x, it.value, hist[:min(hl.value, history_cap)].copy()


def countsketch_indices(m, s):
    b = DeviceBuffer(m * 4)
    g = DeviceBuffer(m)
    C.call("tomos_countsketch_indices", m, s, b.ptr, g.ptr, None)
    return b.to_numpy(np.int32, (m,)), g.to_numpy(np.int8, (m,))


def countsketch_apply(a_type, tier, m, n, a_buf, wr_buf, lda, s):
    y = DeviceBuffer(s * n * 8)
    C.call("tomos_countsketch_apply", a_type, tier, m, n, a_buf.ptr, wr_buf.ptr if wr_buf else None,
           lda, s, y.ptr, s, None, 0, None)
    return y.to_numpy(np.float64, (s, n))


def geqrf_r(A):
    A = np.asfortranarray(A, dtype=np.float64)
    k, n = A.shape
    ad = dev(A)
    r = DeviceBuffer(n * n * 8)
    C.call("tomos_geqrf_r_f64", k, n, ad.ptr, k, r.ptr, n, None, 0, None)
    return r.to_numpy(np.float64, (n, n))


def trtri(R):
    R = np.asfortranarray(R, dtype=np.float64)
    n = R.shape[0]
    rd = dev(R)
    C.call("tomos_trtri_upper_f64", n, rd.ptr, n, None)
    return rd.to_numpy(np.float64, (n, n))


def sketch_rinv(a_type, tier, m, n, a_buf, wr_buf, lda, s):
    """CountSketch (s x n) of A on the GPU, then R^-1 of its QR (GPU)."""
    y = DeviceBuffer(s * n * 8)
    C.call("tomos_countsketch_apply", a_type, tier, m, n, a_buf.ptr, wr_buf.ptr if wr_buf else None,
           lda, s, y.ptr, s, None, 0, None)
    C.call("tomos_sketch_qr_precond_f64", s, n, y.ptr, s, None)
    full = y.to_numpy(np.float64, (s, n))
    return np.triu(full[:n, :n]), y


# ---- CSR -----------------------------------------------------------------------
VAL_NP = {C.F64: np.float64, C.F32: np.float32, C.F16: np.float16}


def csr_vals_bits(vt, vals):
    """Host values in the device encoding of vt (bf16 rounds to nearest even)."""
    v = np.asarray(vals, np.float64)
    if vt == C.BF16:
        u = v.astype(np.float32).view(np.uint32).astype(np.uint64)
        return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return v.astype(VAL_NP[vt])


def csr_decoded(vt, vals):
    """The stored values as FP64 (what the device products multiply)."""
    b = csr_vals_bits(vt, vals)
    if vt == C.BF16:
        return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return b.astype(np.float64)


class Csr:
    """A CSR matrix on the device (values encoded as vt, indices as it)."""

    def __init__(self, m, n, row_ptrs, col_idxs, vals, vt=C.F64, it=C.INDEX_INT32):
        idt = np.int64 if it == C.INDEX_INT64 else np.int32
        self.m, self.n, self.vt, self.it = m, n, vt, it
        self.nnz = len(col_idxs)
        self.rp = dev(np.asarray(row_ptrs, idt))
        self.ci = dev(np.asarray(col_idxs, idt)) if self.nnz else DeviceBuffer(8)
        self.v = dev(csr_vals_bits(vt, vals)) if self.nnz else DeviceBuffer(8)
        self.host = (np.asarray(row_ptrs, np.int64), np.asarray(col_idxs, np.int64), csr_decoded(vt, vals))

    def args(self):
        return (self.vt, self.it, self.m, self.n, self.nnz)


def csr_mv(A, trans, alpha, x, beta=0.0, y0=None, xt=C.F64, yt=C.F64):
    xdt = np.float64 if xt == C.F64 else np.float32
    ydt = np.float64 if yt == C.F64 else np.float32
    ylen = A.n if trans else A.m
    xd = dev(np.asarray(x, xdt)) if len(x) else DeviceBuffer(8)
    yd = dev(np.asarray(y0, ydt)) if y0 is not None else DeviceBuffer(max(ylen, 1) * np.dtype(ydt).itemsize)
    C.call("tomos_csr_mv", A.vt, A.it, trans, alpha, A.m, A.n, A.nnz, A.v.ptr, A.rp.ptr, A.ci.ptr, xt, xd.ptr,
           beta, yt, yd.ptr, None, 0, None)
    return yd.to_numpy(ydt, (ylen,))


def csr_transpose(A):
    idt = np.int64 if A.it == C.INDEX_INT64 else np.int32
    vdt = np.uint16 if A.vt == C.BF16 else VAL_NP[A.vt]
    vt, pt, it = DeviceBuffer(max(A.nnz, 1) * 8), DeviceBuffer((A.n + 1) * 8), DeviceBuffer(max(A.nnz, 1) * 8)
    C.call("tomos_csr_transpose", *A.args(), A.v.ptr, A.rp.ptr, A.ci.ptr, vt.ptr, pt.ptr, it.ptr, None, 0, None)
    return (pt.to_numpy(idt, (A.n + 1,)), it.to_numpy(idt, (A.nnz,)), vt.to_numpy(vdt, (A.nnz,)))


def lsqr_csr(A, b, value_type=C.F64, refine_type=C.F64, rinv=None, precond_type=-1, tol=1e-10,
             max_iters=400, max_relres=None, restarts=1, check_every=1, ref=None, history_cap=100000):
    """tomos_lsqr_solve_csr; ref = (ref_type, values) for the refine duplicate."""
    if max_relres is None:
        max_relres = tol
    rdt = np.float64 if refine_type == C.F64 else np.float32
    bd = dev(np.asarray(b, dtype=rdt))
    xd = DeviceBuffer(max(A.n, 1) * np.dtype(rdt).itemsize)
    rd = None
    if rinv is not None:
        pdt = np.float64 if precond_type == C.F64 else np.float32
        rd = dev(np.asarray(rinv, dtype=pdt))
    opt = C.LsqrOptions(A.vt, C.TIER_DEFAULT, C.TIER_DEFAULT, value_type, refine_type, precond_type,
                        tol, max_relres, max_iters, restarts, check_every, 0)
    ab = C.lib().tomos_lsqr_arena_size(ctypes.byref(opt), A.m, A.n)
    sb = C.lib().tomos_lsqr_csr_scratch_size(ctypes.byref(opt), A.it, A.m, A.n, A.nnz)
    assert ab > 0 and sb > 0, C.lib().tomos_last_error()
    arena, scr = DeviceBuffer(ab), DeviceBuffer(sb)
    it = ctypes.c_int(0)
    hl = ctypes.c_int(0)
    hist = np.zeros(history_cap, np.float64)
    rt, rv = -1, None
    if ref is not None:
        rt = ref[0]
        rv = dev(np.asarray(ref[1], np.float64 if rt == C.F64 else np.float32))
    C.call("tomos_lsqr_solve_csr", ctypes.byref(opt), A.it, A.m, A.n, A.nnz, A.v.ptr, A.rp.ptr, A.ci.ptr, rt,
           rv.ptr if rv is not None else None, bd.ptr, xd.ptr, rd.ptr if rd is not None else None, max(A.n, 1),
           arena.ptr, ab, scr.ptr, sb, ctypes.byref(it), hist.ctypes.data, history_cap, ctypes.byref(hl), None)
    x = xd.to_numpy(rdt, (A.n,))
    return x, it.value, hist[:min(hl.value, history_cap)].copy()
