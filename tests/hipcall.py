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
