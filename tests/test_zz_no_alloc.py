"""tomos_gemv_noalloc (the GEMV whose partial sums live in a fixed device
array instead of a caller or library allocation) against tomos_gemv with an
explicit scratch buffer: the same kernels on the same inputs, so the results
agree bit for bit.  Each case makes two calls in a row, the second on a
smaller problem, so stale partials of the first call in the shared array
would show in the second."""
import numpy as np
import pytest

from tomos_amd import _capi as C
from tomos_amd._capi import DeviceBuffer

pytestmark = pytest.mark.gpu


def _problem(rng, a_type, m, n):
    A = np.asfortranarray(rng.standard_normal((m, n)).astype(np.float32))
    if a_type == C.F32:
        return DeviceBuffer.from_numpy(A), None
    src = DeviceBuffer.from_numpy(A)
    wh, wr = DeviceBuffer(m * n * 2), DeviceBuffer(m * n * 2)
    variant = C.PACK_FP32_16 if a_type == C.FP32_16 else C.PACK_FP32_BF16
    C.call("tomos_pack", variant, m, n, src.ptr, m, wh.ptr, wr.ptr, m, None)
    return wh, wr


def _both(a_type, tier, trans, m, n, seed, y_type):
    rng = np.random.default_rng(seed)
    a, wr = _problem(rng, a_type, m, n)
    rows, inner = (n, m) if trans else (m, n)
    ydt = np.float64 if y_type == C.F64 else np.float32
    x = DeviceBuffer.from_numpy(rng.standard_normal(inner).astype(np.float32))
    y0 = rng.standard_normal(rows).astype(ydt)
    y_ref, y_na = DeviceBuffer.from_numpy(y0), DeviceBuffer.from_numpy(y0)
    need = C.lib().tomos_gemv_scratch_size(a_type, tier, trans, m, n)
    sc = DeviceBuffer(max(need, 1))
    wr_ptr = wr.ptr if wr is not None else None
    C.call("tomos_gemv", a_type, tier, trans, 1.25, m, n, a.ptr, wr_ptr, m, C.F32, x.ptr, -0.5, y_type,
           y_ref.ptr, sc.ptr, need, None)
    C.call("tomos_gemv_noalloc", a_type, tier, trans, 1.25, m, n, a.ptr, wr_ptr, m, C.F32, x.ptr, -0.5,
           y_type, y_na.ptr, None)
    C.call("tomos_device_synchronize")
    return need, y_ref.to_numpy(ydt, (rows,)), y_na.to_numpy(ydt, (rows,))


# (type, tier, trans, first shape, second shape): TRANSPOSE chunk partials in
# FP32 and FP64, and NO_TRANSPOSE column-split partials (short m)
CASES = [
    (C.FP32_BF16, C.TIER_DEFAULT, C.TRANSPOSE, (300001, 1000), (70001, 333)),
    (C.FP32_16, C.TIER_EXACT, C.TRANSPOSE, (100000, 513), (20000, 100)),
    (C.F32, C.TIER_FAST, C.NO_TRANSPOSE, (5000, 700), (3001, 257)),
    (C.FP32_BF16, C.TIER_DEFAULT, C.NO_TRANSPOSE, (9000, 640), (4097, 65)),
]


@pytest.mark.parametrize("a_type,tier,trans,first,second", CASES)
@pytest.mark.parametrize("y_type", [C.F32, C.F64])
def test_noalloc_matches_gemv(capi, a_type, tier, trans, first, second, y_type):
    print("noalloc case type=%d tier=%d trans=%d %s then %s y=%d" % (a_type, tier, trans, first, second, y_type),
          flush=True)
    for seed, (m, n) in enumerate((first, second)):
        need, ref, got = _both(a_type, tier, trans, m, n, 40 + seed, y_type)
        assert need > 0  # every case goes through the scratch array
        assert np.array_equal(ref, got), (m, n, np.max(np.abs(ref - got)))


def test_noalloc_rejects_oversized_scratch(capi):
    # TRANSPOSE with one row and 5e6 columns needs 5e6 FP64 chunk sums
    # (40 MB), above the fixed 32 MiB: rejected before anything is launched
    m, n = 1, 5_000_000
    need = C.lib().tomos_gemv_scratch_size(C.F64, C.TIER_EXACT, C.TRANSPOSE, m, n)
    assert need > (32 << 20)
    a, x, y = DeviceBuffer(8 * n), DeviceBuffer(8), DeviceBuffer(8 * n)
    rc = C.lib().tomos_gemv_noalloc(C.F64, C.TIER_EXACT, C.TRANSPOSE, 1.0, m, n, a.ptr, None, 1, C.F64, x.ptr,
                                    0.0, C.F64, y.ptr, None)
    assert rc == 4  # TOMOS_ERR_WORKSPACE
    assert b"32" in C.lib().tomos_last_error() or b"33554432" in C.lib().tomos_last_error()
