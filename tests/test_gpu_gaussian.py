"""GPU dense Gaussian sketch (tomos_gaussian_fill) against the oracle's
seeded counter layout, plus the reference's contract itself: i.i.d. N(0,1)
entries (create_gaussian_sketch(Dense&), sketch.hpp:11-13,
gaussian_sketch_registries.cpp:16-30) in FP64, FP32, FP16 and BF16.

Tolerance: the kernel runs Box-Muller in FP32 (logf, sqrtf, sincospif),
the oracle in FP64 with libm: |gpu - oracle| <= 1e-6 (1 + |x|) for
FP32/FP64 output, a few FP32 rounding errors; the 16-bit types are within
one unit in the last place of the rounded oracle value (RNE of a value that
differs by ~1e-7 can land on the neighbouring 16-bit number)."""
import ctypes
import math

import numpy as np
import pytest
from scipy import stats

from tomos_amd import _capi as C
from tomos_amd._capi import DeviceBuffer

pytestmark = pytest.mark.gpu

ESZ = {C.F64: 8, C.F32: 4, C.F16: 2, C.BF16: 2}


def fill(t, rows, cols, ld, seed, offset_elems=0, sentinel=0xA5):
    buf = DeviceBuffer((ld * cols + offset_elems) * ESZ[t] + 16)
    host = np.full(buf.nbytes, sentinel, np.uint8)
    C.call("tomos_memcpy_htod", buf.ptr, host.ctypes.data, buf.nbytes, None)
    ptr = buf.ptr + offset_elems * ESZ[t]
    C.call("tomos_gaussian_fill", t, rows, cols, ptr, ld, seed, None)
    C.call("tomos_device_synchronize")
    raw = buf.to_numpy(np.uint8, (buf.nbytes,))
    body = raw[offset_elems * ESZ[t]: offset_elems * ESZ[t] + ld * cols * ESZ[t]]
    dt = {C.F64: np.float64, C.F32: np.float32, C.F16: np.uint16, C.BF16: np.uint16}[t]
    M = body.view(dt).reshape((ld, cols), order="F")
    # nothing outside the rows x cols block was written
    pad = np.ones(raw.size, bool)
    for j in range(cols):
        s0 = (offset_elems + j * ld) * ESZ[t]
        pad[s0: s0 + rows * ESZ[t]] = False
    assert np.all(raw[pad] == sentinel)
    return M[:rows]


def to_f64(t, M):
    if t == C.F16:
        return M.view(np.float16).astype(np.float64)
    if t == C.BF16:
        return (M.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return M.astype(np.float64)


def ulp16(t, x):
    x = np.abs(x)
    if t == C.F16:
        e = np.floor(np.log2(np.maximum(x, 2.0 ** -14)))
        return 2.0 ** (e - 10)
    e = np.floor(np.log2(np.maximum(x, 2.0 ** -126)))
    return 2.0 ** (e - 7)


@pytest.mark.parametrize("t", [C.F64, C.F32, C.F16, C.BF16])
@pytest.mark.parametrize("rows,cols,ld,off", [
    (1000, 37, 1000, 0),      # even ld, aligned: paired stores
    (999, 5, 1001, 0),        # odd rows and ld: element stores, last pair half-filled
    (64, 3, 64, 1),           # misaligned base: element stores
    (1, 7, 1, 0),             # one row
    (3, 70000, 4, 0),         # more columns than grid.y: the column loop
])
def test_matches_oracle(capi, orc, t, rows, cols, ld, off):
    seed = 0x1234ABCD5678EF01
    got = to_f64(t, fill(t, rows, cols, ld, seed, off))
    ref = orc.gaussian_fill(rows, cols, seed)
    if t in (C.F64, C.F32):
        err = np.abs(got - ref) / (1 + np.abs(ref))
        assert err.max() <= 1e-6, err.max()
    else:
        assert np.all(np.abs(got - ref) <= ulp16(t, ref) + 1e-6 * (1 + np.abs(ref)))


def test_deterministic_and_seeded(capi):
    a = fill(C.F32, 513, 9, 514, 5)
    assert np.array_equal(a, fill(C.F32, 513, 9, 514, 5))
    # the values depend on (seed, rows, cols) only, not on ld or alignment
    assert np.array_equal(a, fill(C.F32, 513, 9, 515, 5, offset_elems=1))
    assert np.mean(a == fill(C.F32, 513, 9, 514, 6)) < 1e-3


def test_distribution_and_sketch_property(capi):
    rows, cols = 4096, 1024  # 4M samples
    v = fill(C.F32, rows, cols, rows, 2024).astype(np.float64)
    x = v.ravel(order="F")
    N = x.size
    assert abs(x.mean()) < 5 / math.sqrt(N)
    assert abs(x.var() - 1) < 5 * math.sqrt(2 / N)
    assert abs(stats.kurtosis(x)) < 5 * math.sqrt(24 / N)
    assert stats.kstest(x[:1000000], "norm").pvalue > 1e-4
    lim = 5 / math.sqrt(x.size / 2)
    assert abs(np.corrcoef(v[0::2].ravel(), v[1::2].ravel())[0, 1]) < lim
    assert abs(np.corrcoef(v[0::2].ravel() ** 2, v[1::2].ravel() ** 2)[0, 1]) < lim
    # Johnson-Lindenstrauss: S / sqrt(k) preserves norms in expectation
    k, m, trials = 256, 4096, 200
    S = fill(C.F64, k, m, k, 99) / math.sqrt(k)
    X = np.random.default_rng(3).standard_normal((m, trials))
    ratio = np.linalg.norm(S @ X, axis=0) ** 2 / np.linalg.norm(X, axis=0) ** 2
    assert abs(ratio.mean() - 1) < 0.03 and ratio.std() < 0.15


def test_errors(capi):
    buf = DeviceBuffer(64)
    with pytest.raises(C.TomosError, match="ld < rows"):
        C.call("tomos_gaussian_fill", C.F32, 8, 1, buf.ptr, 7, 1, None)
    with pytest.raises(C.TomosError, match="F64, F32, F16 or BF16"):
        C.call("tomos_gaussian_fill", C.FP32_16, 2, 2, buf.ptr, 2, 1, None)
    C.call("tomos_gaussian_fill", C.F32, 0, 5, None, 1, 1, None)  # empty: no-op
