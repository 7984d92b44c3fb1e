"""GPU encoders on every fp32 bit pattern: all 2^32 inputs, in chunks of
2^27, through tomos_pack for each of the five encoders, compared element by
element with the oracle's encoder (orc_pack_compare_range, OpenMP on the
host).  Bit-exact: zero mismatches allowed."""
import ctypes

import numpy as np
import pytest

from tomos_amd import _capi as C
from tomos_amd._capi import DeviceBuffer

pytestmark = pytest.mark.gpu

CHUNK = 1 << 27
ENCODERS = [(C.PACK_FP32_16, 0), (C.PACK_FP32_16_CLAMPED, 1), (C.PACK_FP32_16_TRUNCATED, 2),
            (C.PACK_FP32_BF16, 3), (C.PACK_FP32_BF16_ROUNDED, 4)]


def test_pack_all_2_32_patterns(capi, orc):
    src = DeviceBuffer(CHUNK * 4)
    wh, wr = DeviceBuffer(CHUNK * 2), DeviceBuffer(CHUNK * 2)
    hh = np.empty(CHUNK, np.uint16)
    hr = np.empty(CHUNK, np.uint16)
    L = orc.lib()
    fb = ctypes.c_int64(0)
    total_bad = {}
    for c in range((1 << 32) // CHUNK):
        start = c * CHUNK
        bits = np.arange(start, start + CHUNK, dtype=np.uint64).astype(np.uint32)
        C.call("tomos_memcpy_htod", src.ptr, bits.ctypes.data, CHUNK * 4, None)
        for variant, ov in ENCODERS:
            C.call("tomos_pack", variant, CHUNK, 1, src.ptr, CHUNK, wh.ptr, wr.ptr, CHUNK, None)
            C.call("tomos_memcpy_dtoh", hh.ctypes.data, wh.ptr, CHUNK * 2, None)
            C.call("tomos_memcpy_dtoh", hr.ctypes.data, wr.ptr, CHUNK * 2, None)
            bad = L.orc_pack_compare_range(ov, start, CHUNK, hh.ctypes.data_as(orc.u16p),
                                           hr.ctypes.data_as(orc.u16p), ctypes.byref(fb))
            assert bad == 0, "encoder %d: %d mismatches from pattern 0x%08x" % (variant, bad, fb.value)
            total_bad[variant] = total_bad.get(variant, 0) + bad
        if c % 8 == 7:
            print("pack exhaustive: patterns below 0x%09x checked" % (start + CHUNK), flush=True)
    assert all(v == 0 for v in total_bad.values())
