"""tomos_matvec_pair: u <- A p + c u, t <- A^T u, |u|^2 from one read of A,
vs FP64 products of the decoded (hi-only) matrix, over every launch shape
the tuner can pick (row groups 2/4/8, blocks per CU)."""
import ctypes

import numpy as np
import pytest

from tests import hipcall as H
from tomos_amd import _capi as C
from tomos_amd._capi import DeviceBuffer

pytestmark = pytest.mark.gpu

HI_DECODER = {C.FP32_BF16: C.DECODE_FP32_BF16_HI, C.FP32_16: C.DECODE_FP32_16_HALF}
PACKER = {C.FP32_BF16: C.PACK_FP32_BF16, C.FP32_16: C.PACK_FP32_16}
MAX_N = {(8, 256): 1024, (4, 256): 1024, (2, 256): 2048, (0, 0): 2048, (8, 512): 1024, (4, 512): 2048,
         (32, 1024): 2048}


def run_pair(a_type, A, lda, p, u0, coef):
    m, n = A.shape
    if a_type == C.F32:
        Ap = np.zeros((lda, n), np.float32, order="F")
        Ap[:m] = A
        abuf, keep = H.dev(Ap), None
    else:
        abuf, keep = H.pack(PACKER[a_type], A, ld=lda)
    pd, ud = H.dev(p), H.dev(u0)
    td, nd = DeviceBuffer(n * 8), DeviceBuffer(8)
    C.call("tomos_matvec_pair", a_type, m, n, abuf.ptr, lda, pd.ptr, coef, ud.ptr, td.ptr, nd.ptr,
           None, 0, None)
    return ud.to_numpy(np.float32, (m,)), td.to_numpy(np.float64, (n,)), nd.to_numpy(np.float64, (1,))[0]


def used_matrix(orc, a_type, A):
    if a_type == C.F32:
        return A.astype(np.float64)
    eh, er = orc.pack(PACKER[a_type], A)
    return orc.unpack(HI_DECODER[a_type], eh, er).astype(np.float64)


@pytest.mark.parametrize("a_type", [C.F32, C.FP32_BF16, C.FP32_16])
@pytest.mark.parametrize("m,lda,n", [(1000, 1000, 37), (4099, 4104, 1000), (777, 784, 64),
                                      (20000, 20000, 1500), (3001, 3008, 2048), (64, 64, 512)])
@pytest.mark.parametrize("groups,threads,bpc", [(0, 0, 0), (2, 256, 0), (4, 256, 1), (8, 256, 0), (4, 256, 3),
                                               (8, 512, 0), (4, 512, 2), (32, 1024, 0), (32, 1024, 2)])
def test_pair_vs_fp64(capi, orc, a_type, m, lda, n, groups, threads, bpc):
    if n > MAX_N[(groups, threads)]:
        pytest.skip("no kernel for %d row groups x %d threads at n = %d" % (groups, threads, n))
    rng = np.random.default_rng(m + n + groups)
    A = (rng.standard_normal((m, n)) * 2.0 ** rng.integers(-8, 9, (m, n))).astype(np.float32, order="F")
    p = rng.standard_normal(n)
    u0 = rng.standard_normal(m).astype(np.float32)
    coef = -0.37
    C.call("tomos_fused_config_set", a_type, n, groups, bpc, threads)
    try:
        u, t, nrm2 = run_pair(a_type, A, lda, p, u0, coef)
    finally:
        C.call("tomos_fused_config_clear")
    Au = used_matrix(orc, a_type, A)
    u_ref = Au @ p.astype(np.float32).astype(np.float64) + coef * u0.astype(np.float64)
    # row sums accumulate n products in FP32: bound relative to sum |a_ij p_j|
    scale = np.abs(Au) @ np.abs(p) + np.abs(coef * u0)
    assert np.max(np.abs(u - u_ref) / scale) < 4e-6
    ud = u.astype(np.float64)
    t_ref = Au.T @ ud
    tscale = np.abs(Au.T) @ np.abs(ud)
    assert np.max(np.abs(t - t_ref) / tscale) < 4e-6
    assert abs(nrm2 - ud @ ud) <= 1e-12 * (ud @ ud)


def test_pair_deterministic_and_errors(capi):
    rng = np.random.default_rng(1)
    m, n = 50000, 700
    A = rng.standard_normal((m, n)).astype(np.float32, order="F")
    p = rng.standard_normal(n)
    u0 = rng.standard_normal(m).astype(np.float32)
    r1 = run_pair(C.FP32_BF16, A, m, p, u0, 0.5)
    r2 = run_pair(C.FP32_BF16, A, m, p, u0, 0.5)
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1]) and r1[2] == r2[2]
    # unsupported: lda not a multiple of the 16-byte vector, n too large, bad config
    wh, _ = H.pack(C.PACK_FP32_BF16, A[:, :10], ld=m + 1)
    pd, ud, td = H.dev(p[:10]), H.dev(u0), DeviceBuffer(80)
    rc = C.lib().tomos_matvec_pair(C.FP32_BF16, m, 10, ctypes.c_void_p(wh.ptr), m + 1, ctypes.c_void_p(pd.ptr),
                                   0.0, ctypes.c_void_p(ud.ptr), ctypes.c_void_p(td.ptr), None, None, 0, None)
    assert rc == 3  # TOMOS_ERR_UNSUPPORTED
    assert C.lib().tomos_fused_config_set(C.FP32_BF16, 2000, 8, 0, 256) == 1  # G=8 x 256 covers n <= 1024
    assert C.lib().tomos_fused_config_set(C.FP32_BF16, 100, 3, 0, 256) == 1
    assert C.lib().tomos_fused_config_set(C.FP32_BF16, 100, 4, 0, 384) == 1
    assert C.lib().tomos_fused_config_set(C.F64, 100, 4, 0, 256) == 1


def test_tuning_one_format_keeps_the_other(capi):
    """tune_fused installs its best shape and leaves other formats' entries."""
    from tomos_amd import autotune
    rng = np.random.default_rng(3)
    m, n = 20000, 100
    A = rng.standard_normal((m, n)).astype(np.float32, order="F")
    wh, _ = H.pack(C.PACK_FP32_BF16, A)
    af = H.dev(A)
    C.call("tomos_fused_config_clear")
    try:
        best_b, res_b = autotune.tune_fused(C.FP32_BF16, m, n, wh.ptr, m, reps=2,
                                            cands=[(8, 256, 0), (4, 512, 0)])
        best_f, _ = autotune.tune_fused(C.F32, m, n, af.ptr, m, reps=2, cands=[(2, 256, 0), (8, 256, 1)])
        assert len(res_b) == 2
        got = autotune.installed(C.FP32_BF16, n)
        assert got == {k: best_b[k] for k in ("row_groups", "blocks_per_cu", "threads")}
        got = autotune.installed(C.F32, n)
        assert got == {k: best_f[k] for k in ("row_groups", "blocks_per_cu", "threads")}
    finally:
        C.call("tomos_fused_config_clear")


@pytest.mark.parametrize("a_type", [C.F32, C.FP32_BF16])
def test_load_policy_does_not_change_results(capi, a_type):
    """Non-temporal and cacheable loads give bit-identical products (pair and GEMV)."""
    rng = np.random.default_rng(5)
    m, n = 30000, 300
    A = rng.standard_normal((m, n)).astype(np.float32, order="F")
    p = rng.standard_normal(n)
    u0 = rng.standard_normal(m).astype(np.float32)
    outs = []
    try:
        for mode in (1, 2):
            C.call("tomos_stream_loads_set", mode)
            outs.append(run_pair(a_type, A, m, p, u0, 0.25))
    finally:
        C.call("tomos_stream_loads_set", 0)
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert C.lib().tomos_stream_loads_set(3) == 1


def test_kernel_timer_counts_single_pass_launches(capi):
    """tomos_kernel_timer_begin/_end bracket each single-pass kernel launch
    with HIP events: the count equals the launches made (capped at the
    capacity), durations are positive, and the timer leaves the
    products unchanged."""
    m, n = 20000, 300
    rng = np.random.default_rng(5)
    A = rng.standard_normal((m, n)).astype(np.float32, order="F")
    p = rng.standard_normal(n)
    u0 = rng.standard_normal(m).astype(np.float32)
    u_ref, t_ref, _ = run_pair(C.F32, A, m, p, u0, 0.5)
    C.call("tomos_kernel_timer_begin", 3)
    outs = [run_pair(C.F32, A, m, p, u0, 0.5) for _ in range(5)]
    cnt = ctypes.c_int()
    dur = np.zeros(8, np.float32)
    C.call("tomos_kernel_timer_end", dur.ctypes.data, 8, ctypes.byref(cnt))
    assert cnt.value == 3
    assert np.all(dur[:3] > 0.0) and np.all(dur[3:] == 0.0)
    for u, t, _ in outs:
        assert np.array_equal(u, u_ref) and np.array_equal(t, t_ref)
    # no timer running: end is an error, begin rejects a bad capacity
    assert C.lib().tomos_kernel_timer_end(dur.ctypes.data, 8, ctypes.byref(cnt)) != 0
    assert C.lib().tomos_kernel_timer_begin(0) != 0
