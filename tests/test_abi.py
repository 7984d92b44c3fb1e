"""C-ABI library: loads and exports every symbol include/tomos_hip.h declares
(no compute calls; runs without a GPU)."""
import ctypes
import re

from tomos_amd import _capi


def test_library_exports_every_header_symbol():
    L = ctypes.CDLL(_capi.LIB_PATH)
    declared = _capi.header_symbols()
    assert len(declared) > 20
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_bindings_cover_header():
    assert sorted(_capi._SIGS) == _capi.header_symbols()


def test_library_is_gfx950_only():
    """The embedded offload bundle names only the gfx950 target."""
    blob = open(_capi.LIB_PATH, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-z]+)", blob))
    assert targets == {b"gfx950"}, targets


def test_last_error_message_on_bad_argument():
    L = _capi.lib()
    rc = L.tomos_pack(99, 4, 4, None, 4, None, None, 4, None)
    assert rc == 1
    assert b"unknown variant" in L.tomos_last_error() or b"null" in L.tomos_last_error()


def test_adjoint_batched_rejects_bad_arguments():
    """Argument checks run before any device work (no GPU needed)."""
    L = _capi.lib()
    opt = _capi.LsqrOptions(_capi.FP32_BF16, _capi.TIER_DEFAULT, _capi.TIER_EXACT, _capi.F32, _capi.F64,
                            _capi.F64, 1e-6, 1e-6, 30, 20, 8, 1)
    prob = (_capi.LsqProblem * 1)(_capi.LsqProblem(100, 10, 100, None, None, None, None, None, 40))
    for workers in (0, 17):
        assert L.tomos_lsq_adjoint_batched(ctypes.byref(opt), 1, prob, None, None, None, None, None, workers,
                                           None, 0, None, None) != 0
        assert L.tomos_lsq_adjoint_batched_scratch_size(ctypes.byref(opt), 1, prob, workers) == 0
    # null pointers of a problem are rejected, an empty batch is a no-op
    assert L.tomos_lsq_adjoint_batched(ctypes.byref(opt), 1, prob, None, None, None, None, None, 2,
                                       None, 0, None, None) != 0
    assert L.tomos_lsq_adjoint_batched(ctypes.byref(opt), 0, None, None, None, None, None, None, 2,
                                       None, 0, None, None) == 0
