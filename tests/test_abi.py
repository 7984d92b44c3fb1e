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
