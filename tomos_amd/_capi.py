"""ctypes binding of libtomos_hip.so (the C ABI declared in include/tomos_hip.h).

Loading fails loudly when the shared library has not been built: there is
no fallback implementation of any entry point.

The HIP runtime: if PyTorch is importable it is imported before the library
is loaded, so that the process holds a single HIP runtime (torch's
libamdhip64 carries the same soname the library links against).
"""
import ctypes
import os

import numpy as np

try:  # one HIP runtime per process; see module docstring
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is optional plumbing
    torch = None

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtomos_hip.so")

# ---- constants (mirror include/tomos_hip.h) ---------------------------------
F64, F32, F16, FP32_16, BF16, FP32_BF16 = 0, 1, 2, 3, 4, 5
NO_TRANSPOSE, TRANSPOSE = 0, 1
PACK_FP32_16, PACK_FP32_16_CLAMPED, PACK_FP32_16_TRUNCATED = 0, 1, 2
PACK_FP32_BF16, PACK_FP32_BF16_ROUNDED = 3, 4
(DECODE_FP32_16_EXACT, DECODE_FP32_16_HALF, DECODE_FP32_16_HALF_FAST, DECODE_FP32_16_BF16,
 DECODE_FP32_BF16_EXACT, DECODE_FP32_BF16_HI, DECODE_FP32_BF16_EXACT_ROUNDED,
 DECODE_FP32_16_EXACT_TRUNCATED) = range(8)
TIER_DEFAULT, TIER_FAST, TIER_EXACT = 0, 1, 2
TIER_MATERIALIZE = 3  # GEMM only: the materialize slot as a 3-term split on matrix cores
INDEX_INT32, INDEX_INT64 = 0, 1
ERR_NAMES = {1: "INVALID_ARGUMENT", 2: "HIP", 3: "UNSUPPORTED", 4: "WORKSPACE",
             5: "NOT_CONVERGED", 6: "SINGULAR", 7: "INTERNAL"}

i64 = ctypes.c_int64
vp = ctypes.c_void_p
sz = ctypes.c_size_t
dbl = ctypes.c_double
cint = ctypes.c_int


class TomosError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%s)" % (msg, ERR_NAMES.get(code, code)))
        self.code = code


class QrProblem(ctypes.Structure):
    _fields_ = [("k", i64), ("n", i64), ("lda", i64), ("a", vp)]


class LsqProblem(ctypes.Structure):
    _fields_ = [("m", i64), ("n", i64), ("lda", i64), ("a", vp), ("wr", vp), ("b", vp), ("x", vp),
                ("rinv", vp), ("ldr", i64)]


class CsProblem(ctypes.Structure):
    _fields_ = [("m", i64), ("n", i64), ("lda", i64), ("s", i64), ("ldy", i64), ("a", vp), ("wr", vp),
                ("y", vp)]


class LsqrOptions(ctypes.Structure):
    _fields_ = [("a_type", cint), ("mv_tier", cint), ("restart_tier", cint),
                ("value_type", cint), ("refine_type", cint), ("precond_type", cint),
                ("tol", dbl), ("max_relres", dbl), ("max_iters", cint), ("restarts", cint),
                ("check_every", cint), ("fused", cint), ("confirm", cint)]


# name -> (restype, argtypes)
_SIGS = {
    "tomos_last_error": (ctypes.c_char_p, []),
    "tomos_version": (cint, []),
    "tomos_set_device": (cint, [cint]),
    "tomos_get_device_info": (cint, [ctypes.POINTER(cint), ctypes.POINTER(cint),
                                     ctypes.POINTER(cint), ctypes.POINTER(sz)]),
    "tomos_malloc": (cint, [ctypes.POINTER(vp), sz]),
    "tomos_free": (cint, [vp]),
    "tomos_memcpy_htod": (cint, [vp, vp, sz, vp]),
    "tomos_memcpy_dtoh": (cint, [vp, vp, sz, vp]),
    "tomos_memcpy_dtod": (cint, [vp, vp, sz, vp]),
    "tomos_memset": (cint, [vp, cint, sz, vp]),
    "tomos_stream_synchronize": (cint, [vp]),
    "tomos_device_synchronize": (cint, []),
    "tomos_stream_create": (cint, [ctypes.POINTER(vp)]),
    "tomos_stream_destroy": (cint, [vp]),
    "tomos_event_create": (cint, [ctypes.POINTER(vp)]),
    "tomos_event_destroy": (cint, [vp]),
    "tomos_event_record": (cint, [vp, vp]),
    "tomos_event_elapsed_ms": (cint, [ctypes.POINTER(ctypes.c_float), vp, vp]),
    "tomos_copy_kernel": (cint, [vp, vp, sz, vp]),
    "tomos_read_kernel": (cint, [vp, sz, vp, vp]),
    "tomos_fill_normal_f32": (cint, [vp, i64, i64, i64, ctypes.c_uint64, vp, cint, cint, vp]),
    "tomos_convert": (cint, [cint, cint, i64, i64, vp, i64, vp, i64, vp]),
    "tomos_pack": (cint, [cint, i64, i64, vp, i64, vp, vp, i64, vp]),
    "tomos_unpack": (cint, [cint, i64, i64, vp, vp, i64, vp, i64, vp]),
    "tomos_decode_wh": (cint, [cint, i64, vp, vp, vp]),
    "tomos_gemv_scratch_size": (sz, [cint, cint, cint, i64, i64]),
    "tomos_gemv": (cint, [cint, cint, cint, dbl, i64, i64, vp, vp, i64, cint, vp, dbl, cint,
                          vp, vp, sz, vp]),
    "tomos_gemv_noalloc": (cint, [cint, cint, cint, dbl, i64, i64, vp, vp, i64, cint, vp, dbl,
                                  cint, vp, vp]),
    "tomos_gemm": (cint, [cint, cint, cint, cint, dbl, i64, i64, i64, vp, vp, i64, vp, i64,
                          dbl, vp, i64, vp]),
    "tomos_gemm_split_scratch_size": (sz, [i64, i64, i64]),
    "tomos_matvec_pair_scratch_size": (sz, [cint, i64, i64]),
    "tomos_fused_config_set": (cint, [cint, i64, cint, cint, cint]),
    "tomos_fused_config_clear": (cint, []),
    "tomos_stream_loads_set": (cint, [cint]),
    "tomos_fused_config_get": (cint, [cint, i64, ctypes.POINTER(cint), ctypes.POINTER(cint), ctypes.POINTER(cint)]),
    "tomos_lsq_adjoint_scratch_size": (sz, [ctypes.POINTER(LsqrOptions), i64, i64]),
    "tomos_lsq_adjoint_solve": (cint, [ctypes.POINTER(LsqrOptions), i64, i64, vp, vp, i64, vp, vp, i64,
                                       vp, vp, sz, ctypes.POINTER(cint), vp, cint, ctypes.POINTER(cint),
                                       vp]),
    "tomos_lsq_grad_scratch_size": (sz, [cint, cint, i64, i64]),
    "tomos_sketch_qr_precond_batched_scratch_size": (sz, [cint, vp]),
    "tomos_lsqr_batched_scratch_size": (sz, [ctypes.POINTER(LsqrOptions), cint, vp, cint, cint, cint]),
    "tomos_lsqr_batched_solve": (cint, [ctypes.POINTER(LsqrOptions), cint, vp, cint, cint, cint, cint, vp,
                                        sz, vp, vp, vp]),
    "tomos_sketch_qr_precond_batched_f64": (cint, [cint, vp, vp, sz, vp]),
    "tomos_lsq_grad": (cint, [cint, cint, i64, i64, vp, vp, i64, vp, vp, vp, vp, vp, i64, vp, sz, vp]),
    "tomos_lsq_adjoint_batched_scratch_size": (sz, [ctypes.POINTER(LsqrOptions), cint, vp, cint]),
    "tomos_lsq_adjoint_batched": (cint, [ctypes.POINTER(LsqrOptions), cint, vp, vp, vp, vp, vp, vp, cint, vp,
                                         sz, vp, vp]),
    "tomos_matvec_pair": (cint, [cint, i64, i64, vp, i64, vp, dbl, vp, vp, vp, vp, sz, vp]),
    "tomos_kernel_timer_begin": (cint, [cint]),
    "tomos_kernel_timer_end": (cint, [vp, cint, vp]),
    "tomos_gemm_split": (cint, [cint, i64, i64, i64, dbl, vp, vp, i64, vp, vp, i64, dbl, vp, i64, vp,
                                sz, vp]),
    "tomos_gaussian_fill": (cint, [cint, i64, i64, vp, i64, ctypes.c_uint64, vp]),
    "tomos_csr_check": (cint, [cint, cint, i64, i64, i64, vp, vp, vp]),
    "tomos_csr_sketch": (cint, [cint, cint, cint, i64, i64, vp, vp, vp, ctypes.c_uint64, vp]),
    "tomos_csr_mv_scratch_size": (sz, [cint, cint, cint, i64, i64, i64]),
    "tomos_csr_mv": (cint, [cint, cint, cint, ctypes.c_double, i64, i64, i64, vp, vp, vp, cint, vp,
                            ctypes.c_double, cint, vp, vp, sz, vp]),
    "tomos_csr_mm": (cint, [cint, cint, ctypes.c_double, i64, i64, i64, vp, vp, vp, i64, cint, vp, i64,
                            ctypes.c_double, cint, vp, i64, vp]),
    "tomos_csr_transpose_scratch_size": (sz, [cint, cint, i64, i64, i64]),
    "tomos_csr_transpose": (cint, [cint, cint, i64, i64, i64, vp, vp, vp, vp, vp, vp, vp, sz, vp]),
    "tomos_countsketch_indices": (cint, [i64, i64, vp, vp, vp]),
    "tomos_countsketch_scratch_size": (sz, [i64, i64, i64]),
    "tomos_countsketch_apply": (cint, [cint, cint, i64, i64, vp, vp, i64, i64, vp, i64, vp, sz,
                                       vp]),
    "tomos_countsketch_apply_batched_scratch_size": (sz, [cint, vp]),
    "tomos_countsketch_apply_batched": (cint, [cint, cint, cint, vp, vp, sz, vp]),
    "tomos_geqrf_scratch_size": (sz, [i64, i64]),
    "tomos_geqrf_r_f64": (cint, [i64, i64, vp, i64, vp, i64, vp, sz, vp]),
    "tomos_trtri_upper_f64": (cint, [i64, vp, i64, vp]),
    "tomos_qr_panels_set": (cint, [cint]),
    "tomos_sketch_qr_precond_f64": (cint, [i64, i64, vp, i64, vp]),
    "tomos_lsqr_arena_size": (sz, [ctypes.POINTER(LsqrOptions), i64, i64]),
    "tomos_lsqr_scratch_size": (sz, [ctypes.POINTER(LsqrOptions), i64, i64]),
    "tomos_lsqr_solve": (cint, [ctypes.POINTER(LsqrOptions), i64, i64, vp, vp, i64, vp, vp, vp,
                                i64, vp, sz, vp, sz, ctypes.POINTER(cint), vp, cint,
                                ctypes.POINTER(cint), vp]),
    "tomos_lsqr_solve_refined": (cint, [ctypes.POINTER(LsqrOptions), i64, i64, vp, vp, i64, cint, vp, i64, vp,
                                        vp, vp, i64, vp, sz, vp, sz, ctypes.POINTER(cint), vp, cint,
                                        ctypes.POINTER(cint), vp]),
    "tomos_lsqr_csr_scratch_size": (sz, [ctypes.POINTER(LsqrOptions), cint, i64, i64, i64]),
    "tomos_lsqr_solve_csr": (cint, [ctypes.POINTER(LsqrOptions), cint, i64, i64, i64, vp, vp, vp, cint, vp, vp,
                                    vp, vp, i64, vp, sz, vp, sz, ctypes.POINTER(cint), vp, cint,
                                    ctypes.POINTER(cint), vp]),
}

_lib = None


def header_symbols():
    """Function names declared in include/tomos_hip.h."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "tomos_hip.h")
    text = open(hdr).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tomos_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libtomos_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:  # export completeness is checked by the test suite
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().tomos_last_error()
        raise TomosError(rc, msg.decode() if msg else "error")
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))


# ---- device buffers ----------------------------------------------------------
class DeviceBuffer:
    """Owning device allocation (hipMalloc via the C ABI)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = vp()
        call("tomos_malloc", ctypes.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    def free(self):
        if self.ptr:
            lib().tomos_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @classmethod
    def from_numpy(cls, arr):
        arr = np.asarray(arr)
        if not (arr.flags.f_contiguous or arr.flags.c_contiguous):
            arr = np.asfortranarray(arr)
        b = cls(arr.nbytes)
        if arr.nbytes:
            call("tomos_memcpy_htod", b.ptr, arr.ctypes.data, arr.nbytes, None)
        return b

    def to_numpy(self, dtype, shape, order="F"):
        out = np.empty(shape, dtype=dtype, order=order)
        if out.nbytes:
            assert out.nbytes <= self.nbytes
            call("tomos_memcpy_dtoh", out.ctypes.data, self.ptr, out.nbytes, None)
        return out

    def zero(self):
        call("tomos_memset", self.ptr, 0, self.nbytes, None)
        return self


def synchronize():
    call("tomos_device_synchronize")
