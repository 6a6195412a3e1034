"""ctypes binding of libscarl_hip.so (declared in include/scarl_hip.h).

The product path: every call goes to the HIP library.  There is no CPU fallback -- if the
library is missing, importing anything that needs it raises ``ScarlLibraryMissing``.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscarl_hip.so")

SCARL_OK, SCARL_E_INVALID, SCARL_E_HIP, SCARL_E_UNSUPPORTED, SCARL_E_NOMEM = 0, 1, 2, 3, 4
CM_GAUSSIAN, CM_VON_KARMAN, CM_EXPONENTIAL = 1, 2, 3
FOM_NONE, FOM_GAUSSIAN, FOM_LOGNORMAL = 0, 1, 2
PREC_FP32, PREC_FP64 = 32, 64


class ScarlLibraryMissing(RuntimeError):
    pass


class ScarlError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"scarl error {code}: {msg}")
        self.code = code


class FieldDesc(C.Structure):
    """Mirror of ``scarl_field_desc``."""
    _fields_ = [
        ("n", C.c_int32 * 3),
        ("dx", C.c_double * 3),
        ("model", C.c_int32),
        ("corrL", C.c_double * 3),
        ("nu", C.c_double),
        ("sigma", C.c_double),
        ("margi", C.c_int32),
        ("avg", C.c_double),
        ("std_dev", C.c_double),
        ("seed", C.c_uint64),
        ("stream", C.c_uint32),
        ("normalize", C.c_int32),
    ]


# scarl_hip_set_tuning keys
TUNE_FP32_ELEMS, TUNE_MFMA_DFT16, TUNE_AUTOTUNE, TUNE_LAG_FFT, TUNE_SCHEDULE = 1, 2, 3, 4, 5

# symbol -> (restype, argtypes); the list is the complete ABI of include/scarl_hip.h
_vp, _i32, _i64, _d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
_pi32, _pd = C.POINTER(C.c_int32), C.POINTER(C.c_double)
SIGNATURES = {
    "scarl_hip_init": [_i32],
    "scarl_hip_finalize": [],
    "scarl_hip_last_error": [C.c_char_p, _i32],
    "scarl_hip_set_timing": [_i32],
    "scarl_hip_timing_report": [C.c_char_p, _i32],
    "scarl_hip_timing_reset": [],
    "scarl_hip_get_stream": [C.POINTER(C.c_void_p)],
    "scarl_hip_set_tuning": [_i32, _i32],
    "scarl_hip_gen_std_gauss_fft": [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _vp, _i32],
    "scarl_hip_gen_std_gauss_fft_mpi": [_vp, _vp, _vp, _i32, _i32, _i32, _vp, _vp, _i32, _vp, _i32],
    "scarl_hip_gen_std_gauss_fft_dev": [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _vp, _i32],
    "scarl_hip_gen_std_gauss_fft_dev_noalloc": [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _vp, _i32],
    "scarl_hip_set_marginal": [_i32, _d, _d],
    "scarl_hip_build_sk_query": [_i32, _vp, _vp, _vp, _vp, _i32, _i32, _vp],
    "scarl_hip_build_sk": [_i32, _vp, _vp, _vp, _vp, _i32, _i32, _vp],
    "scarl_hip_gen_std_gauss_shino_fft": [_vp, _vp, _vp, _vp, _vp, _i32, _i32, _i32],
    "scarl_hip_r8_uniform_01": [_pi32, _pd],
    "scarl_hip_r8vec_uniform_01": [_i32, _pi32, _vp],
    "scarl_hip_r8vec_normal_01": [_i32, _pi32, _vp],
    "scarl_hip_apply_unity_partition": [_vp, _vp, _vp, _i32, _vp, _vp],
    "scarl_hip_multivariate_transformation": [_d, _d, _i32, _vp, _i64],
    "scarl_hip_normalize_field": [_vp, _vp, _vp, _vp, _vp, _pd, _pd],
    "scarl_hip_moments": [_vp, _i64, _i32, _vp, _vp, _pd, _pd],
    "scarl_hip_rebuild_sk_fft": [_vp, _vp, _i32, _i32, _d, _d, _vp],
    "scarl_hip_rebuild_corrl": [_vp, _vp, _i32, _vp, _vp],
    "scarl_hip_lag_covariance": [_vp, _vp, _d, _i32, _vp],
    "scarl_hip_kv": [_d, _vp, _i32, _vp],
    "scarl_hip_target_corr": [_i32, _d, _vp, _i32, _vp],
    "scarl_hip_batch_create": [C.POINTER(FieldDesc), _i32, _i32, C.POINTER(C.c_void_p)],
    "scarl_hip_blend_create": [_vp, _vp, _vp, _vp, C.POINTER(FieldDesc), _i32, C.POINTER(C.c_void_p)],
    "scarl_hip_batch_set_noise": [_vp, _i32, _vp],
    "scarl_hip_batch_run": [_vp, _i32],
    "scarl_hip_batch_sync": [_vp],
    "scarl_hip_batch_output": [_vp, _i32, C.POINTER(C.c_void_p), _vp],
    "scarl_hip_batch_copy_output": [_vp, _i32, _vp, _i64],
    "scarl_hip_batch_moments": [_vp, _vp, _i32, _pi32],
    "scarl_hip_batch_lag_cov": [_vp, _i32, _vp],
    "scarl_hip_batch_field_info": [_vp, _i32, _pd, _pd],
    "scarl_hip_batch_plan_report": [_vp, C.c_char_p, _i32],
    "scarl_hip_batch_bytes": [_vp, C.POINTER(C.c_int64)],
    "scarl_hip_batch_destroy": [_vp],
    "scarl_hip_gen_field_h": [_vp, C.POINTER(FieldDesc), _i32, _vp],
}

_lib = None


def lib():
    """Load libscarl_hip.so (built in-tree); raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ScarlLibraryMissing(
                f"{LIB_PATH} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().scarl_hip_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int):
    if rc != SCARL_OK:
        raise ScarlError(rc, last_error())


def call(name, *args):
    check(getattr(lib(), name)(*args))


class _ArrayArg:
    """A numpy array passed as a pointer argument; holds a reference to the array so that a
    temporary stays alive until the C call that receives it returns."""
    __slots__ = ("array", "_as_parameter_")

    def __init__(self, a):
        self.array = a
        self._as_parameter_ = C.c_void_p(a.ctypes.data)


def ptr(a):
    """Pointer argument for a numpy array (None passes NULL)."""
    if a is None:
        return None
    return _ArrayArg(a)
