"""ctypes binding of libscarl_h5.so (declared in include/scarl_h5.h): HDF5 and XDMF output.

The library is host code over the HDF5 1.10 C API. It is built next to libscarl_hip.so by the
same build step. If it is missing, every call raises ``ScarlLibraryMissing``.
"""
from __future__ import annotations

import ctypes as C
import os

from ._capi import ScarlLibraryMissing

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscarl_h5.so")

H5_OK, H5_E_INVALID, H5_E_IO = 0, 1, 2
LABEL_LEN = 20

_vp, _i32, _i64, _d, _s = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_char_p
_pi32, _pd = C.POINTER(C.c_int32), C.POINTER(C.c_double)
# symbol -> argtypes; the list is the complete ABI of include/scarl_h5.h
SIGNATURES = {
    "scarl_h5_last_error": [_s, _i32],
    "scarl_h5_create": [_s, C.POINTER(_vp)],
    "scarl_h5_open": [_s, _i32, C.POINTER(_vp)],
    "scarl_h5_close": [_vp],
    "scarl_h5_create_dataset": [_vp, _s, _pi32, _i32],
    "scarl_h5_write_block": [_vp, _s, _vp, _i32, _i64, _i64, _pi32, _pi32],
    "scarl_h5_attr_i32": [_vp, _s, _s, _pi32, _i32],
    "scarl_h5_attr_f64": [_vp, _s, _s, _pd, _i32],
    "scarl_h5_write_times": [_vp, _s, _pd, _i32, _i32],
    "scarl_h5_dataset_info": [_vp, _s, _pi32, _pi32],
    "scarl_h5_read_block": [_vp, _s, _vp, _i32, _i64, _i64, _pi32, _pi32],
    "scarl_h5_read_attr_i32": [_vp, _s, _s, _pi32, _i32, _pi32],
    "scarl_h5_read_attr_f64": [_vp, _s, _s, _pd, _i32, _pi32],
    "scarl_h5_attr_count": [_vp, _s, _pi32],
    "scarl_h5_attr_name": [_vp, _s, _i32, _s, _i32],
    "scarl_h5_read_times": [_vp, _pd, _i32, _s, _i32, _pi32, _pi32],
    "scarl_xmf_write_elements": [_s, _s, _s, _s, _s, _pd, _pd, _pi32, _i32],
    "scarl_xmf_write_global": [_s, C.POINTER(_s), _pd, _pd, _pi32, _i32],
    "scarl_fortran_fixed": [_d, _i32, _i32, _s, _i32],
}

_lib = None


class H5Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"scarl_h5 error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ScarlLibraryMissing(
                f"{LIB_PATH} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, argtypes in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = C.c_int32
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().scarl_h5_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != H5_OK:
        raise H5Error(rc, last_error())
    return rc


def b(s) -> bytes:
    return s.encode() if isinstance(s, str) else bytes(s)
