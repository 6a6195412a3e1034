"""Drop-in for the reference module ``normal_ScaRL`` (SRC_F/normal_ScaRL.f90).

The Park-Miller stream is produced on the GPU by modular jump-ahead (seed_i = seed*16807^i mod
2^31-1), bit-exact with the reference's Schrage recursion.  Fortran's intent(inout) seed is
returned as the second element of each result.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A


def r8_uniform_01(seed):
    """r8_uniform_01 (normal_ScaRL.f90:746-841) -> (r, seed)."""
    s = C.c_int32(int(seed))
    r = C.c_double()
    A.call("scarl_hip_r8_uniform_01", C.byref(s), C.byref(r))
    return r.value, s.value


def r8vec_uniform_01(n, seed):
    """r8vec_uniform_01 (normal_ScaRL.f90:1435-1515) -> (r, seed)."""
    s = C.c_int32(int(seed))
    r = np.empty(max(int(n), 0), np.float64)
    A.call("scarl_hip_r8vec_uniform_01", int(n), C.byref(s), A.ptr(r) if n > 0 else None)
    return r, s.value


def r8vec_normal_01(n, seed):
    """r8vec_normal_01 (normal_ScaRL.f90:1138-1259) -> (x, seed)."""
    s = C.c_int32(int(seed))
    x = np.empty(max(int(n), 0), np.float64)
    A.call("scarl_hip_r8vec_normal_01", int(n), C.byref(s), A.ptr(x) if n > 0 else None)
    return x, s.value
