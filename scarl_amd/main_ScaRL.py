"""GPU versions of the hot element routines of ``program main_ScaRL`` (SRC_F/main_ScaRL.f90).

The reference keeps them as internal procedures of the program; here they are callable.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A


def apply_UnityPartition_mtx(Np, Np_ovlp, partitionType, topo_pos, topo_shape, randField_3D):
    """apply_UnityPartition_mtx (main_ScaRL.f90:1216-1301): randField_3D *= sqrt(PoU), in place."""
    f = randField_3D
    if not (f.dtype == np.float64 and f.flags.f_contiguous):
        raise ValueError("randField_3D must be a Fortran-ordered float64 array")
    i3 = lambda v: np.ascontiguousarray(np.asarray(v, np.int32).reshape(3))
    A.call("scarl_hip_apply_unity_partition", A.ptr(f), A.ptr(i3(Np)), A.ptr(i3(Np_ovlp)), int(partitionType),
           A.ptr(i3(topo_pos)), A.ptr(i3(topo_shape)))
    return f


def normalize_field(topo_pos, topo_shape, Np, Np_ovlp, rank, comm_group, randField):
    """normalize_field (main_ScaRL.f90:903-953) on one rank, in place; returns (avg, std_dev).
    The reference reduces over every rank of comm_group: here topo_shape must be (1, 1, 1)."""
    del rank, comm_group  # one rank: the ALLREDUCE is the identity
    f = randField
    if not (f.dtype == np.float64 and f.flags.f_contiguous):
        raise ValueError("randField must be a Fortran-ordered float64 array")
    i3 = lambda v: np.ascontiguousarray(np.asarray(v, np.int32).reshape(3))
    avg, sd = C.c_double(), C.c_double()
    A.call("scarl_hip_normalize_field", A.ptr(f), A.ptr(i3(Np)), A.ptr(i3(Np_ovlp)), A.ptr(i3(topo_pos)),
           A.ptr(i3(topo_shape)), C.byref(avg), C.byref(sd))
    return avg.value, sd.value


def multiVariateTransformation(avg, std_dev, margiFirst, randField):
    """multiVariateTransformation (main_ScaRL.f90:1355-1395), in place."""
    f = randField
    if not (f.dtype == np.float64 and (f.flags.f_contiguous or f.flags.c_contiguous)):
        raise ValueError("randField must be a contiguous float64 array")
    A.call("scarl_hip_multivariate_transformation", float(avg), float(std_dev), int(margiFirst), A.ptr(f),
           int(f.size))
    return f
