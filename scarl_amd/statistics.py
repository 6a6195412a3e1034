"""GPU moments of ``Stat_ScaRL`` (SRC_F/Stat_ScaRL/statistics_RF.f90:18-174)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A


def calculate_average_and_stdVar(randField):
    """randField: (xNTotal, Nmc) -> (evntAvg, evntStdDev, globalAvg, globalStdDev), with
    FP64 compensated sums on the device and the reference's formulas (:80-87)."""
    rf = np.asfortranarray(randField, dtype=np.float64)
    if rf.ndim == 1:
        rf = rf.reshape(-1, 1, order="F")
    n, nmc = rf.shape
    ea = np.empty(nmc)
    es = np.empty(nmc)
    ga = C.c_double()
    gs = C.c_double()
    A.call("scarl_hip_moments", A.ptr(rf), int(n), int(nmc), A.ptr(ea), A.ptr(es), C.byref(ga), C.byref(gs))
    return ea, es, ga.value, gs.value


def kv(nu, x):
    """K_nu(x) in FP64 on the device (the ikv scheme, SRC_F/bessel.f90:13522-13785)."""
    x = np.ascontiguousarray(np.asarray(x, np.float64).ravel())
    out = np.empty_like(x)
    A.call("scarl_hip_kv", float(nu), A.ptr(x), int(x.size), A.ptr(out))
    return out


def target_corr(model, rho, nu=0.5):
    """Target correlation C(rho)/sigma^2 on the device (FP64, K_nu for von Karman)."""
    rho = np.ascontiguousarray(np.asarray(rho, np.float64).ravel())
    out = np.empty_like(rho)
    A.call("scarl_hip_target_corr", int(model), float(nu), A.ptr(rho), int(rho.size), A.ptr(out))
    return out
