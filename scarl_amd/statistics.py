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


def rebuild_Sk_FFT(randFieldVec, xNStep, dir, nDim, avg, stdDev):
    """rebuild_Sk_FFT (SRC_F/Stat_ScaRL/statistics_RF.f90:258-462): Sk_part of direction ``dir``
    (1-based) for the field ``randFieldVec`` of shape ``xNStep`` (Fortran order), normalised by
    (x - avg)/stdDev and tapered along ``dir`` when nDim = 3.  Returns xNStep(dir) values."""
    n = [int(v) for v in xNStep][:nDim] + [1] * (3 - nDim)
    x = np.asfortranarray(np.asarray(randFieldVec, np.float64).reshape(n, order="F"))
    out = np.empty(n[dir - 1])
    A.call("scarl_hip_rebuild_sk_fft", A.ptr(x), A.ptr(np.asarray(n, np.int32)), int(dir), int(nDim), float(avg),
           float(stdDev), A.ptr(out))
    return out


def rebuild_Sk(randField, L_Loc, nDim, globalAvg, globalStdDev):
    """rebuild_Sk for one rank (SRC_F/Stat_ScaRL/statistics_RF.f90:180-252): the first
    L_Loc(d)/2 values of Sk_part along each direction, concatenated (SkTot_Dir); with one rank
    the ALLREDUCE average is the identity.  Returns (SkTot_Dir, nhalf)."""
    parts, nhalf = [], []
    for d in range(1, nDim + 1):
        sk = rebuild_Sk_FFT(randField, L_Loc, d, nDim, globalAvg, globalStdDev)
        h = int(L_Loc[d - 1]) // 2
        parts.append(sk[:h])
        nhalf.append(h)
    return np.concatenate(parts), nhalf


def rebuild_corrL(SkTot_Dir, nhalf, xStep):
    """rebuild_corrL (SRC_F/Stat_ScaRL/statistics_RF.f90:468-486)."""
    nd = len(nhalf)
    sk = np.ascontiguousarray(np.asarray(SkTot_Dir, np.float64))
    nh = np.zeros(3, np.int32)
    nh[:nd] = nhalf
    xs = np.zeros(3)
    xs[:nd] = np.asarray(xStep, np.float64)[:nd]
    out = np.empty(nd)
    A.call("scarl_hip_rebuild_corrl", A.ptr(sk), A.ptr(nh), int(nd), A.ptr(xs), A.ptr(out))
    return out


def lag_covariance(field, mean, nlag):
    """Per-axis centred covariance of a 3D field (Fortran order) at lags 0..nlag-1 on the device:
    array (3, nlag); NaN where the lag reaches past the extent."""
    x = np.asfortranarray(np.asarray(field, np.float64))
    if x.ndim == 2:
        x = x.reshape(x.shape + (1,), order="F")
    out = np.empty((3, int(nlag)))
    A.call("scarl_hip_lag_covariance", A.ptr(x), A.ptr(np.asarray(x.shape, np.int32)), float(mean), int(nlag),
           A.ptr(out))
    return out
