"""Drop-in for the reference module ``spectra_ScaRL`` (SRC_F/spectra_ScaRL.f90) and the dead
Shinozuka consumer gen_std_gauss_Shino_FFT (SRC_F/fftw_ScaRL.f90:157-255), on the GPU."""
from __future__ import annotations

import numpy as np

from . import _capi as A


def _i3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(3))


def _d3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


def build_Sk_mtx(corrMod, xRange, corrL, Np, sz_c, seed, rank):
    """build_Sk_mtx (spectra_ScaRL.f90:10-142) -> newly allocated complex Sk_mtx (Fortran order)."""
    shape = np.zeros(3, np.int32)
    A.call("scarl_hip_build_sk_query", int(corrMod), A.ptr(_d3(xRange)), A.ptr(_d3(corrL)), A.ptr(_i3(Np)),
           A.ptr(_i3(sz_c)), int(seed), int(rank), A.ptr(shape))
    sk = np.empty(tuple(int(v) for v in shape), np.complex128, order="F")
    A.call("scarl_hip_build_sk", int(corrMod), A.ptr(_d3(xRange)), A.ptr(_d3(corrL)), A.ptr(_i3(Np)),
           A.ptr(_i3(sz_c)), int(seed), int(rank), A.ptr(sk))
    return sk


def gen_std_gauss_Shino_FFT(data_real_3D, Np, xRange_in, corrL, pointsPerCorrL, corrMod, seed, rank):
    """gen_std_gauss_Shino_FFT (fftw_ScaRL.f90:157-255): real part of the padded C2C backward
    transform of the Sk quadrants, written into data_real_3D (Fortran float64, shape Np)."""
    Np = _i3(Np)
    if not (data_real_3D.dtype == np.float64 and data_real_3D.flags.f_contiguous
            and data_real_3D.shape == tuple(int(v) for v in Np)):
        raise ValueError("data_real_3D must be a Fortran-ordered float64 array of shape Np")
    A.call("scarl_hip_gen_std_gauss_shino_fft", A.ptr(data_real_3D), A.ptr(Np), A.ptr(_d3(xRange_in)),
           A.ptr(_d3(corrL)), A.ptr(_i3(pointsPerCorrL)), int(corrMod), int(seed), int(rank))
    return data_real_3D
