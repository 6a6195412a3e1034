"""Drop-in for the reference module ``fftw_ScaRL`` (SRC_F/fftw_ScaRL.f90) on the GPU.

Same procedure names and dummy arguments; arrays are Fortran-ordered numpy arrays.  The
reference's random_number draws are replaced by a host-exported phase array ``phi`` (for
parity) or, when ``phi`` is None, by device Philox4x32-10 phases keyed by seed + rank.
"""
from __future__ import annotations

import numpy as np

from . import _capi as A


def _i3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(3))


def _d3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


def gen_std_gauss_FFT(data_real_3D, Np, xRange, corrL, corrMod, seed, rank, phi=None, fuse_marginal=False):
    """gen_std_gauss_FFT (SRC_F/fftw_ScaRL.f90:261-418).

    data_real_3D: float64 array of shape Np, Fortran order, intent(inout) -- overwritten with
    the unnormalised REDFT01^3 of sqrt(S) cos(2 PI phi) on the octant [1:Np/2]^3.
    phi: (Np1/2, Np2/2, Np3/2) phases in [0,1), or None for device noise.
    """
    Np = _i3(Np)
    if not (isinstance(data_real_3D, np.ndarray) and data_real_3D.dtype == np.float64
            and data_real_3D.flags.f_contiguous and data_real_3D.shape == tuple(int(v) for v in Np)):
        raise ValueError("data_real_3D must be a Fortran-ordered float64 array of shape Np")
    if phi is not None:
        phi = np.asfortranarray(phi, dtype=np.float64)
        if phi.shape != tuple(int(v) // 2 for v in Np):
            raise ValueError(f"phi must have shape Np/2 = {tuple(int(v) // 2 for v in Np)}")
    A.call("scarl_hip_gen_std_gauss_fft", A.ptr(data_real_3D), A.ptr(Np), A.ptr(_d3(xRange)), A.ptr(_d3(corrL)),
           int(corrMod), int(seed), int(rank), A.ptr(phi), 1 if fuse_marginal else 0)
    return data_real_3D


def gen_Std_Gauss_FFT_MPI(xRange, corrL, corrMod, seed, rank, L, Np, comm_size=1, phi=None):
    """gen_Std_Gauss_FFT_MPI (SRC_F/fftw_ScaRL.f90:424-583) for one rank holding the whole field.

    Returns the newly allocated (L1, L2, L3) field (the reference's allocatable intent(out)).
    phi: (L1/2, L2/2, L3) phases, or None for device noise.
    """
    L = _i3(L)
    Np = _i3(Np)
    out = np.empty(tuple(int(v) for v in L), np.float64, order="F")
    if phi is not None:
        phi = np.asfortranarray(phi, dtype=np.float64)
        if phi.shape != (int(L[0]) // 2, int(L[1]) // 2, int(L[2])):
            raise ValueError("phi must have shape (L1/2, L2/2, L3)")
    A.call("scarl_hip_gen_std_gauss_fft_mpi", A.ptr(out), A.ptr(_d3(xRange)), A.ptr(_d3(corrL)), int(corrMod),
           int(seed), int(rank), A.ptr(L), A.ptr(Np), int(comm_size), A.ptr(phi), 0)
    return out


def set_marginal(margiFirst, avg, std_dev):
    """Marginal fused into gen_std_gauss_FFT when fuse_marginal=True."""
    A.call("scarl_hip_set_marginal", int(margiFirst), float(avg), float(std_dev))
