"""scarl_amd -- MI355X (gfx950) spectral random-field generator with ScaRL's entry points.

Host-side mirror of the reference modules on the hot path (all compute runs in
libscarl_hip.so through the C ABI of include/scarl_hip.h):

* ``fftw_ScaRL``   -- gen_std_gauss_FFT, gen_Std_Gauss_FFT_MPI (SRC_F/fftw_ScaRL.f90)
* ``normal_ScaRL`` -- r8_uniform_01, r8vec_uniform_01, r8vec_normal_01 (SRC_F/normal_ScaRL.f90)
* ``main_ScaRL``   -- apply_UnityPartition_mtx, normalize_field, multiVariateTransformation, and
                      the program itself (``main``): decks in, HDF5/XDMF files out (SRC_F/main_ScaRL.f90)
* ``statistics``   -- calculate_average_and_stdVar, rebuild_Sk, rebuild_corrL
                      (SRC_F/Stat_ScaRL/statistics_RF.f90)
* ``main_Stat``    -- the statistics program (SRC_F/Stat_ScaRL/main_Stat.f90)
* ``read_input``   -- the deck readers (SRC_F/read_input.f90)
* ``write_output`` -- the HDF5/XDMF writers over libscarl_h5.so (SRC_F/write_output.f90, hdf5_ScaRL.f90)
* ``geometry``     -- the subdomain geometry and topology of a sample
* ``generator``    -- Hermitian spectral generator, ensembles and blended subdomain batches
"""
from ._capi import (CM_EXPONENTIAL, CM_GAUSSIAN, CM_VON_KARMAN, FOM_GAUSSIAN, FOM_LOGNORMAL, FOM_NONE,
                    PREC_FP32, PREC_FP64, ScarlError, ScarlLibraryMissing, lib)

__all__ = ["CM_GAUSSIAN", "CM_VON_KARMAN", "CM_EXPONENTIAL", "FOM_NONE", "FOM_GAUSSIAN", "FOM_LOGNORMAL",
           "PREC_FP32", "PREC_FP64", "ScarlError", "ScarlLibraryMissing", "lib"]
