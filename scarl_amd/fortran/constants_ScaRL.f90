! constants_ScaRL -- enum codes and constants with the reference's values
! (SRC_F/constants_ScaRL.f90:5-27); cm_EXPONENTIAL is new in this build.
module constants_ScaRL
    implicit none
    double precision, parameter :: PI = 3.1415926535898d0
    double precision, parameter :: SQRT_2PI = 2.50662827463d0
    double precision, parameter :: ZERO = 0d0
    double precision, parameter :: TOLERANCE = 0.000000001
    double precision, parameter :: MIN_DOUBLE = -2.0D+307
    double precision, parameter :: MAX_DOUBLE = 2.0D+307
    integer, parameter :: FFT = 1, FFT_MPI = 2
    integer, parameter :: cm_GAUSSIAN = 1, cm_VON_KARMAN = 2, cm_EXPONENTIAL = 3
    integer, parameter :: fom_GAUSSIAN = 1, fom_LOGNORMAL = 2
    integer, parameter :: msh_AUTO = 1, msh_UNV = 2
    integer, parameter :: SCREEN = 6
    integer, parameter :: NATIVE = 1, SEM = 2
end module constants_ScaRL
