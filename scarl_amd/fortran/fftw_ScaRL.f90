! fftw_ScaRL -- GPU drop-in for the reference module of the same name (SRC_F/fftw_ScaRL.f90).
! Same module name, procedure names and dummy arguments; the bodies call libscarl_hip.so.
! The reference's random_number draws (:355-365) become device Philox phases keyed by
! seed + rank, or the optional host-exported phase array phi (for parity runs).
module fftw_ScaRL
    use iso_c_binding
    use constants_ScaRL
    use normal_ScaRL
    use spectra_ScaRL
    use scarl_hip_iface
    implicit none
contains

    ! gen_std_gauss_Shino_FFT (SRC_F/fftw_ScaRL.f90:157-255; the reference driver does not call
    ! it): build_Sk_mtx on xRange_in/corrL, its four quadrants placed on the
    ! 2**ceiling(log2(2 Np)) padded grid, a 3D C2C backward transform, and the real part of the
    ! first Np points into data_real_3D.  pointsPerCorrL is unused, as in the reference.
    subroutine gen_std_gauss_Shino_FFT(data_real_3D, Np, xRange_in, corrL, pointsPerCorrL, corrMod, seed, rank)
        integer, dimension(3), intent(in) :: Np
        double precision, dimension(3), intent(in) :: xRange_in
        integer, intent(in) :: corrMod
        double precision, dimension(3), intent(in) :: corrL
        integer, dimension(3), intent(in) :: pointsPerCorrL
        integer, intent(in) :: seed
        integer, intent(in) :: rank
        double precision, dimension(:,:,:), intent(inout), target, contiguous :: data_real_3D
        if (any(shape(data_real_3D) /= Np)) stop "gen_std_gauss_Shino_FFT: data_real_3D shape /= Np"
        call scarl_check(scarl_hip_gen_std_gauss_shino_fft(c_loc(data_real_3D), int(Np, c_int32_t), xRange_in, &
                         corrL, int(pointsPerCorrL, c_int32_t), int(corrMod, c_int32_t), int(seed, c_int32_t), &
                         int(rank, c_int32_t)), "gen_std_gauss_Shino_FFT")
    end subroutine gen_std_gauss_Shino_FFT

    ! gen_std_gauss_FFT (SRC_F/fftw_ScaRL.f90:261-418)
    subroutine gen_std_gauss_FFT(data_real_3D, Np, xRange, corrL, corrMod, seed, rank, phi)
        integer, dimension(3), intent(in) :: Np
        double precision, dimension(3), intent(in) :: xRange
        integer, intent(in) :: corrMod
        double precision, dimension(3), intent(in) :: corrL
        integer, intent(in) :: seed
        integer, intent(in) :: rank
        double precision, dimension(:,:,:), intent(inout), target, contiguous :: data_real_3D
        double precision, dimension(:,:,:), intent(in), target, contiguous, optional :: phi
        integer(c_int32_t) :: np_c(3)
        type(c_ptr) :: pphi
        if (any(shape(data_real_3D) /= Np)) stop "gen_std_gauss_FFT: data_real_3D shape /= Np"
        np_c = int(Np, c_int32_t)
        pphi = c_null_ptr
        if (present(phi)) then
            if (any(shape(phi) /= Np/2)) stop "gen_std_gauss_FFT: phi shape /= Np/2"
            pphi = c_loc(phi)
        end if
        call scarl_check(scarl_hip_gen_std_gauss_fft(c_loc(data_real_3D), np_c, xRange, corrL, &
                         int(corrMod, c_int32_t), int(seed, c_int32_t), int(rank, c_int32_t), pphi, 0_c_int32_t), &
                         "gen_std_gauss_FFT")
    end subroutine gen_std_gauss_FFT

    ! gen_Std_Gauss_FFT_MPI (SRC_F/fftw_ScaRL.f90:424-583) with one rank holding the whole
    ! field; comm is accepted for interface compatibility (its group must have one rank).
    subroutine gen_Std_Gauss_FFT_MPI(data_real_3D_out, xRange, corrL, corrMod, seed, rank, L, Np, comm, phi)
        integer, dimension(3), intent(in) :: L
        integer, intent(in) :: rank, comm
        double precision, dimension(3), intent(in) :: xRange
        integer, intent(in) :: corrMod
        double precision, dimension(3), intent(in) :: corrL
        integer, intent(in) :: seed
        integer, dimension(3), intent(in) :: Np
        double precision, dimension(:,:,:), allocatable, intent(out), target :: data_real_3D_out
        double precision, dimension(:,:,:), intent(in), target, contiguous, optional :: phi
        type(c_ptr) :: pphi
        integer :: dummy
        dummy = comm
        allocate(data_real_3D_out(L(1), L(2), L(3)))
        pphi = c_null_ptr
        if (present(phi)) pphi = c_loc(phi)
        call scarl_check(scarl_hip_gen_std_gauss_fft_mpi(c_loc(data_real_3D_out), xRange, corrL, &
                         int(corrMod, c_int32_t), int(seed, c_int32_t), int(rank, c_int32_t), &
                         int(L, c_int32_t), int(Np, c_int32_t), 1_c_int32_t, pphi, 0_c_int32_t), &
                         "gen_Std_Gauss_FFT_MPI")
    end subroutine gen_Std_Gauss_FFT_MPI

end module fftw_ScaRL
