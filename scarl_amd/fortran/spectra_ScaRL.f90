! spectra_ScaRL -- GPU drop-in for the reference module of the same name (SRC_F/spectra_ScaRL.f90).
! Same module name, procedure name and dummy arguments; the body calls libscarl_hip.so.
! The reference allocates Sk_mtx inside build_Sk_mtx (:74-75), so this wrapper asks the library
! for the shape first, allocates, then has the device fill it.
module spectra_ScaRL
    use iso_c_binding
    use constants_ScaRL
    use normal_ScaRL
    use scarl_hip_iface
    implicit none
contains

    ! build_Sk_mtx (SRC_F/spectra_ScaRL.f90:10-142): Sk_mtx(sz_Sk(1), 2 sz_Sk(2), 2 sz_Sk(3)) with
    ! the phases of r8vec_uniform_01(seed + 1000000) (bit-exact LCG on the device), the zeroed
    ! first/last planes and, for cm_GAUSSIAN, the Gaussian amplitude of the four mirrored quadrants.
    ! An Sk_mtx allocated on entry is released first (the reference would stop on the allocate).
    subroutine build_Sk_mtx(corrMod, xRange, corrL, Np, sz_c, seed, rank, Sk_mtx)
        integer, dimension(3), intent(in) :: Np
        double precision, dimension(3), intent(in) :: xRange, corrL
        integer, intent(in) :: corrMod
        integer, intent(in) :: seed
        integer, intent(in) :: rank
        integer, dimension(3) :: sz_c
        double complex, dimension(:,:,:), allocatable, target :: Sk_mtx
        integer(c_int32_t) :: shp(3), np_c(3), szc_c(3)
        np_c = int(Np, c_int32_t)
        szc_c = int(sz_c, c_int32_t)
        if (allocated(Sk_mtx)) deallocate(Sk_mtx)
        call scarl_check(scarl_hip_build_sk_query(int(corrMod, c_int32_t), xRange, corrL, np_c, szc_c, &
                         int(seed, c_int32_t), int(rank, c_int32_t), shp), "build_Sk_mtx")
        allocate(Sk_mtx(shp(1), shp(2), shp(3)))
        call scarl_check(scarl_hip_build_sk(int(corrMod, c_int32_t), xRange, corrL, np_c, szc_c, &
                         int(seed, c_int32_t), int(rank, c_int32_t), c_loc(Sk_mtx)), "build_Sk_mtx")
    end subroutine build_Sk_mtx

end module spectra_ScaRL
