! bind(C) shim over the reference's own routines (compiled from /root/reference sources
! into oracle/_ref by `make ref`).  Test infrastructure only: used to pin the oracle's
! LCG (normal_ScaRL.f90:746-841,1138-1259,1435-1515) and K_nu (bessel.f90:13522-13785).
module scarl_ref_shim
    use iso_c_binding
    use bessel, only: ikv
    use normal_ScaRL, only: r8vec_uniform_01, r8vec_normal_01
    implicit none
contains
    subroutine ref_ikv(v, x, nmax, bk) bind(C, name="ref_ikv")
        real(c_double), value :: v, x
        integer(c_int), value :: nmax
        real(c_double), intent(out) :: bk(0:nmax)
        real(8) :: vm
        real(8), allocatable :: bi(:), di(:), dkk(:), bkk(:)
        allocate(bi(0:nmax+300), di(0:nmax+300), bkk(0:nmax+300), dkk(0:nmax+300))
        call ikv(v, x, vm, bi, di, bkk, dkk)
        bk(0:nmax) = bkk(0:nmax)
    end subroutine
    subroutine ref_uniform(n, seed, r) bind(C, name="ref_r8vec_uniform_01")
        integer(c_int), value :: n
        integer(c_int), intent(inout) :: seed
        real(c_double), intent(out) :: r(n)
        call r8vec_uniform_01(n, seed, r)
    end subroutine
    subroutine ref_normal(n, seed, x) bind(C, name="ref_r8vec_normal_01")
        integer(c_int), value :: n
        integer(c_int), intent(inout) :: seed
        real(c_double), intent(out) :: x(n)
        call r8vec_normal_01(n, seed, x)
    end subroutine
end module
