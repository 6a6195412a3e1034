! normal_ScaRL -- GPU drop-in for the reference's Park-Miller / Box-Muller routines
! (SRC_F/normal_ScaRL.f90:746-841,1138-1259,1435-1515); seed is intent(inout) as there.
module normal_ScaRL
    use iso_c_binding
    use scarl_hip_iface
    implicit none
contains

    function r8_uniform_01(seed)
        integer(kind=4), intent(inout) :: seed
        real(kind=8) :: r8_uniform_01
        real(kind=8), target :: r(1)
        call scarl_check(scarl_hip_r8vec_uniform_01(1_c_int32_t, seed, c_loc(r)), "r8_uniform_01")
        r8_uniform_01 = r(1)
    end function r8_uniform_01

    subroutine r8vec_uniform_01(n, seed, r)
        integer(kind=4), intent(in) :: n
        integer(kind=4), intent(inout) :: seed
        real(kind=8), intent(out), target :: r(n)
        call scarl_check(scarl_hip_r8vec_uniform_01(int(n, c_int32_t), seed, c_loc(r)), "r8vec_uniform_01")
    end subroutine r8vec_uniform_01

    subroutine r8vec_normal_01(n, seed, x)
        integer(kind=4), intent(in) :: n
        integer(kind=4), intent(inout) :: seed
        real(kind=8), intent(out), target :: x(n)
        call scarl_check(scarl_hip_r8vec_normal_01(int(n, c_int32_t), seed, c_loc(x)), "r8vec_normal_01")
    end subroutine r8vec_normal_01

end module normal_ScaRL
