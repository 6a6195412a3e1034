! scarl_hip_iface -- ISO_C_BINDING interfaces to libscarl_hip.so (include/scarl_hip.h).
module scarl_hip_iface
    use iso_c_binding
    implicit none

    interface
        integer(c_int) function scarl_hip_init(device) bind(C, name="scarl_hip_init")
            import :: c_int, c_int32_t
            integer(c_int32_t), value :: device
        end function
        integer(c_int) function scarl_hip_last_error(buf, length) bind(C, name="scarl_hip_last_error")
            import :: c_int, c_int32_t, c_char
            character(kind=c_char), dimension(*) :: buf
            integer(c_int32_t), value :: length
        end function
        integer(c_int) function scarl_hip_gen_std_gauss_fft(out, Np, xRange, corrL, corrMod, seed, rank, phi, flags) &
                bind(C, name="scarl_hip_gen_std_gauss_fft")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: out
            integer(c_int32_t), dimension(3), intent(in) :: Np
            real(c_double), dimension(3), intent(in) :: xRange, corrL
            integer(c_int32_t), value :: corrMod, seed, rank
            type(c_ptr), value :: phi
            integer(c_int32_t), value :: flags
        end function
        integer(c_int) function scarl_hip_gen_std_gauss_fft_mpi(out, xRange, corrL, corrMod, seed, rank, L, Np, &
                comm_size, phi, flags) bind(C, name="scarl_hip_gen_std_gauss_fft_mpi")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: out
            real(c_double), dimension(3), intent(in) :: xRange, corrL
            integer(c_int32_t), value :: corrMod, seed, rank
            integer(c_int32_t), dimension(3), intent(in) :: L, Np
            integer(c_int32_t), value :: comm_size
            type(c_ptr), value :: phi
            integer(c_int32_t), value :: flags
        end function
        integer(c_int) function scarl_hip_r8vec_uniform_01(n, seed, r) bind(C, name="scarl_hip_r8vec_uniform_01")
            import :: c_int, c_int32_t, c_ptr
            integer(c_int32_t), value :: n
            integer(c_int32_t), intent(inout) :: seed
            type(c_ptr), value :: r
        end function
        integer(c_int) function scarl_hip_r8vec_normal_01(n, seed, x) bind(C, name="scarl_hip_r8vec_normal_01")
            import :: c_int, c_int32_t, c_ptr
            integer(c_int32_t), value :: n
            integer(c_int32_t), intent(inout) :: seed
            type(c_ptr), value :: x
        end function
    end interface

contains

    ! map a non-zero status to the reference's error convention: stop with the message
    subroutine scarl_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), dimension(512) :: buf
        character(len=512) :: msg
        integer :: i, ierr
        if (rc == 0) return
        ierr = scarl_hip_last_error(buf, 512_c_int32_t)
        msg = ' '
        do i = 1, 512
            if (buf(i) == c_null_char) exit
            msg(i:i) = buf(i)
        end do
        write(*, '(a,a,a,i0,a,a)') 'ERROR in ', where, ' (status ', rc, '): ', trim(msg)
        stop 1
    end subroutine

end module scarl_hip_iface
