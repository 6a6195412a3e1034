! scarl_hip_iface -- ISO_C_BINDING interfaces to libscarl_hip.so (include/scarl_hip.h).
module scarl_hip_iface
    use iso_c_binding
    implicit none

    ! scarl_field_desc (include/scarl_hip.h); the member order gives the same C layout
    type, bind(C) :: scarl_field_desc
        integer(c_int32_t) :: n(3)
        real(c_double) :: dx(3)
        integer(c_int32_t) :: model
        real(c_double) :: corrL(3)
        real(c_double) :: nu
        real(c_double) :: sigma
        integer(c_int32_t) :: margi
        real(c_double) :: avg, std_dev
        integer(c_int64_t) :: seed
        integer(c_int32_t) :: stream
        integer(c_int32_t) :: normalize
    end type
    integer(c_int32_t), parameter :: SCARL_PREC_FP32 = 32, SCARL_PREC_FP64 = 64

    interface
        integer(c_int) function scarl_hip_init(device) bind(C, name="scarl_hip_init")
            import :: c_int, c_int32_t
            integer(c_int32_t), value :: device
        end function
        integer(c_int) function scarl_hip_finalize() bind(C, name="scarl_hip_finalize")
            import :: c_int
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
        ! spectra_ScaRL: build_Sk_mtx as a shape query and a fill; the Shinozuka consumer
        integer(c_int) function scarl_hip_build_sk_query(corrMod, xRange, corrL, Np, sz_c, seed, rank, out_shape) &
                bind(C, name="scarl_hip_build_sk_query")
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: corrMod
            real(c_double), dimension(3), intent(in) :: xRange, corrL
            integer(c_int32_t), dimension(3), intent(in) :: Np, sz_c
            integer(c_int32_t), value :: seed, rank
            integer(c_int32_t), dimension(3), intent(out) :: out_shape
        end function
        integer(c_int) function scarl_hip_build_sk(corrMod, xRange, corrL, Np, sz_c, seed, rank, sk) &
                bind(C, name="scarl_hip_build_sk")
            import :: c_int, c_int32_t, c_double, c_ptr
            integer(c_int32_t), value :: corrMod
            real(c_double), dimension(3), intent(in) :: xRange, corrL
            integer(c_int32_t), dimension(3), intent(in) :: Np, sz_c
            integer(c_int32_t), value :: seed, rank
            type(c_ptr), value :: sk
        end function
        integer(c_int) function scarl_hip_gen_std_gauss_shino_fft(out, Np, xRange, corrL, pointsPerCorrL, corrMod, &
                seed, rank) bind(C, name="scarl_hip_gen_std_gauss_shino_fft")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: out
            integer(c_int32_t), dimension(3), intent(in) :: Np
            real(c_double), dimension(3), intent(in) :: xRange, corrL
            integer(c_int32_t), dimension(3), intent(in) :: pointsPerCorrL
            integer(c_int32_t), value :: corrMod, seed, rank
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
        ! Stat_ScaRL moments and corrL estimator (include/scarl_hip.h)
        integer(c_int) function scarl_hip_moments(rf, n, nmc, evnt_avg, evnt_std, g_avg, g_std) &
                bind(C, name="scarl_hip_moments")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: rf
            integer(c_int64_t), value :: n
            integer(c_int32_t), value :: nmc
            real(c_double), dimension(*), intent(out) :: evnt_avg, evnt_std
            real(c_double), intent(out) :: g_avg, g_std
        end function
        integer(c_int) function scarl_hip_rebuild_sk_fft(field, xNStep, dir, nDim, avg, stdDev, sk_part) &
                bind(C, name="scarl_hip_rebuild_sk_fft")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: field
            integer(c_int32_t), dimension(3), intent(in) :: xNStep
            integer(c_int32_t), value :: dir, nDim
            real(c_double), value :: avg, stdDev
            real(c_double), dimension(*), intent(out) :: sk_part
        end function
        integer(c_int) function scarl_hip_rebuild_corrl(sk_tot, nhalf, nDim, xStep, corrL) &
                bind(C, name="scarl_hip_rebuild_corrl")
            import :: c_int, c_int32_t, c_double
            real(c_double), dimension(*), intent(in) :: sk_tot
            integer(c_int32_t), dimension(3), intent(in) :: nhalf
            integer(c_int32_t), value :: nDim
            real(c_double), dimension(3), intent(in) :: xStep
            real(c_double), dimension(*), intent(out) :: corrL
        end function
        ! batched Hermitian generator (include/scarl_hip.h)
        integer(c_int) function scarl_hip_batch_create(fields, count, precision, h) bind(C, name="scarl_hip_batch_create")
            import :: c_int, c_int32_t, c_ptr, scarl_field_desc
            type(scarl_field_desc), dimension(*), intent(in) :: fields
            integer(c_int32_t), value :: count, precision
            type(c_ptr), intent(out) :: h
        end function
        integer(c_int) function scarl_hip_blend_create(gdim, nb, starts, exts, params, precision, h) &
                bind(C, name="scarl_hip_blend_create")
            import :: c_int, c_int32_t, c_ptr, scarl_field_desc
            integer(c_int32_t), dimension(3), intent(in) :: gdim, nb
            integer(c_int32_t), dimension(*), intent(in) :: starts, exts
            type(scarl_field_desc), intent(in) :: params
            integer(c_int32_t), value :: precision
            type(c_ptr), intent(out) :: h
        end function
        integer(c_int) function scarl_hip_batch_run(h, use_graph) bind(C, name="scarl_hip_batch_run")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: use_graph
        end function
        integer(c_int) function scarl_hip_batch_sync(h) bind(C, name="scarl_hip_batch_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function scarl_hip_batch_copy_output(h, field, dst, nbytes) &
                bind(C, name="scarl_hip_batch_copy_output")
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: field
            type(c_ptr), value :: dst
            integer(c_int64_t), value :: nbytes
        end function
        integer(c_int) function scarl_hip_batch_destroy(h) bind(C, name="scarl_hip_batch_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
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
