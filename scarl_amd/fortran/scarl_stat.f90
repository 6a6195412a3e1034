! scarl_stat -- program main_Stat (SRC_F/Stat_ScaRL/main_Stat.f90) for one process and one GPU.
!
! Input on stdin, read list-directed one record at a time as in the reference (:38-71): the
! number of files, calculateCorrL, deleteSampleInTheEnd, then a path and a dataset name per file.
! With one process the reference's block choice (estimate_nFields_2, set_Local_Range) is one block
! over the whole field, so the whole dataset is read (read_RF_h5_File_Table, :567-608). The
! moments (calculate_average_and_stdVar_MPI), rebuild_Sk and rebuild_corrL run on the GPU through
! libscarl_hip.so. The statistics go to the sample's root (write_StatisticsOnH5, :200-279) and
! to STAT_<name> (makeStatFile, :610-787) through libscarl_h5.so.
program scarl_stat
    use iso_c_binding
    use scarl_hip_iface
    use scarl_h5_iface
    implicit none
    integer :: nFiles, calculateCorrL, deleteSampleInTheEnd, i
    character(len=1024) :: resPath, dset_name

    call scarl_check(scarl_hip_init(0_c_int32_t), "scarl_hip_init")
    read (*, *) nFiles
    read (*, *) calculateCorrL
    read (*, *) deleteSampleInTheEnd
    do i = 1, nFiles
        read (*, *) resPath
        read (*, *) dset_name
        call process_file(trim(adjustl(resPath)), trim(adjustl(dset_name)))
    end do
    call scarl_check(scarl_hip_finalize(), "scarl_hip_finalize")
    print *, "EXIT CODE: OK"

contains

    integer function read_int(h, name)
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: name
        integer(c_int32_t), dimension(1) :: v
        integer(c_int32_t) :: n
        call h5_check(scarl_h5_read_attr_i32(h, cstr(""), cstr(name), v, 1_c_int32_t, n), name)
        read_int = v(1)
    end function

    subroutine read_int_vec(h, name, v)
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: name
        integer, dimension(3), intent(out) :: v
        integer(c_int32_t), dimension(3) :: t
        integer(c_int32_t) :: n
        call h5_check(scarl_h5_read_attr_i32(h, cstr(""), cstr(name), t, 3_c_int32_t, n), name)
        v = t
    end subroutine

    subroutine read_real_vec(h, name, v)
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: name
        double precision, dimension(3), intent(out) :: v
        integer(c_int32_t) :: n
        call h5_check(scarl_h5_read_attr_f64(h, cstr(""), cstr(name), v, 3_c_int32_t, n), name)
    end subroutine

    subroutine put_real(h, name, v)
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: name
        double precision, dimension(:), intent(in) :: v
        call h5_check(scarl_h5_attr_f64(h, cstr(""), cstr(name), v, int(size(v), c_int32_t)), name)
    end subroutine

    subroutine put_real_scalar(h, name, x)
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: name
        double precision, intent(in) :: x
        double precision, dimension(1) :: v
        v(1) = x
        call h5_check(scarl_h5_attr_f64(h, cstr(""), cstr(name), v, 0_c_int32_t), name)
    end subroutine

    subroutine put_int(h, name, x)
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: name
        integer, intent(in) :: x
        integer(c_int32_t), dimension(1) :: v
        v(1) = x
        call h5_check(scarl_h5_attr_i32(h, cstr(""), cstr(name), v, 0_c_int32_t), name)
    end subroutine

    subroutine put_int_vec(h, name, x)
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: name
        integer, dimension(3), intent(in) :: x
        call h5_check(scarl_h5_attr_i32(h, cstr(""), cstr(name), int(x, c_int32_t), 3_c_int32_t), name)
    end subroutine

    ! write_StatisticsOnH5 (:200-279); existing attributes are replaced
    subroutine write_statistics(path, evntAvg, evntStdDev, gAvg, gStd, withCorrL, corrL_out, SkTot, nhalf, nDim)
        character(len=*), intent(in) :: path
        double precision, dimension(:), intent(in) :: evntAvg, evntStdDev, corrL_out, SkTot
        double precision, intent(in) :: gAvg, gStd
        logical, intent(in) :: withCorrL
        integer, dimension(3), intent(in) :: nhalf
        integer, intent(in) :: nDim
        type(c_ptr) :: h
        integer :: d, i0
        character(len=8) :: nm
        call h5_check(scarl_h5_open(cstr(path), 1_c_int32_t, h), "h5fopen_f "//path)
        call put_real_scalar(h, "globalAvg", gAvg)
        call put_real_scalar(h, "globalStdDev", gStd)
        call put_real(h, "evntAvg", evntAvg)
        call put_real(h, "evntStdDev", evntStdDev)
        if (withCorrL) then
            call put_real(h, "corrL_out", corrL_out(1:nDim))
            i0 = 0
            do d = 1, nDim
                write (nm, '(A,I0)') "Sk_", d
                call put_real(h, trim(nm), SkTot(i0 + 1:i0 + nhalf(d)))
                i0 = i0 + nhalf(d)
            end do
        end if
        call h5_check(scarl_h5_close(h), "h5fclose_f")
    end subroutine

    subroutine process_file(path, dset)
        character(len=*), intent(in) :: path, dset
        type(c_ptr) :: h, hn
        integer :: nDim, d, i0, k, slash, dot
        integer, dimension(3) :: L, v3, nhalf
        integer(c_int32_t), dimension(3) :: dims
        integer(c_int32_t) :: prec
        double precision, dimension(3) :: xStep, r3
        double precision, dimension(:, :, :), allocatable, target :: field
        double precision, dimension(:), allocatable :: sk_part, SkTot
        double precision, dimension(1) :: ea, es
        double precision, dimension(3) :: corrL_out
        double precision :: ga, gs, mb
        integer(8) :: fsize
        character(len=1024) :: newPath, base
        character(len=16), dimension(7) :: ints = [character(len=16) :: "nb_procs", "nDim", "Nmc", "method", &
                                                   "seedStart", "corrMod", "margiFirst"]
        character(len=16), dimension(4) :: ivecs = [character(len=16) :: "nFields", "Np", "Np_ovlp", "L"]
        character(len=16), dimension(5) :: rvecs = [character(len=16) :: "xMinGlob", "xMaxGlob", "xStep", "corrL", &
                                                    "overlap"]

        ! read_RF_h5_File_Attributes, read_RF_h5_File_Table
        call h5_check(scarl_h5_open(cstr(path), 0_c_int32_t, h), "ERROR OPENING FILE "//path)
        nDim = read_int(h, "nDim")
        call read_int_vec(h, "L", L)
        call read_real_vec(h, "xStep", xStep)
        call h5_check(scarl_h5_dataset_info(h, cstr(dset), dims, prec), dset)
        if (any(dims /= L)) stop "scarl_stat: the dataset's extents differ from the attribute L"
        allocate(field(L(1), L(2), L(3)))
        call h5_check(scarl_h5_read_block(h, cstr(dset), c_loc(field), 64_c_int32_t, int(L(1), c_int64_t), &
                                          int(L(1), c_int64_t) * L(2), [0_c_int32_t, 0_c_int32_t, 0_c_int32_t], &
                                          int(L, c_int32_t)), dset)
        call h5_check(scarl_h5_close(h), "h5fclose_f")

        ! calculate_average_and_stdVar_MPI, rebuild_Sk, rebuild_corrL
        call scarl_check(scarl_hip_moments(c_loc(field), int(size(field), c_int64_t), 1_c_int32_t, ea, es, ga, gs), &
                         "calculate_average_and_stdVar")
        nhalf = 0
        corrL_out = 0d0
        allocate(SkTot(sum(L(1:nDim) / 2) + 1))
        if (calculateCorrL == 1) then
            i0 = 0
            do d = 1, nDim
                allocate(sk_part(L(d)))
                call scarl_check(scarl_hip_rebuild_sk_fft(c_loc(field), int(L, c_int32_t), int(d, c_int32_t), &
                                                          int(nDim, c_int32_t), ga, gs, sk_part), "rebuild_Sk_FFT")
                nhalf(d) = L(d) / 2
                SkTot(i0 + 1:i0 + nhalf(d)) = sk_part(1:nhalf(d))
                i0 = i0 + nhalf(d)
                deallocate(sk_part)
            end do
            call scarl_check(scarl_hip_rebuild_corrl(SkTot, int(nhalf, c_int32_t), int(nDim, c_int32_t), xStep, &
                                                     corrL_out), "rebuild_corrL")
            call write_statistics(path, ea, es, ga, gs, .true., corrL_out, SkTot, nhalf, nDim)
        end if

        ! makeStatFile: STAT_<name> beside the sample
        slash = index(path, "/", back=.true.)
        newPath = path(1:slash)//"STAT_"//path(slash + 1:)
        inquire(file=path, size=fsize)
        mb = dble(fsize) / dble(1024d0**2d0)
        call h5_check(scarl_h5_open(cstr(path), 0_c_int32_t, h), "h5fopen_f "//path)
        call h5_check(scarl_h5_create(cstr(newPath), hn), "h5fcreate_f "//trim(newPath))
        do k = 1, size(ints)
            call put_int(hn, trim(ints(k)), read_int(h, trim(ints(k))))
        end do
        call put_real_scalar(hn, "old_file_mb_size", mb)
        do k = 1, size(ivecs)
            call read_int_vec(h, trim(ivecs(k)), v3)
            call put_int_vec(hn, trim(ivecs(k)), v3)
        end do
        do k = 1, size(rvecs)
            call read_real_vec(h, trim(rvecs(k)), r3)
            call put_real(hn, trim(rvecs(k)), r3)
        end do
        call h5_check(scarl_h5_close(hn), "h5fclose_f")
        call h5_check(scarl_h5_close(h), "h5fclose_f")
        call write_statistics(trim(newPath), ea, es, ga, gs, calculateCorrL == 1, corrL_out, SkTot, nhalf, nDim)
        if (deleteSampleInTheEnd == 1) then
            ! rm <path up to the file name's first dot>*
            dot = index(path(slash + 1:), ".")
            base = path
            if (dot > 0) base = path(1:slash + dot - 1)
            call execute_command_line("rm "//trim(base)//"*")
        end if
        print *, "INFO Output on: ", trim(newPath)
    end subroutine process_file

end program scarl_stat
