! scarl_h5_iface -- ISO_C_BINDING interfaces to libscarl_h5.so (include/scarl_h5.h), and the
! status check that turns a failure into the reference's `stop "msg"`.
module scarl_h5_iface
    use iso_c_binding
    implicit none

    interface
        integer(c_int) function scarl_h5_last_error(buf, length) bind(C, name="scarl_h5_last_error")
            import :: c_int, c_int32_t, c_char
            character(kind=c_char), dimension(*) :: buf
            integer(c_int32_t), value :: length
        end function
        integer(c_int) function scarl_h5_create(path, h) bind(C, name="scarl_h5_create")
            import :: c_int, c_char, c_ptr
            character(kind=c_char), dimension(*), intent(in) :: path
            type(c_ptr), intent(out) :: h
        end function
        integer(c_int) function scarl_h5_open(path, writable, h) bind(C, name="scarl_h5_open")
            import :: c_int, c_int32_t, c_char, c_ptr
            character(kind=c_char), dimension(*), intent(in) :: path
            integer(c_int32_t), value :: writable
            type(c_ptr), intent(out) :: h
        end function
        integer(c_int) function scarl_h5_close(h) bind(C, name="scarl_h5_close")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function scarl_h5_create_dataset(h, name, dims, prec) bind(C, name="scarl_h5_create_dataset")
            import :: c_int, c_int32_t, c_char, c_ptr
            type(c_ptr), value :: h
            character(kind=c_char), dimension(*), intent(in) :: name
            integer(c_int32_t), dimension(3), intent(in) :: dims
            integer(c_int32_t), value :: prec
        end function
        integer(c_int) function scarl_h5_write_block(h, name, data, prec, ld1, ld2, offset, count) &
                bind(C, name="scarl_h5_write_block")
            import :: c_int, c_int32_t, c_int64_t, c_char, c_ptr
            type(c_ptr), value :: h
            character(kind=c_char), dimension(*), intent(in) :: name
            type(c_ptr), value :: data
            integer(c_int32_t), value :: prec
            integer(c_int64_t), value :: ld1, ld2
            integer(c_int32_t), dimension(3), intent(in) :: offset, count
        end function
        integer(c_int) function scarl_h5_attr_i32(h, object, name, v, n) bind(C, name="scarl_h5_attr_i32")
            import :: c_int, c_int32_t, c_char, c_ptr
            type(c_ptr), value :: h
            character(kind=c_char), dimension(*), intent(in) :: object, name
            integer(c_int32_t), dimension(*), intent(in) :: v
            integer(c_int32_t), value :: n
        end function
        integer(c_int) function scarl_h5_attr_f64(h, object, name, v, n) bind(C, name="scarl_h5_attr_f64")
            import :: c_int, c_int32_t, c_double, c_char, c_ptr
            type(c_ptr), value :: h
            character(kind=c_char), dimension(*), intent(in) :: object, name
            real(c_double), dimension(*), intent(in) :: v
            integer(c_int32_t), value :: n
        end function
        integer(c_int) function scarl_h5_dataset_info(h, name, dims, prec) bind(C, name="scarl_h5_dataset_info")
            import :: c_int, c_int32_t, c_char, c_ptr
            type(c_ptr), value :: h
            character(kind=c_char), dimension(*), intent(in) :: name
            integer(c_int32_t), dimension(3), intent(out) :: dims
            integer(c_int32_t), intent(out) :: prec
        end function
        integer(c_int) function scarl_h5_read_block(h, name, data, prec, ld1, ld2, offset, count) &
                bind(C, name="scarl_h5_read_block")
            import :: c_int, c_int32_t, c_int64_t, c_char, c_ptr
            type(c_ptr), value :: h
            character(kind=c_char), dimension(*), intent(in) :: name
            type(c_ptr), value :: data
            integer(c_int32_t), value :: prec
            integer(c_int64_t), value :: ld1, ld2
            integer(c_int32_t), dimension(3), intent(in) :: offset, count
        end function
        integer(c_int) function scarl_h5_read_attr_i32(h, object, name, v, cap, n) bind(C, name="scarl_h5_read_attr_i32")
            import :: c_int, c_int32_t, c_char, c_ptr
            type(c_ptr), value :: h
            character(kind=c_char), dimension(*), intent(in) :: object, name
            integer(c_int32_t), dimension(*), intent(out) :: v
            integer(c_int32_t), value :: cap
            integer(c_int32_t), intent(out) :: n
        end function
        integer(c_int) function scarl_h5_read_attr_f64(h, object, name, v, cap, n) bind(C, name="scarl_h5_read_attr_f64")
            import :: c_int, c_int32_t, c_double, c_char, c_ptr
            type(c_ptr), value :: h
            character(kind=c_char), dimension(*), intent(in) :: object, name
            real(c_double), dimension(*), intent(out) :: v
            integer(c_int32_t), value :: cap
            integer(c_int32_t), intent(out) :: n
        end function
        integer(c_int) function scarl_h5_write_times(h, labels, times, ntimes, nprocs) &
                bind(C, name="scarl_h5_write_times")
            import :: c_int, c_int32_t, c_double, c_char, c_ptr
            type(c_ptr), value :: h
            character(kind=c_char), dimension(*), intent(in) :: labels
            real(c_double), dimension(*), intent(in) :: times
            integer(c_int32_t), value :: ntimes, nprocs
        end function
        integer(c_int) function scarl_xmf_write_elements(folder, xmf_name, h5_name, h5_path, ds_name, xMin, xMax, L, &
                prec) bind(C, name="scarl_xmf_write_elements")
            import :: c_int, c_int32_t, c_double, c_char
            character(kind=c_char), dimension(*), intent(in) :: folder, xmf_name, h5_name, h5_path, ds_name
            real(c_double), dimension(3), intent(in) :: xMin, xMax
            integer(c_int32_t), dimension(3), intent(in) :: L
            integer(c_int32_t), value :: prec
        end function
        integer(c_int) function scarl_xmf_write_global(path, h5_list, coord_0_list, xStep, xNStep, nb_procs) &
                bind(C, name="scarl_xmf_write_global")
            import :: c_int, c_int32_t, c_double, c_char, c_ptr
            character(kind=c_char), dimension(*), intent(in) :: path
            type(c_ptr), dimension(*), intent(in) :: h5_list
            real(c_double), dimension(*), intent(in) :: coord_0_list
            real(c_double), dimension(3), intent(in) :: xStep
            integer(c_int32_t), dimension(3), intent(in) :: xNStep
            integer(c_int32_t), value :: nb_procs
        end function
    end interface

contains

    ! a C string from a Fortran string (trailing blanks trimmed)
    function cstr(s) result(c)
        character(len=*), intent(in) :: s
        character(kind=c_char, len=len_trim(s) + 1) :: c
        c = trim(s)//c_null_char
    end function

    subroutine h5_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), dimension(512) :: buf
        character(len=512) :: msg
        integer :: i, rc2
        if (rc == 0) return
        buf = c_null_char
        rc2 = scarl_h5_last_error(buf, 512_c_int32_t)
        msg = ""
        do i = 1, 512
            if (buf(i) == c_null_char) exit
            msg(i:i) = buf(i)
        end do
        print *, "ERROR in ", where, ": ", trim(msg)
        stop "HDF5 output failed"
    end subroutine

end module scarl_h5_iface
