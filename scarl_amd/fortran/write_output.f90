! write_output -- the HDF5/XDMF writers of ScaRL for a one-process run (SRC_F/write_output.f90),
! over libscarl_h5.so (HDF5 C API).
!
! Same module name, procedure names and dummy arguments as the reference. The reference gathers
! every rank's block at rank 0 in a token ring (rank r waits for r - 1, :373-431). Here all ranks
! live in one process, and the caller calls the multi-rank writers once per rank, in rank order,
! which is the token ring's order:
! * rank 0 creates the file and writes its block;
! * every later rank opens the file and adds its block;
! * the last rank writes the closing attributes and the .xmf.
! comm is accepted and not used.
module write_output
    use iso_c_binding
    use scarl_h5_iface
    implicit none
    private
    public :: write_hdf5_multi_proc_3D_1ds, write_hdf5_multi_proc_3D, write_hdf5_single_proc_3D, &
              write_HDF5_attributes, write_time_on_hdf5, write_XMF_elements, write_XMF_global, &
              write_info_file, write_constant_field

    double precision, dimension(3), save :: xMin_rank0 = 0d0

contains

    function join(folder, name) result(p)
        character(len=*), intent(in) :: folder, name
        character(len=2048) :: p
        p = trim(adjustl(folder))//"/"//trim(adjustl(name))
    end function

    function num2str(n) result(s)
        integer, intent(in) :: n
        character(len=30) :: s
        write (s, '(I30)') n
        s = adjustl(s)
    end function

    ! write_pos_on_HDF5_dataset (:700-716)
    subroutine pos_attrs(h, ds, pos_0, pos_N)
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: ds
        integer, dimension(3), intent(in) :: pos_0, pos_N
        call h5_check(scarl_h5_attr_i32(h, cstr(ds), cstr("pos_0"), int(pos_0, c_int32_t), 3_c_int32_t), "pos_0")
        call h5_check(scarl_h5_attr_i32(h, cstr(ds), cstr("pos_N"), int(pos_N, c_int32_t), 3_c_int32_t), "pos_N")
    end subroutine

    subroutine real_attr(h, ds, name, v)
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: ds, name
        double precision, dimension(:), intent(in) :: v
        call h5_check(scarl_h5_attr_f64(h, cstr(ds), cstr(name), v, int(size(v), c_int32_t)), name)
    end subroutine

    subroutine int_attr(h, name, v)
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: name
        integer, intent(in) :: v
        integer(c_int32_t), dimension(1) :: a
        a(1) = v
        call h5_check(scarl_h5_attr_i32(h, cstr(""), cstr(name), a, 0_c_int32_t), name)
    end subroutine

    subroutine int_vec_attr(h, name, v)
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: name
        integer, dimension(:), intent(in) :: v
        call h5_check(scarl_h5_attr_i32(h, cstr(""), cstr(name), int(v, c_int32_t), int(size(v), c_int32_t)), name)
    end subroutine

    ! h5sselect_hyperslab_f + h5dwrite_f of one block at the 0-based offset
    subroutine write_block(h, ds, field, offset)
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: ds
        double precision, dimension(:, :, :), intent(in) :: field
        integer, dimension(3), intent(in) :: offset
        double precision, dimension(:, :, :), allocatable, target :: tmp
        integer(c_int32_t), dimension(3) :: cnt
        tmp = field
        cnt = int(shape(field), c_int32_t)
        call h5_check(scarl_h5_write_block(h, cstr(ds), c_loc(tmp), 64_c_int32_t, int(cnt(1), c_int64_t), &
                                           int(cnt(1), c_int64_t) * cnt(2), int(offset, c_int32_t), cnt), ds)
    end subroutine

    subroutine open_or_create(path, create, h)
        character(len=*), intent(in) :: path
        logical, intent(in) :: create
        type(c_ptr), intent(out) :: h
        if (create) then
            call h5_check(scarl_h5_create(cstr(path), h), "h5fcreate_f "//trim(path))
        else
            call h5_check(scarl_h5_open(cstr(path), 1_c_int32_t, h), "h5fopen_f "//trim(path))
        end if
    end subroutine

    subroutine write_hdf5_multi_proc_3D_1ds(xMin, xMax, xMinGlob, xMaxGlob, pos_0, pos_N, L, Np, Np_ovlp, &
                                            randField_3D, HDF5_name, XMF_name, res_folder, rank, nb_procs, comm)
        double precision, dimension(3), intent(in) :: xMin, xMax, xMinGlob, xMaxGlob
        integer, dimension(3), intent(in) :: pos_0, pos_N, L, Np, Np_ovlp
        double precision, dimension(:, :, :), intent(in) :: randField_3D
        character(len=*) :: HDF5_name, XMF_name, res_folder
        integer, intent(in) :: rank, nb_procs, comm
        type(c_ptr) :: h
        character(len=*), parameter :: ds = "samples"

        call open_or_create(join(res_folder, HDF5_name), rank == 0, h)
        if (rank == 0) call h5_check(scarl_h5_create_dataset(h, cstr(ds), int(L, c_int32_t), 64_c_int32_t), ds)
        call write_block(h, ds, randField_3D, pos_0 - 1)
        if (rank == 0) then
            call real_attr(h, ds, "xMin", xMin)
            xMin_rank0 = xMin
        else if (rank == nb_procs - 1) then
            call real_attr(h, ds, "xMax", xMax)
        end if
        if (rank == nb_procs - 1) call pos_attrs(h, ds, [1, 1, 1], L)
        call h5_check(scarl_h5_close(h), "h5fclose_f")
        if (rank == nb_procs - 1) call write_XMF_elements(HDF5_name, xMin_rank0, xMax, L, XMF_name, res_folder, ".", ds)
    end subroutine write_hdf5_multi_proc_3D_1ds

    subroutine write_hdf5_multi_proc_3D(xMin, xMax, xMinGlob, xMaxGlob, pos_0, pos_N, randField_3D, HDF5_name, &
                                        res_folder, rank, nb_procs, comm)
        double precision, dimension(3), intent(in) :: xMin, xMax, xMinGlob, xMaxGlob
        integer, dimension(3), intent(in) :: pos_0, pos_N
        double precision, dimension(:, :, :), intent(in) :: randField_3D
        character(len=*) :: HDF5_name, res_folder
        integer, intent(in) :: rank, nb_procs, comm
        type(c_ptr) :: h
        character(len=64) :: ds

        ds = "sample_1_p"//trim(num2str(rank))
        call open_or_create(join(res_folder, HDF5_name), rank == 0, h)
        call h5_check(scarl_h5_create_dataset(h, cstr(ds), int(shape(randField_3D), c_int32_t), 64_c_int32_t), ds)
        call write_block(h, ds, randField_3D, [0, 0, 0])
        call real_attr(h, ds, "xMin", xMin)
        call real_attr(h, ds, "xMax", xMax)
        if (rank > 0) call pos_attrs(h, ds, pos_0, pos_N)   ! rank 0's dataset gets none (:95-97)
        call h5_check(scarl_h5_close(h), "h5fclose_f")
    end subroutine write_hdf5_multi_proc_3D

    subroutine write_hdf5_single_proc_3D(xMin, xMax, xMinGlob, xMaxGlob, pos_0, pos_N, randField_3D, HDF5_name, &
                                         XMF_name, res_folder, rank)
        double precision, dimension(3), intent(in) :: xMin, xMax, xMinGlob, xMaxGlob
        integer, dimension(3), intent(in) :: pos_0, pos_N
        double precision, dimension(:, :, :), intent(in) :: randField_3D
        character(len=*) :: HDF5_name, XMF_name, res_folder
        integer, intent(in), optional :: rank
        type(c_ptr) :: h
        character(len=64) :: ds

        ds = "samples"
        if (present(rank)) ds = "sample_1_p"//trim(num2str(rank))
        call open_or_create(join(res_folder, HDF5_name), .true., h)
        call h5_check(scarl_h5_create_dataset(h, cstr(ds), int(shape(randField_3D), c_int32_t), 64_c_int32_t), ds)
        call write_block(h, ds, randField_3D, [0, 0, 0])
        call real_attr(h, ds, "xMin", xMin)
        call real_attr(h, ds, "xMax", xMax)
        call pos_attrs(h, ds, pos_0, pos_N)
        call h5_check(scarl_h5_close(h), "h5fclose_f")
        call write_XMF_elements(HDF5_name, xMin, xMax, shape(randField_3D), XMF_name, res_folder, ".", ds)
    end subroutine write_hdf5_single_proc_3D

    subroutine write_HDF5_attributes(HDF5Path, nb_procs, nDim, Nmc, method, seedStart, corrMod, margiFirst, nFields, &
                                     xMinGlob, xMaxGlob, xStep, corrL, overlap, L, Np, Np_ovlp, opened)
        character(len=*), intent(in) :: HDF5Path
        integer, intent(in) :: nb_procs, nDim, Nmc, method, seedStart, corrMod, margiFirst
        double precision, dimension(:), intent(in) :: xMinGlob, xMaxGlob, xStep, corrL, overlap
        integer, dimension(:), intent(in) :: nFields, L, Np, Np_ovlp
        logical, intent(in) :: opened
        type(c_ptr) :: h

        call h5_check(scarl_h5_open(cstr(HDF5Path), 1_c_int32_t, h), "h5fopen_f "//trim(HDF5Path))
        call int_attr(h, "nb_procs", nb_procs)
        call int_attr(h, "nDim", nDim)
        call int_attr(h, "Nmc", Nmc)
        call int_attr(h, "method", method)
        call int_attr(h, "seedStart", seedStart)
        call int_attr(h, "corrMod", corrMod)
        call int_attr(h, "margiFirst", margiFirst)
        call int_vec_attr(h, "nFields", nFields)
        call int_vec_attr(h, "Np", Np)
        call int_vec_attr(h, "Np_ovlp", Np_ovlp)
        call int_vec_attr(h, "L", L)
        call real_attr(h, "", "xMinGlob", xMinGlob)
        call real_attr(h, "", "xMaxGlob", xMaxGlob)
        call real_attr(h, "", "xStep", xStep)
        call real_attr(h, "", "corrL", corrL)
        call real_attr(h, "", "overlap", overlap)
        call h5_check(scarl_h5_close(h), "h5fclose_f")
    end subroutine write_HDF5_attributes

    ! With oneFile the reference gathers every rank's times at rank 0 (:857-865). One process has
    ! one set of times: rank 0 writes it into every rank's column.
    subroutine write_time_on_hdf5(time_label, time_trace, HDF5_name, res_folder, rank, nb_procs, comm, oneFile)
        character(len=*), dimension(:), intent(in) :: time_label
        double precision, dimension(:), intent(in) :: time_trace
        character(len=*), intent(in) :: HDF5_name, res_folder
        integer, intent(in) :: rank, nb_procs, comm
        logical, intent(in) :: oneFile
        double precision, dimension(:, :), allocatable :: time_all
        character(kind=c_char, len=20), dimension(:), allocatable :: lab
        type(c_ptr) :: h
        integer :: i, ncol

        if (oneFile .and. rank /= 0) return
        ncol = 1
        if (oneFile) ncol = nb_procs
        allocate(time_all(size(time_trace), ncol), lab(size(time_label)))
        do i = 1, ncol
            time_all(:, i) = time_trace
        end do
        do i = 1, size(time_label)
            lab(i) = time_label(i)
        end do
        call h5_check(scarl_h5_open(cstr(join(res_folder, HDF5_name)), 1_c_int32_t, h), "h5fopen_f")
        call h5_check(scarl_h5_write_times(h, lab, time_all, int(size(time_trace), c_int32_t), int(ncol, c_int32_t)), &
                      "write_time_on_hdf5")
        call h5_check(scarl_h5_close(h), "h5fclose_f")
    end subroutine write_time_on_hdf5

    subroutine write_XMF_elements(HDF5_name, xMin, xMax, xNStep, XMF_name, XMF_folder, H5_TO_XMF_path, ds_name)
        double precision, dimension(3), intent(in) :: xMin, xMax
        integer, dimension(3), intent(in) :: xNStep
        character(len=*), intent(in) :: HDF5_name, XMF_name, XMF_folder, H5_TO_XMF_path, ds_name
        call h5_check(scarl_xmf_write_elements(cstr(XMF_folder), cstr(adjustl(XMF_name)), cstr(adjustl(HDF5_name)), &
                                               cstr(H5_TO_XMF_path), cstr(adjustl(ds_name)), xMin, xMax, &
                                               int(xNStep, c_int32_t), 64_c_int32_t), "write_XMF_elements")
    end subroutine write_XMF_elements

    subroutine write_XMF_global(globXMF_name, HDF5_list, coord_0_list, xStep, xNStep, nb_procs)
        character(len=*) :: globXMF_name
        integer, intent(in) :: nb_procs
        character(len=*), dimension(nb_procs) :: HDF5_list
        double precision, dimension(3, nb_procs) :: coord_0_list
        double precision, dimension(3), intent(in) :: xStep
        integer, dimension(3), intent(in) :: xNStep
        character(kind=c_char, len=1025), dimension(:), allocatable, target :: names
        type(c_ptr), dimension(:), allocatable :: ptrs
        integer :: i

        allocate(names(nb_procs), ptrs(nb_procs))
        do i = 1, nb_procs
            names(i) = trim(adjustl(HDF5_list(i)))//c_null_char
            ptrs(i) = c_loc(names(i))
        end do
        ! the reference opens "."//"/"//globXMF_name (:589-590)
        call h5_check(scarl_xmf_write_global(cstr("./"//trim(adjustl(globXMF_name))), ptrs, coord_0_list, xStep, &
                                             int(xNStep, c_int32_t), int(nb_procs, c_int32_t)), "write_XMF_global")
    end subroutine write_XMF_global

    subroutine write_info_file(info_file_name, res_folder, rank)
        character(len=*), intent(in) :: info_file_name, res_folder
        integer, intent(in) :: rank
        type(c_ptr) :: h
        call h5_check(scarl_h5_create(cstr(join(res_folder, info_file_name)), h), "write_info_file")
        call h5_check(scarl_h5_close(h), "h5fclose_f")
    end subroutine write_info_file

    subroutine write_constant_field(xMinGlob, xMaxGlob, avg, output_name, res_folder)
        double precision, dimension(3), intent(in) :: xMinGlob, xMaxGlob
        double precision, intent(in) :: avg
        character(len=*), intent(in) :: output_name, res_folder
        double precision, dimension(2, 2, 2) :: constant_field
        character(len=1024) :: HDF5_name, XMF_name

        HDF5_name = trim(adjustl(output_name))//".h5"
        XMF_name = trim(adjustl(output_name))//".xmf"
        constant_field = avg
        call write_hdf5_single_proc_3D(xMinGlob, xMaxGlob, xMinGlob, xMaxGlob, [1, 1, 1], [2, 2, 2], constant_field, &
                                       HDF5_name, XMF_name, res_folder)
        call write_HDF5_attributes(join(res_folder, HDF5_name), 1, 3, 1, -1, -1, -1, -1, [1, 1, 1], xMinGlob, &
                                   xMaxGlob, xMaxGlob - xMinGlob, [-1d0, -1d0, -1d0], [0d0, 0d0, 0d0], [2, 2, 2], &
                                   [2, 2, 2], [0, 0, 0], .false.)
    end subroutine write_constant_field

end module write_output
