! read_input -- the deck readers of ScaRL for a one-process run (SRC_F/read_input.f90).
!
! Same module name, procedure names and dummy arguments as the reference: read_config_ScaRL
! (:26-105), read_extra_config_ScaRL (:110-147), read_input_ScaRL (:152-348) and getLine
! (:565-585). The values are read with Fortran's own list-directed READ of each line that getLine
! returns. In a one-process run, rank 0 is the only rank, so there is nothing to broadcast:
! comm_group is accepted and not used. The SEM path (random.spec) is not supported; it stops with
! a message.
module read_input
    implicit none
    private
    public :: read_config_ScaRL, read_extra_config_ScaRL, read_input_ScaRL, getLine

contains

    logical function file_exists(path)
        character(len=*), intent(in) :: path
        inquire(file=trim(adjustl(path)), exist=file_exists)
    end function

    subroutine read_config_ScaRL(config_file, gen_meth, c_info_file, d_sample, one_file, one_hdf_ds, &
                                 input_file, rank, comm_group)
        character(len=*), intent(in) :: config_file
        integer, intent(in) :: rank, comm_group
        integer, intent(out) :: gen_meth
        logical, intent(out) :: c_info_file, d_sample, one_file, one_hdf_ds
        character(len=*), intent(out) :: input_file
        integer :: fid, tmp_int, error
        character(len=1024) :: buffer

        if (file_exists(config_file)) then
            open(newunit=fid, file=config_file, action='read', iostat=error)
            buffer = getLine(fid, '#')
            read(buffer, *) gen_meth
            buffer = getLine(fid, '#')
            read(buffer, *) tmp_int
            c_info_file = tmp_int == 1
            buffer = getLine(fid, '#')
            read(buffer, *) tmp_int
            d_sample = tmp_int == 1
            buffer = getLine(fid, '#')
            read(buffer, *) tmp_int
            one_file = tmp_int == 1
            buffer = getLine(fid, '#')
            read(buffer, *) tmp_int
            one_hdf_ds = tmp_int == 1
            buffer = getLine(fid, '#')
            read(buffer, *) input_file
            close(fid)
        else
            if (rank == 0) print *, "FILE: '"//trim(adjustl(config_file))//"' not found. Taking default parameters."
            gen_meth = 1
            c_info_file = .false.
            d_sample = .false.
            one_file = .true.
            one_hdf_ds = .true.
            input_file = "./input_ScaRL.txt"
        end if
    end subroutine read_config_ScaRL

    subroutine read_extra_config_ScaRL(x_config_file, extra_config, topo_shape_in, rank, comm_group)
        character(len=*), intent(in) :: x_config_file
        integer, intent(in) :: rank, comm_group
        logical, intent(out) :: extra_config
        integer, dimension(3), intent(out) :: topo_shape_in
        integer :: fid, error
        character(len=1024) :: buffer

        topo_shape_in = -1
        extra_config = .false.
        if (file_exists(x_config_file)) then
            open(newunit=fid, file=x_config_file, action='read', iostat=error)
            extra_config = .true.
            buffer = getLine(fid, '#')
            read(buffer, *) topo_shape_in
            close(fid)
            if (rank == 0) print *, "topo_shape_in    = ", topo_shape_in
        end if
    end subroutine read_extra_config_ScaRL

    subroutine read_input_ScaRL(file_path_in, rank, nSamples, output_folder, xMinGlob, xMaxGlob, corrL, corrMod, &
                                margiFirst, seedBase, output_name, avg, std_dev, overlap, pointsPerCorrL, comm_group)
        character(len=*), intent(in) :: file_path_in
        integer, intent(in) :: rank, comm_group
        integer, intent(out) :: nSamples
        character(len=1024), intent(out) :: output_folder
        double precision, dimension(:, :), allocatable, intent(out) :: xMinGlob, xMaxGlob, corrL, overlap
        integer, dimension(:), allocatable, intent(out) :: corrMod, margiFirst, seedBase
        character(len=1024), dimension(:), allocatable, intent(out) :: output_name
        double precision, dimension(:), allocatable, intent(out) :: avg, std_dev
        integer, dimension(:, :), allocatable, intent(out) :: pointsPerCorrL
        character(len=1024) :: buffer, file_path
        integer :: fid, s, error

        ! the reference opens getcwd()//"/"//file_path_in: the same file as the relative path
        if (file_exists("random.spec")) stop "read_input_ScaRL: the SEM input path is not supported"
        file_path = trim(adjustl(file_path_in))
        if (.not. file_exists(file_path)) stop "'input_Scarl.txt' NOT FOUND"
        open(newunit=fid, file=file_path, action='read', iostat=error)
        buffer = getLine(fid, '#')
        read(buffer, *) output_folder
        buffer = getLine(fid, '#')
        read(buffer, *) nSamples
        allocate(output_name(nSamples), xMinGlob(3, nSamples), xMaxGlob(3, nSamples), corrL(3, nSamples))
        allocate(overlap(3, nSamples), pointsPerCorrL(3, nSamples), corrMod(nSamples), margiFirst(nSamples))
        allocate(avg(nSamples), std_dev(nSamples), seedBase(nSamples))
        do s = 1, nSamples
            buffer = getLine(fid, '#')
            read(buffer, *) output_name(s)
            buffer = getLine(fid, '#')
            read(buffer, *) xMinGlob(:, s)
            buffer = getLine(fid, '#')
            read(buffer, *) xMaxGlob(:, s)
            buffer = getLine(fid, '#')
            read(buffer, *) corrL(:, s)
            buffer = getLine(fid, '#')
            read(buffer, *) overlap(:, s)
            buffer = getLine(fid, '#')
            read(buffer, *) pointsPerCorrL(:, s)
            buffer = getLine(fid, '#')
            read(buffer, *) corrMod(s)
            buffer = getLine(fid, '#')
            read(buffer, *) margiFirst(s)
            buffer = getLine(fid, '#')
            read(buffer, *) avg(s)
            buffer = getLine(fid, '#')
            read(buffer, *) std_dev(s)
            buffer = getLine(fid, '#')
            read(buffer, *) seedBase(s)
        end do
        close(fid)
        if (rank == 0) print *, "Number of Samples = ", nSamples
    end subroutine read_input_ScaRL

    ! The next line whose first character after left adjustment is not comment_Tag; "eof_gl" at
    ! the end of the file; the 200th line read if every one was a comment.
    function getLine(fid, comment_Tag) result(nextLine)
        integer, intent(in) :: fid
        character(len=1), intent(in) :: comment_Tag
        character(len=1024) :: nextLine
        integer :: i, stat

        do i = 1, 200
            read(fid, fmt="(A)", iostat=stat) nextLine
            nextLine = adjustl(nextLine)
            if (stat /= 0) then
                nextLine = "eof_gl"
                exit
            else if (nextLine(1:1) /= comment_Tag) then
                exit
            end if
        end do
    end function getLine

end module read_input
