! scarl_main -- program main_ScaRL (SRC_F/main_ScaRL.f90:1-220) for one process and one GPU.
!
! Usage, from the folder that holds config_ScaRL.txt and the sample deck:
!     scarl_main <nb_procs> [32|64]
! nb_procs is the number of MPI ranks the reference would run on; it fixes the subdomain grid
! (extra_config_ScaRL.txt, or decide_topo_shape). The second argument is the generation
! precision (default 32).
!
! What is restated from the reference:
! * the geometry chain (:100-176, 455-468) and the topology (decide_topo_shape, find_factors,
!   get_topo_pos, :747-978);
! * the input checks of verify_inputs (:286-333);
! * the output branches, names and attributes of create_fields (:405-735).
! Every subdomain of a sample is generated in one batched GPU call (scarl_hip_blend_create):
! generation, partition of unity, overlap sum and marginal, fused. FFT_MPI generates the one
! global field. The field then goes through the write_output module, rank by rank.
program scarl_main
    use iso_c_binding
    use scarl_hip_iface
    use read_input
    use write_output
    implicit none

    integer, parameter :: FFT = 1, FFT_MPI = 2
    integer :: nb_procs_in, prec, gen_meth, nSamples, s, nb_procs
    logical :: c_info_file, d_sample, one_file, one_hdf_ds, extra_config
    character(len=1024) :: input_file, output_folder, arg
    double precision, dimension(:, :), allocatable :: xMinGlob, xMaxGlob, corrL, overlap
    double precision, dimension(:), allocatable :: avg, std_dev
    integer, dimension(:), allocatable :: corrMod, margiFirst, seedBase
    integer, dimension(:, :), allocatable :: pointsPerCorrL
    character(len=1024), dimension(:), allocatable :: output_name
    integer, dimension(3) :: topo_shape_in, topo_shape, L, Np, Np_ovlp
    double precision, dimension(3) :: xStep
    integer(8) :: t0, t_init, rate
    double precision :: t_read

    nb_procs_in = 1
    prec = 32
    if (command_argument_count() >= 1) then
        call get_command_argument(1, arg)
        read (arg, *) nb_procs_in
    end if
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg)
        read (arg, *) prec
    end if
    call system_clock(t_init, rate)
    t0 = t_init
    call scarl_check(scarl_hip_init(0_c_int32_t), "scarl_hip_init")
    call read_config_ScaRL("./config_ScaRL.txt", gen_meth, c_info_file, d_sample, one_file, one_hdf_ds, input_file, 0, 0)
    call read_input_ScaRL(input_file, 0, nSamples, output_folder, xMinGlob, xMaxGlob, corrL, corrMod, margiFirst, &
                          seedBase, output_name, avg, std_dev, overlap, pointsPerCorrL, 0)
    call read_extra_config_ScaRL("extra_config_ScaRL.txt", extra_config, topo_shape_in, 0, 0)
    call execute_command_line("mkdir -p "//trim(adjustl(output_folder)))

    do s = 1, nSamples
        t_read = elapsed(t0)
        call system_clock(t0)
        call verify_inputs(xMinGlob(:, s), xMaxGlob(:, s), corrL(:, s), overlap(:, s), avg(s), std_dev(s), corrMod(s), &
                           margiFirst(s), pointsPerCorrL(:, s), gen_meth)
        xStep = corrL(:, s) / dble(pointsPerCorrL(:, s) - 1)
        L = 1 + ceiling((xMaxGlob(:, s) - xMinGlob(:, s)) / xStep)
        if (gen_meth == FFT_MPI) then
            Np_ovlp = 0
        else
            Np_ovlp = 1 + ceiling(overlap(:, s) * dble(pointsPerCorrL(:, s) - 1))
        end if
        if (extra_config) then
            if (product(topo_shape_in) /= nb_procs_in) stop 'ERROR! extra_config_ScaRL.txt found but procs grid '// &
                'dimensions do not agree with the total number of procs in MPI process'
            topo_shape = topo_shape_in
        else
            call decide_topo_shape(nb_procs_in, L, Np_ovlp, gen_meth, topo_shape)
        end if
        nb_procs = product(topo_shape)
        if (std_dev(s) == 0d0) then
            call write_constant_field(xMinGlob(:, s), xMaxGlob(:, s), avg(s), output_name(s), output_folder)
            cycle
        end if
        Np = ceiling(dble(L + (topo_shape - 1) * Np_ovlp) / dble(topo_shape))
        where (topo_shape == 2 .and. Np < Np_ovlp) Np = Np_ovlp
        where (topo_shape >= 3 .and. Np < 2 * Np_ovlp) Np = 2 * Np_ovlp
        L = Np * topo_shape - (topo_shape - 1) * Np_ovlp
        print *, "sample ", trim(output_name(s)), ": topo_shape =", topo_shape, " Np =", Np, " L =", L
        call create_fields(s, t_read, elapsed(t0))
        call system_clock(t0)
    end do
    call scarl_check(scarl_hip_finalize(), "scarl_hip_finalize")
    print *, "EXIT CODE: OK"

contains

    double precision function elapsed(since)
        integer(8), intent(in) :: since
        integer(8) :: now
        call system_clock(now)
        elapsed = dble(now - since) / dble(rate)
    end function

    function num2str5(n) result(str)
        integer, intent(in) :: n
        character(len=5) :: str
        write (str, '(I5.5)') n
    end function

    ! SRC_F/main_ScaRL.f90:747-764
    subroutine get_topo_pos(rank, shp, pos)
        integer, intent(in) :: rank
        integer, dimension(3), intent(in) :: shp
        integer, dimension(3), intent(out) :: pos
        integer :: rest
        pos(3) = rank / (shp(1) * shp(2))
        rest = rank - pos(3) * shp(1) * shp(2)
        pos(2) = rest / shp(1)
        pos(1) = rest - pos(2) * shp(1)
    end subroutine

    ! SRC_F/main_ScaRL.f90:959-978
    subroutine find_factors(n, d)
        integer, intent(in) :: n
        integer, dimension(:), intent(out) :: d
        integer :: div, nxt, rest, i
        d = 0
        i = 1
        div = 2
        nxt = 3
        rest = n
        do while (rest /= 1)
            do while (mod(rest, div) == 0)
                d(i) = div
                i = i + 1
                rest = rest / div
            end do
            div = nxt
            nxt = nxt + 2
        end do
    end subroutine

    ! SRC_F/main_ScaRL.f90:786-897 with pointsPerBlockIdeal = 1 (:29). vol_surf_factor starts at
    ! 0 (the reference compares against it before any assignment). FFT_MPI: FFTW-MPI's default
    ! slabs of ceiling(L3 / nb_procs) planes.
    subroutine decide_topo_shape(np_total, Lin, nov, meth, shp)
        integer, intent(in) :: np_total, meth
        integer, dimension(3), intent(in) :: Lin, nov
        integer, dimension(3), intent(out) :: shp
        integer, dimension(3) :: nBlocks, nBlocksIdeal, nPoints, per, nFieldsIdeal, nFieldsChosen
        integer, dimension(100) :: factors
        integer :: np, i, pos, blk
        double precision :: vsf, vsf_t

        if (meth == FFT_MPI) then
            blk = (Lin(3) + np_total - 1) / np_total
            shp = [1, 1, (Lin(3) + blk - 1) / blk]
            return
        end if
        nBlocks = 1
        do
            nPoints = Lin + (nBlocks - 1) * nov
            per = ceiling(dble(nPoints) / dble(nBlocks))
            if (product(int(per, 8)) <= 1_8) exit
            if (product(nBlocks) >= np_total) exit
            pos = maxloc(per, 1)
            nBlocks(pos) = nBlocks(pos) + 1
        end do
        nBlocksIdeal = nBlocks
        vsf = 0d0
        nFieldsChosen = 1
        do np = ceiling(0.75 * np_total), np_total
            nBlocks = nBlocksIdeal
            nFieldsIdeal = 1
            call find_factors(np, factors)
            do i = size(factors), 1, -1
                if (factors(i) == 0) cycle
                if (all(nBlocks == 1)) exit
                pos = maxloc(nBlocks, 1)
                nFieldsIdeal(pos) = nFieldsIdeal(pos) * factors(i)
                nBlocks(pos) = ceiling(dble(nBlocks(pos)) / dble(factors(i)))
            end do
            vsf_t = dble(product(nFieldsIdeal)) / dble(2 * sum(nFieldsIdeal * cshift(nFieldsIdeal, shift=1)))
            if (vsf_t >= vsf) then
                vsf = vsf_t
                nFieldsChosen = nFieldsIdeal
            end if
        end do
        shp = nFieldsChosen
    end subroutine

    ! SRC_F/main_ScaRL.f90:286-333
    subroutine verify_inputs(xMin, xMax, cL, ovl, av, sd, cm, mf, ppc, meth)
        double precision, dimension(3), intent(in) :: xMin, xMax, cL, ovl
        double precision, intent(in) :: av, sd
        integer, intent(in) :: cm, mf, meth
        integer, dimension(3), intent(in) :: ppc
        if (any(xMin >= xMax)) stop "ERROR: xMinGlob >= xMaxGlob"
        if (any(cL < 0d0)) stop "ERROR: corrL < 0d0"
        if (any(ovl < 0d0)) print *, "WARNING: overlap < 0d0, the mesh will be discontinuous"
        if (cm /= 1) stop "ERROR: Correlation model not implemented"
        if (mf /= 1 .and. mf /= 2) stop "ERROR: First-order marginal not implemented"
        if (meth /= FFT .and. meth /= FFT_MPI) stop "ERROR: Generation Method  not implemented"
        if (av <= 0d0 .and. mf == 2) stop "ERROR: When using lognormal first-order marginal density the average "// &
            "should be greater than 0"
        if (sd < 0d0) stop "ERROR: Standard deviation should not be smaller than 0d0"
        if (any(ppc < 2)) stop "ERROR: PointsPerCorrL < 2"
    end subroutine

    logical function exists(path)
        character(len=*), intent(in) :: path
        inquire(file=trim(path), exist=exists)
    end function

    ! create_fields (SRC_F/main_ScaRL.f90:343-739) for every rank of sample s
    subroutine create_fields(s, t_read, t_topo)
        integer, intent(in) :: s
        double precision, intent(in) :: t_read, t_topo
        character(len=1024) :: oname, HDF5_name, XMF_name, res_folder, info_name, h5r
        character(len=20), dimension(9) :: labels
        double precision, dimension(9) :: trace
        character(len=1024), dimension(:), allocatable :: HDF5_list
        double precision, dimension(:, :), allocatable :: coord_0_list
        double precision, dimension(:, :, :), allocatable :: field
        real(c_float), dimension(:, :, :), allocatable, target :: fieldf
        double precision, dimension(:, :, :), allocatable, target :: fieldd
        integer(c_int32_t), dimension(:), allocatable :: starts, exts
        integer, dimension(3) :: tp, last_pos, pos_0, pos_N
        double precision, dimension(3) :: coord_0, coord_N, xRange, xMaxG, ovl
        type(scarl_field_desc) :: desc
        type(c_ptr) :: h
        character(len=8) :: dstr
        character(len=10) :: tstr
        integer :: r, a, k, i
        integer(8) :: t1
        real :: u

        res_folder = output_folder
        oname = output_name(s)
        call names(oname, HDF5_name, XMF_name)
        if (exists(trim(res_folder)//"/"//trim(HDF5_name))) then
            call date_and_time(dstr, tstr)
            oname = trim(output_name(s))//"-"//dstr(3:8)//"_"//tstr(1:6)
            call names(oname, HDF5_name, XMF_name)
        end if
        call get_topo_pos(nb_procs - 1, topo_shape, last_pos)
        xRange = dble(Np - 1) * xStep
        xMaxG = dble((Np - Np_ovlp) * last_pos) * xStep + xMinGlob(:, s) + xRange
        ovl = dble(Np_ovlp - 1) * xStep

        ! generation
        call system_clock(t1)
        desc%n = [2, 2, 2]
        desc%dx = xStep
        desc%model = corrMod(s)
        desc%corrL = corrL(:, s)
        desc%nu = 0.5d0
        desc%sigma = 1d0
        desc%margi = margiFirst(s)
        desc%avg = avg(s)
        desc%std_dev = std_dev(s)
        if (seedBase(s) >= 0) then
            desc%seed = int(seedBase(s), c_int64_t) + 1
        else
            call random_seed()
            call random_number(u)
            desc%seed = 1 + int(u * 2147483646.0, c_int64_t)
        end if
        desc%stream = 0
        desc%normalize = 1
        if (gen_meth == FFT_MPI) then
            desc%n = L
            call scarl_check(scarl_hip_batch_create([desc], 1_c_int32_t, int(prec, c_int32_t), h), "batch_create")
            k = 0
        else
            allocate(starts(sum(topo_shape)), exts(sum(topo_shape)))
            i = 0
            do a = 1, 3
                do r = 0, topo_shape(a) - 1
                    i = i + 1
                    starts(i) = (Np(a) - Np_ovlp(a)) * r
                    exts(i) = Np(a)
                end do
            end do
            call scarl_check(scarl_hip_blend_create(int(L, c_int32_t), int(topo_shape, c_int32_t), starts, exts, desc, &
                                                    int(prec, c_int32_t), h), "blend_create")
            k = -1
        end if
        call scarl_check(scarl_hip_batch_run(h, 0_c_int32_t), "batch_run")
        call scarl_check(scarl_hip_batch_sync(h), "batch_sync")
        if (prec == 32) then
            allocate(fieldf(L(1), L(2), L(3)))
            call scarl_check(scarl_hip_batch_copy_output(h, int(k, c_int32_t), c_loc(fieldf), int(size(fieldf), c_int64_t) &
                                                         * 4), "batch_copy_output")
            field = dble(fieldf)
        else
            allocate(fieldd(L(1), L(2), L(3)))
            call scarl_check(scarl_hip_batch_copy_output(h, int(k, c_int32_t), c_loc(fieldd), int(size(fieldd), c_int64_t) &
                                                         * 8), "batch_copy_output")
            field = fieldd
        end if
        call scarl_check(scarl_hip_batch_destroy(h), "batch_destroy")
        trace(3) = elapsed(t1)
        trace(4:7) = 0d0   ! unit partition, overlap sum, normalisation and marginal ran inside the generation

        ! output, rank by rank (the token ring's order)
        call system_clock(t1)
        allocate(HDF5_list(nb_procs), coord_0_list(3, nb_procs))
        do r = 0, nb_procs - 1
            call get_topo_pos(r, topo_shape, tp)
            pos_0 = (Np - Np_ovlp) * tp + 1
            pos_N = pos_0 + Np - 1
            coord_0 = dble(pos_0 - 1) * xStep + xMinGlob(:, s)
            coord_N = dble(pos_N - 1) * xStep + xMinGlob(:, s)
            coord_0_list(:, r + 1) = dble((Np - Np_ovlp) * tp) * xStep + xMinGlob(:, s)
            if (one_file) then
                HDF5_list(r + 1) = trim(oname)//".h5"
                if (one_hdf_ds) then
                    call write_hdf5_multi_proc_3D_1ds(coord_0, coord_N, xMinGlob(:, s), xMaxG, pos_0, pos_N, L, Np, &
                        Np_ovlp, field(pos_0(1):pos_N(1), pos_0(2):pos_N(2), pos_0(3):pos_N(3)), HDF5_name, XMF_name, &
                        res_folder, r, nb_procs, 0)
                else
                    call write_hdf5_multi_proc_3D(coord_0, coord_N, xMinGlob(:, s), xMaxG, pos_0, pos_N, &
                        field(pos_0(1):pos_N(1), pos_0(2):pos_N(2), pos_0(3):pos_N(3)), HDF5_name, res_folder, r, &
                        nb_procs, 0)
                end if
            else
                h5r = trim(oname)//"_proc_"//num2str5(r)//".h5"
                HDF5_list(r + 1) = h5r
                call write_hdf5_single_proc_3D(coord_0, coord_N, xMinGlob(:, s), xMaxG, pos_0, pos_N, &
                    field(pos_0(1):pos_N(1), pos_0(2):pos_N(2), pos_0(3):pos_N(3)), h5r, &
                    trim(oname)//"_proc_"//num2str5(r)//".xmf", res_folder, r)
                call write_HDF5_attributes(trim(res_folder)//"/"//trim(h5r), nb_procs, 3, 1, FFT, seedBase(s), &
                    corrMod(s), margiFirst(s), topo_shape, xMinGlob(:, s), xMaxG, xStep, corrL(:, s), ovl, L, Np, &
                    Np_ovlp, .false.)
            end if
        end do
        if (one_file) call write_HDF5_attributes(trim(res_folder)//"/"//trim(oname)//".h5", nb_procs, 3, 1, FFT, &
            seedBase(s), corrMod(s), margiFirst(s), topo_shape, xMinGlob(:, s), xMaxG, xStep, corrL(:, s), ovl, L, Np, &
            Np_ovlp, .false.)
        if (.not. (one_file .and. one_hdf_ds)) call write_XMF_global(trim(res_folder)//"/"//trim(oname)//".xmf", &
            HDF5_list, coord_0_list, xStep, Np, nb_procs)
        trace(8) = elapsed(t1)
        trace(1) = t_read
        trace(2) = t_topo
        trace(9) = elapsed(t_init)
        labels = [character(len=20) :: "Read_input", "Topology", "Generation", "Apply_unit", "Add_overlap", &
                  "Normalize", "Multivar_trans", "Write_file", "Total_Wall_time"]
        if (one_file) then
            call write_time_on_hdf5(labels, trace, HDF5_name, res_folder, 0, nb_procs, 0, .true.)
        else
            do r = 0, nb_procs - 1
                call write_time_on_hdf5(labels, trace, HDF5_list(r + 1), res_folder, r, nb_procs, 0, .false.)
            end do
        end if
        if (c_info_file) then
            info_name = "INFO-"//trim(oname)//".h5"
            if (exists(trim(res_folder)//"/"//trim(info_name))) then
                call date_and_time(dstr, tstr)
                info_name = "INFO-"//trim(oname)//"-"//dstr(3:8)//"_"//tstr(1:6)//".h5"
            end if
            call write_info_file(info_name, res_folder, 0)
            call write_HDF5_attributes(trim(res_folder)//"/"//trim(info_name), nb_procs, 3, 1, FFT, seedBase(s), &
                corrMod(s), margiFirst(s), topo_shape, xMinGlob(:, s), xMaxG, xStep, corrL(:, s), ovl, L, Np, Np_ovlp, &
                .false.)
            call write_time_on_hdf5(labels, trace, info_name, res_folder, 0, nb_procs, 0, .true.)
        end if
        if (d_sample) call execute_command_line("rm "//trim(res_folder)//"/"//trim(oname)//"*")
    end subroutine create_fields

    subroutine names(oname, HDF5_name, XMF_name)
        character(len=*), intent(in) :: oname
        character(len=*), intent(out) :: HDF5_name, XMF_name
        if (one_file) then
            HDF5_name = trim(oname)//".h5"
            XMF_name = trim(oname)//".xmf"
        else
            HDF5_name = trim(oname)//"_proc_"//num2str5(0)//".h5"
            XMF_name = trim(oname)//"_proc_"//num2str5(0)//".xmf"
        end if
    end subroutine

end program scarl_main
