! dropin_driver -- calls the drop-in modules the way create_fields does
! (SRC_F/main_ScaRL.f90:510-521) and writes the raw results for the parity test.
! usage: dropin_driver <out.bin>    reads Np, xRange, corrL, seed, rank from stdin and the
! phases phi(Np/2) as raw doubles from <out.bin>.phi; then a second stdin record with the
! spectra_ScaRL / Shinozuka case: NpS(3), xRangeS(3), corrLS(3), sz_c(3), seedS.
program dropin_driver
    use iso_c_binding
    use constants_ScaRL
    use fftw_ScaRL
    use normal_ScaRL
    use spectra_ScaRL
    implicit none
    integer :: Np(3), seed, rank, u, iseed, NpS(3), sz_c(3), seedS
    double precision :: xRange(3), corrL(3), xRangeS(3), corrLS(3)
    double precision, allocatable :: k_mtx(:,:,:), phi(:,:,:), r(:), k_mpi(:,:,:), shino(:,:,:)
    double complex, allocatable :: Sk_mtx(:,:,:)
    character(len=512) :: path
    call get_command_argument(1, path)
    read(*, *) Np, xRange, corrL, seed, rank
    read(*, *) NpS, xRangeS, corrLS, sz_c, seedS
    allocate(k_mtx(Np(1), Np(2), Np(3)), phi(Np(1)/2, Np(2)/2, Np(3)/2))
    open(newunit=u, file=trim(path)//'.phi', access='stream', form='unformatted', status='old')
    read(u) phi
    close(u)
    call gen_std_gauss_FFT(k_mtx, Np, xRange, corrL, cm_GAUSSIAN, seed, rank, phi)
    call gen_Std_Gauss_FFT_MPI(k_mpi, xRange, corrL, cm_GAUSSIAN, seed, rank, Np, Np, 0)
    ! spectra_ScaRL::build_Sk_mtx allocates Sk_mtx (a second call re-allocates it)
    call build_Sk_mtx(cm_GAUSSIAN, xRangeS / corrLS, corrLS, NpS, sz_c, seedS, rank, Sk_mtx)
    call build_Sk_mtx(cm_GAUSSIAN, xRangeS / corrLS, corrLS, NpS, sz_c, seedS, rank, Sk_mtx)
    allocate(shino(NpS(1), NpS(2), NpS(3)))
    call gen_std_gauss_Shino_FFT(shino, NpS, xRangeS, corrLS, [5, 5, 5], cm_GAUSSIAN, seedS, rank)
    allocate(r(5))
    iseed = 12345
    call r8vec_uniform_01(5, iseed, r)
    open(newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
    write(u) k_mtx
    write(u) k_mpi
    write(u) r
    write(u) iseed
    write(u) shape(Sk_mtx)
    write(u) Sk_mtx
    write(u) shino
    close(u)
    print *, " EXIT CODE: OK "
end program dropin_driver
