!===============================================================================
! Test driver of the Fortran TDMA drop-in: reads one system (raw stream file:
! nr, ntheta, nz, n_sweep, is_parallel as int32, then aW, aE, aS, aN, aB,
! aT, aP, Su, phi as (-1:n+2)^3 float64 arrays in Fortran order), calls
! tdma_3d_mpi as the momentum assembler does (src/mod_momentum_3d.f90:329)
! and writes the library status (int32) and phi.
!
!   tdma_driver <input> <output>
!===============================================================================
program tdma_driver
    use mod_types_3d, only: mesh_t
    use mod_eafp_solver, only: tdma_3d_mpi, eafp_last_status
    implicit none
    character(len=1024) :: fin, fout
    integer :: u, nr, nth, nz, n_sweep, ipar, st
    real(8), allocatable :: aW(:,:,:), aE(:,:,:), aS(:,:,:), aN(:,:,:)
    real(8), allocatable :: aB(:,:,:), aT(:,:,:), aP(:,:,:), Su(:,:,:), phi(:,:,:)
    type(mesh_t) :: m

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) nr, nth, nz, n_sweep, ipar
    allocate(aW(-1:nr+2, -1:nth+2, -1:nz+2))
    allocate(aE, aS, aN, aB, aT, aP, Su, phi, mold=aW)
    read(u) aW, aE, aS, aN, aB, aT, aP, Su, phi
    close(u)

    m%nr = nr
    m%ntheta = nth
    m%nz = nz
    m%is_parallel = (ipar /= 0)
    m%topo%jloc = nth

    call tdma_3d_mpi(aW, aE, aS, aN, aB, aT, aP, Su, phi, m, n_sweep)
    st = eafp_last_status()

    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) st
    write(u) phi
    close(u)
end program tdma_driver
