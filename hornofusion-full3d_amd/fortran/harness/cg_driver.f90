!===============================================================================
! Test driver of the Fortran drop-in: reads one assembled p' system (raw
! stream file: nr, ntheta, nz, max_iter, is_parallel as int32, tol as
! float64, then aW, aE, aS, aN, aB, aT, aP, Su, phi as (-1:n+2)^3 float64
! arrays in Fortran order), calls cg_3d_mpi exactly as
! solve_pressure_correction does (src/mod_pressure_3d.f90:316-318), and
! writes residual (float64), n_iter and the library status (int32) and phi.
!
!   cg_driver <input> <output>
!===============================================================================
program cg_driver
    use mod_types_3d, only: mesh_t
    use mod_eafp_solver, only: cg_3d_mpi, eafp_last_status
    implicit none
    character(len=1024) :: fin, fout
    integer :: u, nr, nth, nz, max_iter, ipar, n_iter, st
    real(8) :: tol, residual
    real(8), allocatable :: aW(:,:,:), aE(:,:,:), aS(:,:,:), aN(:,:,:)
    real(8), allocatable :: aB(:,:,:), aT(:,:,:), aP(:,:,:), Su(:,:,:), phi(:,:,:)
    type(mesh_t) :: m

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) nr, nth, nz, max_iter, ipar
    read(u) tol
    allocate(aW(-1:nr+2, -1:nth+2, -1:nz+2))
    allocate(aE, aS, aN, aB, aT, aP, Su, phi, mold=aW)
    read(u) aW, aE, aS, aN, aB, aT, aP, Su, phi
    close(u)

    m%nr = nr
    m%ntheta = nth
    m%nz = nz
    m%is_parallel = (ipar /= 0)
    m%topo%jloc = nth

    call cg_3d_mpi(aW, aE, aS, aN, aB, aT, aP, Su, phi, m, max_iter, tol, residual, n_iter)
    st = eafp_last_status()

    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) residual
    write(u) n_iter, st
    write(u) phi
    close(u)
end program cg_driver
