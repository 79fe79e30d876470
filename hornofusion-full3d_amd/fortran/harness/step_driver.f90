!===============================================================================
! Test driver of the Fortran p' step (eafp_pressure_step).  Reads one case
! from a raw stream file:
!   int32  nr, ntheta, nz, max_iter, is_parallel, gas_in_poisson,
!          solve_multiphase, gas_compressibility, has_liq_cont,
!          has_pv_active, has_gas_reach, has_liq_room, solver, precond
!   f64    dt, rho_gas, T_ambient, beta_expansion, alpha_p, tol
!   f64    r(-1:nr+2), theta(-1:ntheta+2), z(-1:nz+2)
!   f64    vol, Ar, Ath, Az, alpha_l, ur_l, uth_l, uz_l, aPr_l, aPth_l, aPz_l,
!          rho_l, T_l, alpha_g, ur_g, uth_g, uz_g, aPr_g, aPth_g, aPz_g, rho_g,
!          T_old_g, p                      ((-1:n+2)^3, Fortran order)
!   int32  cell_type ((-1:n+2)^3), outlet ((-1:nr+2, -1:ntheta+2))
!   then the masks that are present: int32 liq_cont, pv_active, gas_reach,
!   f64 liq_room
! calls eafp_pressure_step as solve_pressure_correction would
! (src/mod_pressure_3d.f90:174-397) and writes int32 status, n_iter, f64
! residual, cg_res, then pp, Fc_r, Fc_th, Fc_z (f64), lk_r, lk_th, lk_z
! (int32), ur_l, uth_l, uz_l, ur_g, uth_g, uz_g, p (f64).
!
!   step_driver <input> <output>
!===============================================================================
program step_driver
    use, intrinsic :: iso_c_binding, only: c_int, c_double
    use mod_types_3d, only: mesh_t
    use mod_eafp_solver, only: eafp_pressure_step, eafp_last_status
    implicit none
    character(len=1024) :: fin, fout
    integer :: u, st, n_iter
    integer(c_int) :: hdr(14)
    real(c_double) :: par(6), residual, cg_res
    real(c_double), allocatable :: r(:), theta(:), z(:)
    real(c_double), allocatable :: vol(:,:,:), Ar(:,:,:), Ath(:,:,:), Az(:,:,:)
    real(c_double), allocatable :: alpha_l(:,:,:), ur_l(:,:,:), uth_l(:,:,:), uz_l(:,:,:)
    real(c_double), allocatable :: aPr_l(:,:,:), aPth_l(:,:,:), aPz_l(:,:,:), rho_l(:,:,:), T_l(:,:,:)
    real(c_double), allocatable :: alpha_g(:,:,:), ur_g(:,:,:), uth_g(:,:,:), uz_g(:,:,:)
    real(c_double), allocatable :: aPr_g(:,:,:), aPth_g(:,:,:), aPz_g(:,:,:), rho_g(:,:,:), T_old_g(:,:,:)
    real(c_double), allocatable :: p(:,:,:), pp(:,:,:), Fc_r(:,:,:), Fc_th(:,:,:), Fc_z(:,:,:), liq_room(:,:,:)
    integer(c_int), allocatable :: cell_type(:,:,:), outlet(:,:), lk_r(:,:,:), lk_th(:,:,:), lk_z(:,:,:)
    integer(c_int), allocatable :: liq_cont(:,:,:), pv_active(:,:,:), gas_reach(:,:,:)
    type(mesh_t) :: m
    integer :: nr, nth, nz

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) hdr
    read(u) par
    nr = hdr(1)
    nth = hdr(2)
    nz = hdr(3)
    allocate(r(-1:nr+2), theta(-1:nth+2), z(-1:nz+2))
    read(u) r, theta, z
    allocate(vol(-1:nr+2, -1:nth+2, -1:nz+2))
    allocate(Ar, Ath, Az, alpha_l, ur_l, uth_l, uz_l, aPr_l, aPth_l, aPz_l, rho_l, T_l, mold=vol)
    allocate(alpha_g, ur_g, uth_g, uz_g, aPr_g, aPth_g, aPz_g, rho_g, T_old_g, p, mold=vol)
    allocate(pp, Fc_r, Fc_th, Fc_z, mold=vol)
    allocate(cell_type(-1:nr+2, -1:nth+2, -1:nz+2), outlet(-1:nr+2, -1:nth+2))
    allocate(lk_r, lk_th, lk_z, mold=cell_type)
    read(u) vol, Ar, Ath, Az, alpha_l, ur_l, uth_l, uz_l, aPr_l, aPth_l, aPz_l, rho_l, T_l
    read(u) alpha_g, ur_g, uth_g, uz_g, aPr_g, aPth_g, aPz_g, rho_g, T_old_g, p
    read(u) cell_type, outlet
    if (hdr(9) /= 0) then
        allocate(liq_cont, mold=cell_type)
        read(u) liq_cont
    end if
    if (hdr(10) /= 0) then
        allocate(pv_active, mold=cell_type)
        read(u) pv_active
    end if
    if (hdr(11) /= 0) then
        allocate(gas_reach, mold=cell_type)
        read(u) gas_reach
    end if
    if (hdr(12) /= 0) then
        allocate(liq_room, mold=vol)
        read(u) liq_room
    end if
    close(u)

    m%nr = nr
    m%ntheta = nth
    m%nz = nz
    m%is_parallel = (hdr(5) /= 0)
    m%topo%jloc = nth

    ! absent masks are passed as unallocated allocatables: not present
    call eafp_pressure_step(m, r, theta, z, vol, Ar, Ath, Az, cell_type, outlet, &
                            alpha_l, ur_l, uth_l, uz_l, aPr_l, aPth_l, aPz_l, rho_l, T_l, &
                            alpha_g, ur_g, uth_g, uz_g, aPr_g, aPth_g, aPz_g, rho_g, T_old_g, p, &
                            par(1), par(2), par(3), par(4), hdr(6) /= 0, hdr(7) /= 0, hdr(8) /= 0, &
                            par(5), int(hdr(4)), par(6), pp, Fc_r, Fc_th, Fc_z, lk_r, lk_th, lk_z, &
                            residual, cg_res, n_iter, liq_cont, pv_active, gas_reach, liq_room, &
                            hdr(13), hdr(14))
    st = eafp_last_status()

    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) int(st, c_int), int(n_iter, c_int)
    write(u) residual, cg_res
    write(u) pp, Fc_r, Fc_th, Fc_z
    write(u) lk_r, lk_th, lk_z
    write(u) ur_l, uth_l, uz_l, ur_g, uth_g, uz_g, p
    close(u)
end program step_driver
