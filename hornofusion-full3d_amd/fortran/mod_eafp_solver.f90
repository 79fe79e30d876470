!===============================================================================
! mod_eafp_solver.f90 - drop-in replacement of EAF3D's p' conjugate gradient
! and of its line-by-line TDMA smoother
!
! cg_3d_mpi keeps the reference subroutine interface exactly
! (src/mod_solver_3d.f90:485-496): seven coefficient arrays and Su with
! explicit lower bound -1 (the gfortran/flang lower-bound pitfall of
! assumed-shape dummies, docs/ARCHITECTURE.md:331-347), phi intent(inout),
! the mesh, max_iter and tol in, residual and n_iter out.  The body hands the
! arrays to eafp_cg_solve in libeafp.so (include/eafp.h) through
! ISO_C_BINDING; the library runs the FP64 Jacobi-PCG on the GPU.
!
! Failure (parallel mesh, HIP error, bad sizes) is reported the way the
! reference's caller already consumes it: residual = NaN, n_iter = 0
! (src/mod_pressure_3d.f90:322).
!
! tdma_3d_mpi keeps the interface of src/mod_solver_3d.f90:252-262 (the same
! eight arrays, phi, the mesh, n_sweep) and hands them to eafp_tdma_solve,
! whose GPU sweeps keep the reference's line order and give phi bit for bit.
! The reference subroutine cannot fail and reports nothing, so a library
! failure stops the program (error stop) rather than return phi untouched.
!
! eafp_pressure_step runs the whole p' step of solve_pressure_correction
! (src/mod_pressure_3d.f90:174-397: sh%pp = 0, assembly, cg_3d_mpi, the
! fluxes, velocity corrections, pressure update and sealed cells) through
! eafp_pp_step, with the p' system kept on the GPU.  Its arguments are the
! reference's own arrays (mesh, both phases, sh%p, the workspace masks and
! fluxes); the velocities and p are updated in place, as the reference does.
! On a library failure nothing is updated and residual = NaN.
!
! In the reference build this module replaces the `use mod_solver_3d` of
! src/mod_pressure_3d.f90:28 for cg_3d_mpi, and of the TDMA callers
! (src/mod_momentum_3d.f90:329, src/mod_energy_3d.f90:320, ...) for
! tdma_3d_mpi; mesh_t comes from mod_types_3d.
!===============================================================================
module mod_eafp_solver
    use, intrinsic :: iso_c_binding, only: c_int, c_double, c_ptr, c_loc, c_null_ptr
    use, intrinsic :: ieee_arithmetic, only: ieee_value, ieee_quiet_nan
    use mod_types_3d, only: mesh_t
    implicit none
    private
    public :: cg_3d_mpi, tdma_3d_mpi, eafp_pressure_step, eafp_last_status

    ! include/eafp.h: EAFP_STEP_*, EAFP_COEF_*, EAFP_PRECOND_*
    integer(c_int), parameter, public :: EAFP_STEP_FP64 = 0, EAFP_STEP_MIXED = 1
    integer(c_int), parameter, public :: EAFP_COEF_BF16 = 1
    integer(c_int), parameter, public :: EAFP_PRECOND_JACOBI = 0, EAFP_PRECOND_ZLINE = 1

    ! include/eafp.h: eafp_desc
    type, bind(C) :: eafp_desc_t
        integer(c_int) :: nr, ntheta, nz
        integer(c_int) :: halo, lbound
        integer(c_int) :: is_parallel, periodic_theta
    end type eafp_desc_t

    ! include/eafp.h: eafp_pp_fields, eafp_pp_config, eafp_pp_update,
    ! eafp_mixed_opts, eafp_pp_step_opts, eafp_stats (the same member order)
    type, bind(C) :: eafp_pp_fields_t
        type(c_ptr) :: r, theta, z, vol, Ar, Ath, Az, cell_type, outlet
        type(c_ptr) :: alpha_l, ur_l, uth_l, uz_l, aPr_l, aPth_l, aPz_l, rho_l, T_l
        type(c_ptr) :: alpha_g, ur_g, uth_g, uz_g, aPr_g, aPth_g, aPz_g, rho_g, T_old_g
        type(c_ptr) :: p, liq_cont, pv_active, gas_reach, liq_room
    end type eafp_pp_fields_t

    type, bind(C) :: eafp_pp_config_t
        real(c_double) :: dt, rho_gas, T_ambient, beta_expansion
        integer(c_int) :: gas_in_poisson, solve_multiphase, gas_compressibility, pad_
    end type eafp_pp_config_t

    type, bind(C) :: eafp_pp_update_t
        type(c_ptr) :: Fc_r, Fc_th, Fc_z, lk_r, lk_th, lk_z
        type(c_ptr) :: ur_l, uth_l, uz_l, ur_g, uth_g, uz_g, p
    end type eafp_pp_update_t

    type, bind(C) :: eafp_mixed_opts_t
        real(c_double) :: tol
        integer(c_int) :: interval, max_steps, coef_format, restart, precond
    end type eafp_mixed_opts_t

    type, bind(C) :: eafp_pp_step_opts_t
        integer(c_int) :: solver, max_iter
        real(c_double) :: tol, alpha_p
        type(eafp_mixed_opts_t) :: mixed
    end type eafp_pp_step_opts_t

    type, bind(C) :: eafp_stats_t
        real(c_double) :: residual
        integer(c_int) :: n_iter, n_refine, status, pad_
        real(c_double) :: true_rel_l2, true_rel_l1, seconds
    end type eafp_stats_t

    interface
        function eafp_pp_step(desc, fields, cfg, opts, pp, io, residual, stats, arrays_on_device) &
                              bind(C, name="eafp_pp_step") result(rc)
            import :: c_int, c_double, c_ptr, eafp_desc_t, eafp_pp_fields_t, eafp_pp_config_t, &
                      eafp_pp_step_opts_t, eafp_pp_update_t, eafp_stats_t
            type(eafp_desc_t), intent(in)         :: desc
            type(eafp_pp_fields_t), intent(in)    :: fields
            type(eafp_pp_config_t), intent(in)    :: cfg
            type(eafp_pp_step_opts_t), intent(in) :: opts
            type(c_ptr), value                    :: pp
            type(eafp_pp_update_t), intent(in)    :: io
            real(c_double), intent(out)           :: residual
            type(eafp_stats_t), intent(out)       :: stats
            integer(c_int), value                 :: arrays_on_device
            integer(c_int)                        :: rc
        end function eafp_pp_step

        function eafp_cg_solve(desc, aW, aE, aS, aN, aB, aT, aP, Su, phi, &
                               max_iter, tol, residual, n_iter) &
                               bind(C, name="eafp_cg_solve") result(rc)
            import :: c_int, c_double, eafp_desc_t
            type(eafp_desc_t), intent(in) :: desc
            real(c_double), intent(in)    :: aW(*), aE(*), aS(*), aN(*)
            real(c_double), intent(in)    :: aB(*), aT(*), aP(*), Su(*)
            real(c_double), intent(inout) :: phi(*)
            integer(c_int), value         :: max_iter
            real(c_double), value         :: tol
            real(c_double), intent(out)   :: residual
            integer(c_int), intent(out)   :: n_iter
            integer(c_int)                :: rc
        end function eafp_cg_solve

        function eafp_tdma_solve(desc, aW, aE, aS, aN, aB, aT, aP, Su, phi, &
                                 n_sweep, arrays_on_device) &
                                 bind(C, name="eafp_tdma_solve") result(rc)
            import :: c_int, c_double, eafp_desc_t
            type(eafp_desc_t), intent(in) :: desc
            real(c_double), intent(in)    :: aW(*), aE(*), aS(*), aN(*)
            real(c_double), intent(in)    :: aB(*), aT(*), aP(*), Su(*)
            real(c_double), intent(inout) :: phi(*)
            integer(c_int), value         :: n_sweep, arrays_on_device
            integer(c_int)                :: rc
        end function eafp_tdma_solve
    end interface

    integer, save :: last_status = 0

contains

    subroutine cg_3d_mpi(aW, aE, aS, aN, aB, aT, aP, Su, phi, m, &
                         max_iter, tol, residual, n_iter)
        real(c_double), intent(in)    :: aW(-1:,-1:,-1:), aE(-1:,-1:,-1:)
        real(c_double), intent(in)    :: aS(-1:,-1:,-1:), aN(-1:,-1:,-1:)
        real(c_double), intent(in)    :: aB(-1:,-1:,-1:), aT(-1:,-1:,-1:)
        real(c_double), intent(in)    :: aP(-1:,-1:,-1:), Su(-1:,-1:,-1:)
        real(c_double), intent(inout) :: phi(-1:,-1:,-1:)
        type(mesh_t), intent(in)      :: m
        integer, intent(in)           :: max_iter
        real(c_double), intent(in)    :: tol
        real(c_double), intent(out)   :: residual
        integer, intent(out)          :: n_iter

        type(eafp_desc_t) :: d
        integer(c_int)    :: rc, nit
        real(c_double)    :: res

        d%nr = m%nr
        d%ntheta = m%ntheta
        d%nz = m%nz
        d%halo = 2
        d%lbound = -1
        d%is_parallel = merge(1_c_int, 0_c_int, m%is_parallel)
        d%periodic_theta = merge(1_c_int, 0_c_int, m%topo%periodic_theta)

        ! every array must have the (-1:n+2)^3 extent the C side assumes
        if (any(shape(phi) /= [m%nr + 4, m%ntheta + 4, m%nz + 4]) .or. &
            any(shape(aP) /= shape(phi)) .or. any(shape(Su) /= shape(phi)) .or. &
            any(shape(aW) /= shape(phi)) .or. any(shape(aE) /= shape(phi)) .or. &
            any(shape(aS) /= shape(phi)) .or. any(shape(aN) /= shape(phi)) .or. &
            any(shape(aB) /= shape(phi)) .or. any(shape(aT) /= shape(phi))) then
            last_status = -1
            residual = ieee_value(residual, ieee_quiet_nan)
            n_iter = 0
            return
        end if

        res = 0.0_c_double
        nit = 0
        rc = eafp_cg_solve(d, aW, aE, aS, aN, aB, aT, aP, Su, phi, &
                           int(max_iter, c_int), tol, res, nit)
        last_status = rc
        if (rc /= 0) then
            residual = ieee_value(residual, ieee_quiet_nan)
            n_iter = 0
        else
            residual = res
            n_iter = nit
        end if
    end subroutine cg_3d_mpi

    subroutine tdma_3d_mpi(aW, aE, aS, aN, aB, aT, aP, Su, phi, m, n_sweep)
        real(c_double), intent(in)    :: aW(-1:,-1:,-1:), aE(-1:,-1:,-1:)
        real(c_double), intent(in)    :: aS(-1:,-1:,-1:), aN(-1:,-1:,-1:)
        real(c_double), intent(in)    :: aB(-1:,-1:,-1:), aT(-1:,-1:,-1:)
        real(c_double), intent(in)    :: aP(-1:,-1:,-1:), Su(-1:,-1:,-1:)
        real(c_double), intent(inout) :: phi(-1:,-1:,-1:)
        type(mesh_t), intent(in)      :: m
        integer, intent(in)           :: n_sweep

        type(eafp_desc_t) :: d
        integer(c_int)    :: rc

        d%nr = m%nr
        d%ntheta = m%ntheta
        d%nz = m%nz
        d%halo = 2
        d%lbound = -1
        d%is_parallel = merge(1_c_int, 0_c_int, m%is_parallel)
        d%periodic_theta = merge(1_c_int, 0_c_int, m%topo%periodic_theta)
        if (any(shape(phi) /= [m%nr + 4, m%ntheta + 4, m%nz + 4]) .or. &
            any(shape(aP) /= shape(phi)) .or. any(shape(Su) /= shape(phi)) .or. &
            any(shape(aW) /= shape(phi)) .or. any(shape(aE) /= shape(phi)) .or. &
            any(shape(aS) /= shape(phi)) .or. any(shape(aN) /= shape(phi)) .or. &
            any(shape(aB) /= shape(phi)) .or. any(shape(aT) /= shape(phi))) then
            last_status = -1
            error stop "tdma_3d_mpi: arrays are not (-1:n+2)^3"
        end if
        rc = eafp_tdma_solve(d, aW, aE, aS, aN, aB, aT, aP, Su, phi, int(n_sweep, c_int), 0_c_int)
        last_status = rc
        if (rc /= 0) then
            write(*, '(A,I0)') 'tdma_3d_mpi: eafp_tdma_solve failed, status ', rc
            error stop 1
        end if
    end subroutine tdma_3d_mpi

    ! The p' step of solve_pressure_correction on the GPU (eafp_pp_step).
    ! In: the mesh (r, theta, z, vol, face areas, cell_type, the top layer's
    ! electrode mask `outlet`), both phases' fields and momentum diagonals,
    ! the config values, alpha_p, max_iter (cfg%max_inner_pres) and tol
    ! (SOR_TOL_PRESSURE); optional workspace masks (absent: not valid).
    ! In/out: the velocities of both phases and p.  Out: pp, the
    ! conservative liquid fluxes and their link flags, the continuity
    ! residual, the solver's residual and iteration count.  solver (optional,
    ! default EAFP_STEP_FP64: cg_3d_mpi's algorithm) and precond (mixed only).
    subroutine eafp_pressure_step(m, r, theta, z, vol, Ar, Ath, Az, cell_type, outlet, &
                                  alpha_l, ur_l, uth_l, uz_l, aPr_l, aPth_l, aPz_l, rho_l, T_l, &
                                  alpha_g, ur_g, uth_g, uz_g, aPr_g, aPth_g, aPz_g, rho_g, T_old_g, p, &
                                  dt, rho_gas, T_ambient, beta_expansion, &
                                  gas_in_poisson, solve_multiphase, gas_compressibility, &
                                  alpha_p, max_iter, tol, pp, Fc_r, Fc_th, Fc_z, lk_r, lk_th, lk_z, &
                                  residual, cg_res, n_iter, &
                                  liq_cont, pv_active, gas_reach, liq_room, solver, precond)
        type(mesh_t), intent(in) :: m
        real(c_double), intent(in), target :: r(-1:), theta(-1:), z(-1:)
        real(c_double), intent(in), target :: vol(-1:,-1:,-1:), Ar(-1:,-1:,-1:), Ath(-1:,-1:,-1:), Az(-1:,-1:,-1:)
        integer(c_int), intent(in), target :: cell_type(-1:,-1:,-1:), outlet(-1:,-1:)
        real(c_double), intent(in), target :: alpha_l(-1:,-1:,-1:), aPr_l(-1:,-1:,-1:), aPth_l(-1:,-1:,-1:)
        real(c_double), intent(in), target :: aPz_l(-1:,-1:,-1:), rho_l(-1:,-1:,-1:), T_l(-1:,-1:,-1:)
        real(c_double), intent(in), target :: alpha_g(-1:,-1:,-1:), aPr_g(-1:,-1:,-1:), aPth_g(-1:,-1:,-1:)
        real(c_double), intent(in), target :: aPz_g(-1:,-1:,-1:), rho_g(-1:,-1:,-1:), T_old_g(-1:,-1:,-1:)
        real(c_double), intent(inout), target :: ur_l(-1:,-1:,-1:), uth_l(-1:,-1:,-1:), uz_l(-1:,-1:,-1:)
        real(c_double), intent(inout), target :: ur_g(-1:,-1:,-1:), uth_g(-1:,-1:,-1:), uz_g(-1:,-1:,-1:)
        real(c_double), intent(inout), target :: p(-1:,-1:,-1:)
        real(c_double), intent(in) :: dt, rho_gas, T_ambient, beta_expansion, alpha_p, tol
        logical, intent(in) :: gas_in_poisson, solve_multiphase, gas_compressibility
        integer, intent(in) :: max_iter
        real(c_double), intent(out), target :: pp(-1:,-1:,-1:), Fc_r(-1:,-1:,-1:), Fc_th(-1:,-1:,-1:)
        real(c_double), intent(out), target :: Fc_z(-1:,-1:,-1:)
        integer(c_int), intent(out), target :: lk_r(-1:,-1:,-1:), lk_th(-1:,-1:,-1:), lk_z(-1:,-1:,-1:)
        real(c_double), intent(out) :: residual, cg_res
        integer, intent(out) :: n_iter
        integer(c_int), intent(in), target, optional :: liq_cont(-1:,-1:,-1:), pv_active(-1:,-1:,-1:)
        integer(c_int), intent(in), target, optional :: gas_reach(-1:,-1:,-1:)
        real(c_double), intent(in), target, optional :: liq_room(-1:,-1:,-1:)
        integer(c_int), intent(in), optional :: solver, precond

        type(eafp_desc_t)         :: d
        type(eafp_pp_fields_t)    :: f
        type(eafp_pp_config_t)    :: c
        type(eafp_pp_update_t)    :: io
        type(eafp_pp_step_opts_t) :: o
        type(eafp_stats_t)        :: st
        real(c_double)            :: res
        integer(c_int)            :: rc
        integer                   :: sh(3)
        logical                   :: bad

        d%nr = m%nr
        d%ntheta = m%ntheta
        d%nz = m%nz
        d%halo = 2
        d%lbound = -1
        d%is_parallel = merge(1_c_int, 0_c_int, m%is_parallel)
        d%periodic_theta = merge(1_c_int, 0_c_int, m%topo%periodic_theta)

        residual = ieee_value(residual, ieee_quiet_nan)
        cg_res = residual
        n_iter = 0
        sh = [m%nr + 4, m%ntheta + 4, m%nz + 4]
        ! every array against the padded shape the library copies from or to
        if (size(r) /= sh(1) .or. size(theta) /= sh(2) .or. size(z) /= sh(3) .or. &
            any(shape(outlet) /= sh(1:2)) .or. any(shape(cell_type) /= sh) .or. &
            any(shape(vol) /= sh) .or. any(shape(Ar) /= sh) .or. any(shape(Ath) /= sh) .or. &
            any(shape(Az) /= sh) .or. any(shape(p) /= sh) .or. any(shape(pp) /= sh) .or. &
            any(shape(alpha_l) /= sh) .or. any(shape(aPr_l) /= sh) .or. any(shape(aPth_l) /= sh) .or. &
            any(shape(aPz_l) /= sh) .or. any(shape(rho_l) /= sh) .or. any(shape(T_l) /= sh) .or. &
            any(shape(alpha_g) /= sh) .or. any(shape(aPr_g) /= sh) .or. any(shape(aPth_g) /= sh) .or. &
            any(shape(aPz_g) /= sh) .or. any(shape(rho_g) /= sh) .or. any(shape(T_old_g) /= sh) .or. &
            any(shape(ur_l) /= sh) .or. any(shape(uth_l) /= sh) .or. any(shape(uz_l) /= sh) .or. &
            any(shape(ur_g) /= sh) .or. any(shape(uth_g) /= sh) .or. any(shape(uz_g) /= sh) .or. &
            any(shape(Fc_r) /= sh) .or. any(shape(Fc_th) /= sh) .or. any(shape(Fc_z) /= sh) .or. &
            any(shape(lk_r) /= sh) .or. any(shape(lk_th) /= sh) .or. any(shape(lk_z) /= sh)) then
            last_status = -1
            return
        end if
        bad = .false.
        if (present(liq_cont)) bad = bad .or. any(shape(liq_cont) /= sh)
        if (present(pv_active)) bad = bad .or. any(shape(pv_active) /= sh)
        if (present(gas_reach)) bad = bad .or. any(shape(gas_reach) /= sh)
        if (present(liq_room)) bad = bad .or. any(shape(liq_room) /= sh)
        if (bad) then
            last_status = -1
            return
        end if

        f%r = c_loc(r)
        f%theta = c_loc(theta)
        f%z = c_loc(z)
        f%vol = c_loc(vol)
        f%Ar = c_loc(Ar)
        f%Ath = c_loc(Ath)
        f%Az = c_loc(Az)
        f%cell_type = c_loc(cell_type)
        f%outlet = c_loc(outlet)
        f%alpha_l = c_loc(alpha_l)
        f%ur_l = c_loc(ur_l)
        f%uth_l = c_loc(uth_l)
        f%uz_l = c_loc(uz_l)
        f%aPr_l = c_loc(aPr_l)
        f%aPth_l = c_loc(aPth_l)
        f%aPz_l = c_loc(aPz_l)
        f%rho_l = c_loc(rho_l)
        f%T_l = c_loc(T_l)
        f%alpha_g = c_loc(alpha_g)
        f%ur_g = c_loc(ur_g)
        f%uth_g = c_loc(uth_g)
        f%uz_g = c_loc(uz_g)
        f%aPr_g = c_loc(aPr_g)
        f%aPth_g = c_loc(aPth_g)
        f%aPz_g = c_loc(aPz_g)
        f%rho_g = c_loc(rho_g)
        f%T_old_g = c_loc(T_old_g)
        f%p = c_loc(p)
        f%liq_cont = c_null_ptr
        f%pv_active = c_null_ptr
        f%gas_reach = c_null_ptr
        f%liq_room = c_null_ptr
        if (present(liq_cont)) f%liq_cont = c_loc(liq_cont)
        if (present(pv_active)) f%pv_active = c_loc(pv_active)
        if (present(gas_reach)) f%gas_reach = c_loc(gas_reach)
        if (present(liq_room)) f%liq_room = c_loc(liq_room)

        c%dt = dt
        c%rho_gas = rho_gas
        c%T_ambient = T_ambient
        c%beta_expansion = beta_expansion
        c%gas_in_poisson = merge(1_c_int, 0_c_int, gas_in_poisson)
        c%solve_multiphase = merge(1_c_int, 0_c_int, solve_multiphase)
        c%gas_compressibility = merge(1_c_int, 0_c_int, gas_compressibility)
        c%pad_ = 0

        ! the velocities and p are updated where they lie (they are also the assembly's inputs)
        io%Fc_r = c_loc(Fc_r)
        io%Fc_th = c_loc(Fc_th)
        io%Fc_z = c_loc(Fc_z)
        io%lk_r = c_loc(lk_r)
        io%lk_th = c_loc(lk_th)
        io%lk_z = c_loc(lk_z)
        io%ur_l = c_loc(ur_l)
        io%uth_l = c_loc(uth_l)
        io%uz_l = c_loc(uz_l)
        io%ur_g = c_loc(ur_g)
        io%uth_g = c_loc(uth_g)
        io%uz_g = c_loc(uz_g)
        io%p = c_loc(p)

        o%solver = EAFP_STEP_FP64
        if (present(solver)) o%solver = solver
        o%max_iter = int(max_iter, c_int)
        o%tol = tol
        o%alpha_p = alpha_p
        o%mixed = eafp_mixed_opts_t(1.0e-8_c_double, 32_c_int, 2000_c_int, EAFP_COEF_BF16, 0_c_int, &
                                    EAFP_PRECOND_JACOBI)
        if (present(precond)) o%mixed%precond = precond

        res = 0.0_c_double
        rc = eafp_pp_step(d, f, c, o, c_loc(pp), io, res, st, 0_c_int)
        last_status = rc
        if (rc /= 0) return
        residual = res
        cg_res = st%residual
        n_iter = st%n_iter
    end subroutine eafp_pressure_step

    ! Status of the last cg_3d_mpi, tdma_3d_mpi or eafp_pressure_step call: 0, EAFP_ERR_* (< 0) or a HIP error (> 0).
    integer function eafp_last_status()
        eafp_last_status = last_status
    end function eafp_last_status

end module mod_eafp_solver
