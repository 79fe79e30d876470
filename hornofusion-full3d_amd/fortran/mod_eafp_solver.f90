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
! In the reference build this module replaces the `use mod_solver_3d` of
! src/mod_pressure_3d.f90:28 for cg_3d_mpi, and of the TDMA callers
! (src/mod_momentum_3d.f90:329, src/mod_energy_3d.f90:320, ...) for
! tdma_3d_mpi; mesh_t comes from mod_types_3d.
!===============================================================================
module mod_eafp_solver
    use, intrinsic :: iso_c_binding, only: c_int, c_double, c_ptr, c_loc
    use, intrinsic :: ieee_arithmetic, only: ieee_value, ieee_quiet_nan
    use mod_types_3d, only: mesh_t
    implicit none
    private
    public :: cg_3d_mpi, tdma_3d_mpi, eafp_last_status

    ! include/eafp.h: eafp_desc
    type, bind(C) :: eafp_desc_t
        integer(c_int) :: nr, ntheta, nz
        integer(c_int) :: halo, lbound
        integer(c_int) :: is_parallel, periodic_theta
    end type eafp_desc_t

    interface
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

    ! Status of the last cg_3d_mpi or tdma_3d_mpi call: 0, EAFP_ERR_* (< 0) or a HIP error (> 0).
    integer function eafp_last_status()
        eafp_last_status = last_status
    end function eafp_last_status

end module mod_eafp_solver
