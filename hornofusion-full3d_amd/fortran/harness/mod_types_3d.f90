!===============================================================================
! Test harness only: the mesh fields that cg_3d_mpi reads, under the module
! and type names of the reference (src/mod_types_3d.f90:143-193, with the
! topology fields of src/mod_mpi_topology.f90), so that mod_eafp_solver
! compiles unchanged against this file here and against the reference's own
! mod_types_3d inside the reference build.
!===============================================================================
module mod_types_3d
    implicit none
    integer, parameter :: dp = kind(1.0d0)

    type :: mpi_topology_t
        integer :: nprocs = 1
        logical :: periodic_theta = .true.
        integer :: jloc = 0
    end type mpi_topology_t

    type :: mesh_t
        logical :: is_parallel = .false.
        type(mpi_topology_t) :: topo
        integer :: nr = 0, ntheta = 0, nz = 0
    end type mesh_t
end module mod_types_3d
