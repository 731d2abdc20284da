! update_rank_boundaries (src/user_subroutines/SMOD_boundary_conditions.f90:17-56)
! on one rank: there is no halo to exchange, and the physical boundaries are
! filled on the device by the engine (codes 2 zero gradient, 3 linear
! extrapolation, 4 reflective wall; code 1, user-defined, is rejected when the
! engine is created, as the reference aborts on it).
submodule(MOD_domain) SMOD_boundary_conditions
    implicit none
contains
    module subroutine update_rank_boundaries(decomposition, mesh, solution)
        use MOD_decomposition, only: decomp_info
        use MOD_device, only: device_fill_ghosts
        type(decomp_info), intent(in) :: decomposition
        type(mesh_fields), intent(in) :: mesh
        type(solution_fields), intent(inout) :: solution
        call device_fill_ghosts()
    end subroutine
end submodule
