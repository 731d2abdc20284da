! Circular dam break (examples/2D_dam_break/cartesian/SMOD_set_initial_conditions.f90
! of the reference): h = 2 inside radius 0.5, 1 outside, at rest, over the
! whole frame including ghosts.  Replaces the base submodule of the same name.
submodule(MOD_user_subroutines) SMOD_set_initial_conditions
    implicit none
contains
    module subroutine set_initial_conditions(decomposition, mesh, solution)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: solution_fields, mesh_fields
        use MOD_solver_parameters, only: num_ghost
        type(decomp_info), intent(in) :: decomposition
        type(mesh_fields), intent(in) :: mesh
        type(solution_fields), intent(inout) :: solution
        integer :: i, j, ilow, ihigh, jlow, jhigh
        call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
        do j = jlow-num_ghost, jhigh+num_ghost
            do i = ilow-num_ghost, ihigh+num_ghost
                if (sqrt(mesh%physical_space%x(i,j)**2 + mesh%physical_space%y(i,j)**2) < 0.5d0) then
                    solution%vars(1,i,j) = 2.d0
                else
                    solution%vars(1,i,j) = 1d0
                end if
                solution%vars(2,i,j) = 0.0d0
                solution%vars(3,i,j) = 0.0d0
            end do
        end do
    end subroutine
end submodule
