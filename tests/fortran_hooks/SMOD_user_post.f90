! Test case of the host's hooks: user_post raises h by 1e-3 in every interior
! cell each time execute_post runs (start-up and snapshots), so the run shows
! whether the changes reach the device state.
submodule(MOD_user_subroutines) SMOD_user_post
    implicit none
contains
    module subroutine user_post(i, j, mesh, solution, decomposition)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: mesh_fields, solution_fields
        integer, intent(in) :: i, j
        type(mesh_fields), intent(in) :: mesh
        type(solution_fields), intent(inout) :: solution
        type(decomp_info), intent(in) :: decomposition
        solution%vars(1,i,j) = solution%vars(1,i,j) + 1.d-3
    end subroutine
end submodule
