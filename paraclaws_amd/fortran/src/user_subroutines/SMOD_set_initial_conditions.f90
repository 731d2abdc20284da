submodule(MOD_user_subroutines) SMOD_set_initial_conditions
    implicit none
contains
    module subroutine set_initial_conditions(decomposition, mesh, solution)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: solution_fields, mesh_fields
        type(decomp_info), intent(in) :: decomposition
        type(mesh_fields), intent(in) :: mesh
        type(solution_fields), intent(inout) :: solution
        solution%vars = 0.0d0
    end subroutine
end submodule
