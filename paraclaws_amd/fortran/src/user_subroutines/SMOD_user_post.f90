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
    end subroutine
end submodule
