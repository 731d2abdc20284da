submodule(MOD_user_subroutines) SMOD_source_function
    implicit none
contains
    module function source_function(i, j, mesh, solution, decomposition) result(phi)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: mesh_fields, solution_fields
        type(mesh_fields), intent(in) :: mesh
        type(solution_fields), intent(in) :: solution
        type(decomp_info), intent(in) :: decomposition
        integer :: i, j
        double precision :: phi(3)
        phi = 0.0d0
    end function
end submodule
