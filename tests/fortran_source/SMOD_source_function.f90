! Test case: a user source_function that returns a nonzero source, with no
! device source selected; the host must refuse to run it.
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
        if (i == 3 .and. j == 2) phi(2) = 1.d-3
    end function
end submodule
