! User hooks (src/user_subroutines/MOD_user_subroutines.f90).  A USER_DIR
! override of any submodule below replaces the base file, as in the reference.
! source_function and user_post are per-cell host callbacks: the GPU path
! cannot call them per step, so a nonzero source_function override is
! detected at start-up and reported (built-in device sources are selected in
! custom.in instead).
module MOD_user_subroutines
    implicit none

    interface
        module subroutine setup_user_input()
        end subroutine

        module subroutine after_step(decomposition, t_ctrl, mesh, solution)
            use MOD_decomposition, only: decomp_info
            use MOD_time_control, only: time_control
            use MOD_domain, only: solution_fields, mesh_fields
            type(decomp_info), intent(in) :: decomposition
            type(time_control), intent(in) :: t_ctrl
            type(mesh_fields), intent(inout) :: mesh
            type(solution_fields), intent(inout) :: solution
        end subroutine

        module subroutine setup_auxiliary_quantities(decomposition, mesh, solution)
            use MOD_decomposition, only: decomp_info
            use MOD_domain, only: solution_fields, mesh_fields
            type(decomp_info), intent(in) :: decomposition
            type(mesh_fields), intent(in) :: mesh
            type(solution_fields), intent(inout) :: solution
        end subroutine

        module subroutine set_initial_conditions(decomposition, mesh, solution)
            use MOD_decomposition, only: decomp_info
            use MOD_domain, only: solution_fields, mesh_fields
            type(decomp_info), intent(in) :: decomposition
            type(mesh_fields), intent(in) :: mesh
            type(solution_fields), intent(inout) :: solution
        end subroutine

        module subroutine user_post(i, j, mesh, solution, decomposition)
            use MOD_decomposition, only: decomp_info
            use MOD_domain, only: mesh_fields, solution_fields
            integer, intent(in) :: i, j
            type(mesh_fields), intent(in) :: mesh
            type(solution_fields), intent(inout) :: solution
            type(decomp_info), intent(in) :: decomposition
        end subroutine

        module function source_function(i, j, mesh, solution, decomposition) result(phi)
            use MOD_decomposition, only: decomp_info
            use MOD_domain, only: mesh_fields, solution_fields
            type(mesh_fields), intent(in) :: mesh
            type(solution_fields), intent(in) :: solution
            type(decomp_info), intent(in) :: decomposition
            integer :: i, j
            double precision :: phi(3)
        end function
    end interface

end module MOD_user_subroutines
