! Test case of the host's hooks: after_step sets hu to zero on the whole frame
! after the fifth accepted step (it reads the step count it keeps itself).
submodule(MOD_user_subroutines) SMOD_setup
    implicit none
contains
    module subroutine setup_user_input()
        use MOD_solver_parameters, only: grav
        grav = 9.81   ! a default-kind literal, as in the reference's setup
    end subroutine

    module subroutine after_step(decomposition, t_ctrl, mesh, solution)
        use MOD_decomposition, only: decomp_info
        use MOD_time_control, only: time_control
        use MOD_domain, only: solution_fields, mesh_fields
        type(decomp_info), intent(in) :: decomposition
        type(time_control), intent(in) :: t_ctrl
        type(mesh_fields), intent(inout) :: mesh
        type(solution_fields), intent(inout) :: solution
        integer, save :: calls = 0
        calls = calls + 1
        if (calls == 5) solution%vars(2,:,:) = 0.d0
    end subroutine

    module subroutine setup_auxiliary_quantities(decomposition, mesh, solution)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: solution_fields, mesh_fields
        type(decomp_info), intent(in) :: decomposition
        type(mesh_fields), intent(in) :: mesh
        type(solution_fields), intent(inout) :: solution
    end subroutine
end submodule
