! User hooks (src/user_subroutines/MOD_user_subroutines.f90).  A USER_DIR
! override of any submodule below replaces the base file, as in the reference.
!
! source_function is a per-cell host callback that the GPU step cannot call.
! The device instead applies a built-in source that the submodule
! SMOD_device_physics selects (device_physics, device_bathymetry); its base
! file selects the reference's phi = 0.  With no device source selected, a
! source_function that is nonzero anywhere on the initial state is rejected
! at start-up (main program).  user_post runs through execute_post on the
! host copy of the state at snapshots, and the changes go back to the device;
! after_step gets the current state, and its changes go back, when
! host_state_hooks (SMOD_device_physics) returns .true. (the base returns
! .false.: no per-step transfer).
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

        ! ---- the GPU path's settings (submodule SMOD_device_physics) ----
        module subroutine device_physics(source_kind, manning_n, dry_tol)
            !! built-in device source that replaces source_function:
            !! source_kind 0 none (the reference's phi = 0), 1 bathymetry slope,
            !! 2 Manning friction, 3 both; manning_n: Manning coefficient;
            !! dry_tol: wet/dry threshold on h (0: off, the reference's behaviour)
            integer, intent(out) :: source_kind
            double precision, intent(out) :: manning_n, dry_tol
        end subroutine

        module subroutine device_bathymetry(decomposition, mesh, b)
            !! bottom elevation b(i, j) of the frame cells i = -1 .. nx+2,
            !! j = -1 .. ny+2, for source kinds 1 and 3
            use MOD_decomposition, only: decomp_info
            use MOD_domain, only: mesh_fields
            type(decomp_info), intent(in) :: decomposition
            type(mesh_fields), intent(in) :: mesh
            double precision, intent(inout) :: b(-1:, -1:)
        end subroutine

        module function host_state_hooks() result(needs)
            !! .true. when after_step reads or changes solution%vars: the host
            !! then fetches the state before each after_step and sends it back
            logical :: needs
        end function
    end interface

end module MOD_user_subroutines
