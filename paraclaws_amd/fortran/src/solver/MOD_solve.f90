! The per-step numerics of the reference (src/solver/MOD_solve.f90) driven on
! the GPU.  solve_dimensional_split keeps the reference's bookkeeping: the
! CFL of each sweep is max(lambda*dt/delta) over the sweep's faces (the
! engine returns the max wave speed, so the CFL is exactly
! smax*dt/delta), the validity test follows each sweep, and ierr is 1 for an
! accepted and -1 for a rejected step.  Both sweeps, the source term and the
! ghost fill run in one device step into a pending state that the caller keeps
! (device_commit(.true.)) or drops (device_commit(.false.)).
module MOD_solve
    implicit none
    public :: solve_dimensional_split, integrate_source_term, execute_post
contains

    subroutine solve_dimensional_split(decomposition, t_ctrl, work, mesh, solution, ierr)
        use MOD_time_control, only: time_control
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: solution_fields, mesh_fields
        use MOD_work_arrays, only: work_arrays
        use MOD_device, only: device_step
        use MOD_error_handling, only: abort_program
        type(decomp_info), intent(in) :: decomposition
        type(time_control), intent(inout) :: t_ctrl
        type(work_arrays), intent(inout) :: work
        type(mesh_fields) :: mesh
        type(solution_fields) :: solution
        integer, intent(inout) :: ierr
        double precision :: smax_xi, smax_eta
        logical :: step_valid, eta_failed

        ierr = -1
        call device_step(t_ctrl%get_dt(), smax_xi, smax_eta, eta_failed)

        ! xi sweep (MOD_solve.f90:71-120)
        call t_ctrl%update_max_CFL(smax_xi*t_ctrl%get_dt()/mesh%computational_space%d_xi)
        step_valid = t_ctrl%check_timestep_validity()
        if (.not. step_valid) return
        ! the eta sweep runs only now in the reference: its NaN/complex check is fatal
        if (eta_failed) call abort_program("NaN in the Jacobian or complex eigenvalues (eta sweep)")

        ! eta sweep (MOD_solve.f90:124-174)
        call t_ctrl%update_max_CFL(smax_eta*t_ctrl%get_dt()/mesh%computational_space%d_eta)
        step_valid = t_ctrl%check_timestep_validity()
        if (step_valid) ierr = 1
    end subroutine

    subroutine integrate_source_term(decomposition, t_ctrl, mesh, solution)
        !! The source term is integrated inside the device step (forward Euler with
        !! the dt of the sweeps, after them, as MOD_solve.f90:295-336 does).
        use MOD_time_control, only: time_control
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: mesh_fields, solution_fields
        type(decomp_info), intent(in) :: decomposition
        type(time_control), intent(in) :: t_ctrl
        type(mesh_fields), intent(in) :: mesh
        type(solution_fields), intent(inout) :: solution
    end subroutine

    subroutine execute_post(decomposition, mesh, solution)
        !! user_post on every interior cell of the host copy (MOD_solve.f90:338-363)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: mesh_fields, solution_fields
        use MOD_user_subroutines, only: user_post
        type(decomp_info), intent(in) :: decomposition
        type(mesh_fields), intent(in) :: mesh
        type(solution_fields), intent(inout) :: solution
        integer :: ilow, ihigh, jlow, jhigh, i, j
        call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
        do j = jlow, jhigh
            do i = ilow, ihigh
                call user_post(i, j, mesh, solution, decomposition)
            end do
        end do
    end subroutine

end module MOD_solve
