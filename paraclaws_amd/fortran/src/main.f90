! Time loop of the reference (src/main.f90:109-301) on one rank with the state
! on the GPU.  Step structure kept: solve -> reject/rollback -> source ->
! boundary update -> updateTime -> snapshot output -> reset_CFL -> after_step.
program paraCLAWs_gpu
    use MOD_decomposition, only: decomp_info, initialize_decomposition
    use MOD_user_subroutines, only: setup_user_input, set_initial_conditions, &
        setup_auxiliary_quantities, after_step, source_function, host_state_hooks
    use MOD_time_control, only: time_control
    use MOD_domain, only: update_rank_boundaries, solution_fields, mesh_fields, write_base_info, &
        setup_mesh
    use MOD_solve, only: solve_dimensional_split, integrate_source_term, execute_post
    use MOD_IO
    use MOD_work_arrays, only: work_arrays
    use MOD_device, only: device_upload, device_download, device_commit, device_totals, &
        device_source_kind, device_loop_start, device_loop_segment, device_time, segment_attempts, &
        device_autotune
    use MOD_pcw, only: pcw_step_record
    use MOD_solver_parameters, only: device_loop
    use MOD_error_handling, only: abort_program
    implicit none

    integer :: ierr, exitstat, cmdstat
    type(time_control) :: t_ctrl
    type(decomp_info) :: decomposition
    type(mesh_fields) :: mesh
    type(solution_fields) :: solution
    type(work_arrays) :: work
    double precision :: totals(3)
    integer :: n_accepted, n_rejected
    integer :: count0, count1, count_rate
    integer :: i, j, k, ilow, ihigh, jlow, jhigh, nrec
    logical :: state_hooks, snapshot_due
    type(pcw_step_record), allocatable :: recs(:)

1   format('paraCLAWs - time: ',f10.8,' current dt: ',f10.8,' CFL(max)=',f10.5,' next dt: ',f10.8,' => snapshot: ', i4)
2   format('paraCLAWs - time: ',f10.8,' current dt: ',f10.8,' CFL(max)=',f10.5,' next dt: ',f10.8)
3   format('totals: mass ',es24.16,' momentum ',es24.16,1x,es24.16)

    write(*,*) "========================================================="
    write(*,*) "paraCLAWs on MI355X (gfx950 engine)"
    write(*,*) "========================================================="
    call execute_command_line('rm -rf output', .true., exitstat, cmdstat)
    call execute_command_line('mkdir -p output', .true., exitstat, cmdstat)

    call read_solver_control(t_ctrl, mesh, solution)
    call read_custom_settings()
    call initialize_decomposition(decomposition, mesh%m_xi, mesh%m_eta)
    call solution%allocate_solution(decomposition)
    call setup_user_input()
    call setup_mesh(mesh, decomposition)
    call work%allocate_work_arrays(decomposition, mesh)
    call work%initialize_work()

    call setup_auxiliary_quantities(decomposition, mesh, solution)
    call set_initial_conditions(decomposition, mesh, solution)
    ! the device cannot call source_function per cell (src/solver/MOD_solve.f90:
    ! 319-327); without a device source selected in SMOD_device_physics, a user
    ! source_function that is nonzero on the initial state is rejected
    if (device_source_kind == 0) then
        call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
        do j = jlow, jhigh
            do i = ilow, ihigh
                if (any(source_function(i, j, mesh, solution, decomposition) /= 0.d0)) &
                    call abort_program("source_function returns a nonzero source, which the GPU step " // &
                        "cannot evaluate per cell: select a built-in device source in SMOD_device_physics")
            end do
        end do
    end if
    state_hooks = host_state_hooks()
    call device_upload(solution%vars)
    call update_rank_boundaries(decomposition, mesh, solution)
    call device_download(solution%vars)
    call execute_post(decomposition, mesh, solution)
    write(*,1) t_ctrl%time, t_ctrl%get_previous_dt(), t_ctrl%get_CFL(), t_ctrl%get_dt(), 0

    call write_base_info(mesh)
    call write_mesh_to_disk(decomposition, mesh)
    call writeData(decomposition, t_ctrl%time, solution)
    call write_vtk_rank(decomposition, mesh, solution, t_ctrl%time)
    call write_pvts_rank(decomposition, t_ctrl%time)
    call device_upload(solution%vars)
    call update_rank_boundaries(decomposition, mesh, solution)
    call device_totals(totals)
    write(*,3) totals
    call device_autotune()

    n_accepted = 0
    n_rejected = 0
    if (device_loop == 1 .and. state_hooks) &
        print *, "custom.in asks for the device time loop, but after_step needs the state " // &
            "every step (host_state_hooks): the loop runs on the host"
    call system_clock(count0, count_rate)
    if (device_loop == 1 .and. .not. state_hooks) then
        ! The same loop with the steps on the device: graph-launched attempts
        ! under the device's copy of the controller, back on the host at every
        ! snapshot.  Each run's attempts are replayed through t_ctrl in order
        ! (their CFL: the same controller decisions, dt sequence and terminal
        ! output as the loop below), and the snapshot step is the run's last.
        allocate(recs(segment_attempts))
        call device_loop_start(t_ctrl)
        do while (t_ctrl%time < t_ctrl%time_final)
            call device_loop_segment(recs, nrec)
            if (nrec == 0) call abort_program("device time loop: a run made no attempt")
            do k = 1, nrec
                call t_ctrl%update_max_CFL(recs(k)%cfl)
                if (t_ctrl%check_timestep_validity() .neqv. (recs(k)%accepted /= 0)) &
                    call abort_program("device time loop: the host controller disagrees with the device's")
                if (recs(k)%accepted == 0) then
                    print *, "***************************************************"
                    print *, "WARNING: time step rejected!"
                    call t_ctrl%updateTime(step_valid=.false.)
                    call t_ctrl%reset_CFL()
                    print *, "***************************************************"
                    n_rejected = n_rejected + 1
                    cycle
                end if
                call t_ctrl%updateTime(step_valid=.true.)
                n_accepted = n_accepted + 1
                snapshot_due = t_ctrl%time >= t_ctrl%snapshot_intervals(t_ctrl%get_snapshot_index())
                if (snapshot_due .neqv. (recs(k)%snapshot > 0)) &
                    call abort_program("device time loop: snapshot steps of host and device differ")
                if (snapshot_due) then
                    if (k /= nrec) call abort_program("device time loop: the run went past a snapshot")
                    call device_download(solution%vars)
                    call execute_post(decomposition, mesh, solution)
                    call device_upload(solution%vars)     ! user_post's changes
                    call writeData(decomposition, t_ctrl%time, solution)
                    call write_vtk_rank(decomposition, mesh, solution, t_ctrl%time)
                    call write_pvts_rank(decomposition, t_ctrl%time)
                    write(*,1) t_ctrl%time, t_ctrl%get_previous_dt(), t_ctrl%get_CFL(), t_ctrl%get_dt(), &
                        t_ctrl%get_snapshot_index()
                    call t_ctrl%advance_snapshot_index()
                else
                    write(*,2) t_ctrl%time, t_ctrl%get_previous_dt(), t_ctrl%get_CFL(), t_ctrl%get_dt()
                end if
                call t_ctrl%reset_CFL()
                call after_step(decomposition, t_ctrl, mesh, solution)
            end do
            if (device_time() /= t_ctrl%time) &
                call abort_program("device time loop: device and host times differ")
        end do
    end if
    do while (t_ctrl%time < t_ctrl%time_final)
        call solve_dimensional_split(decomposition, t_ctrl, work, mesh, solution, ierr)
        if (ierr < 0) then
            print *, "***************************************************"
            print *, "WARNING: time step rejected!"
            call t_ctrl%updateTime(step_valid=.false.)
            call device_commit(.false.)          ! solution = solution_old
            call t_ctrl%reset_CFL()
            print *, "***************************************************"
            n_rejected = n_rejected + 1
            cycle
        end if
        call device_commit(.true.)
        call integrate_source_term(decomposition, t_ctrl, mesh, solution)
        call update_rank_boundaries(decomposition, mesh, solution)
        call t_ctrl%updateTime(step_valid=.true.)
        n_accepted = n_accepted + 1

        if (t_ctrl%time >= t_ctrl%snapshot_intervals(t_ctrl%get_snapshot_index())) then
            call device_download(solution%vars)
            call execute_post(decomposition, mesh, solution)
            call device_upload(solution%vars)     ! user_post's changes
            call writeData(decomposition, t_ctrl%time, solution)
            call write_vtk_rank(decomposition, mesh, solution, t_ctrl%time)
            call write_pvts_rank(decomposition, t_ctrl%time)
            write(*,1) t_ctrl%time, t_ctrl%get_previous_dt(), t_ctrl%get_CFL(), t_ctrl%get_dt(), &
                t_ctrl%get_snapshot_index()
            call t_ctrl%advance_snapshot_index()
        else
            write(*,2) t_ctrl%time, t_ctrl%get_previous_dt(), t_ctrl%get_CFL(), t_ctrl%get_dt()
        end if
        call t_ctrl%reset_CFL()
        if (state_hooks) call device_download(solution%vars)
        call after_step(decomposition, t_ctrl, mesh, solution)
        if (state_hooks) call device_upload(solution%vars)
    end do
    call system_clock(count1)

    call device_totals(totals)
    write(*,3) totals
    write(*,'(a,i0,a,i0,a,f10.4,a)') 'accepted steps: ', n_accepted, '  rejected: ', n_rejected, &
        '  wall: ', real(count1 - count0) / real(count_rate), ' s'
    call device_download(solution%vars)
    open(unit=9, file='output/final_state.bin', access='stream', form='unformatted', status='replace')
    write(9) solution%vars
    close(9)
    call work%deallocate_work_arrays()
    call solution%deallocate_solution()
    print *, "***************************************************"
    print *, "PROGRAM FINISHED SUCCESSFULLY"
    print *, "***************************************************"
end program paraCLAWs_gpu
