! The rank's engine handle and the host<->device transfers of the solution.
! The state lives on the GPU between steps; the host copy of
! solution%vars is refreshed only at start-up, snapshots and user hooks.
module MOD_device
    use iso_c_binding
    use MOD_pcw
    implicit none

    type(c_ptr), save :: handle = c_null_ptr
    integer, save :: m_xi_ = 0, m_eta_ = 0
    integer, save :: device_source_kind = 0   ! the built-in source the engine applies
    integer, parameter :: segment_attempts = 4096  ! attempts of one device-controlled run at most
    ! the device state's ghost cells are those of its interior: a step fills them
    ! (every engine, for the new and the rolled-back state alike), an upload of a
    ! host state does not
    logical, save :: ghosts_current = .false.

contains

    subroutine device_check(rc, what)
        use MOD_error_handling, only: abort_program
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc < 0) call abort_program(what // ": " // pcw_error_message(handle))
    end subroutine

    subroutine device_create(m_xi, m_eta, bc, xi_dimensions, eta_dimensions, grav_, &
            precision_, engine_, higher_order, use_limiter, limiter, source_kind, manning_n, dry_tol)
        !! source_kind, manning_n, dry_tol: the built-in device source of
        !! SMOD_device_physics (PCW_SRC_NONE, 0, 0: the reference's phi = 0)
        use MOD_constants, only: ud_precision
        integer, intent(in) :: m_xi, m_eta, bc(4), precision_, engine_, limiter, source_kind
        double precision, intent(in) :: xi_dimensions(2), eta_dimensions(2), grav_, manning_n, dry_tol
        logical, intent(in) :: higher_order, use_limiter
        type(pcw_params) :: p
        type(pcw_domain) :: d(1)
        integer(c_int) :: rc
        p%precision = precision_
        p%engine = engine_
        p%higher_order = merge(1, 0, higher_order)
        p%use_limiter = merge(1, 0, use_limiter)
        p%limiter = limiter
        p%tile_rows = 0
        p%waves_per_block = 0
        p%pad_ = 0
        p%grav = grav_
        p%eps = ud_precision
        p%dry_tol = dry_tol
        d(1)%nx = m_xi
        d(1)%ny = m_eta
        d(1)%bc = bc
        d(1)%source = source_kind
        d(1)%pad_ = 0
        d(1)%xi_min = xi_dimensions(1)
        d(1)%xi_max = xi_dimensions(2)
        d(1)%eta_min = eta_dimensions(1)
        d(1)%eta_max = eta_dimensions(2)
        d(1)%manning_n = manning_n
        rc = pcw_create(p, d, 1_c_int32_t, handle)
        call device_check(rc, "pcw_create")
        m_xi_ = m_xi
        m_eta_ = m_eta
        device_source_kind = source_kind
    end subroutine

    subroutine device_upload(vars)
        double precision, intent(in), contiguous, target :: vars(:,:,:)
        call device_check(pcw_set_state(handle, 0_c_int32_t, vars), "pcw_set_state")
        ghosts_current = .false.
    end subroutine

    subroutine device_download(vars)
        double precision, intent(inout), contiguous, target :: vars(:,:,:)
        call device_check(pcw_get_state(handle, 0_c_int32_t, vars), "pcw_get_state")
    end subroutine

    subroutine device_set_metrics(kappa, gamma_xi, gamma_eta)
        !! capacity and face length ratios of a mapped grid (frames incl. ghosts)
        double precision, intent(in), contiguous :: kappa(:,:), gamma_xi(:,:), gamma_eta(:,:)
        call device_check(pcw_set_metrics(handle, 0_c_int32_t, kappa, gamma_xi, gamma_eta), &
            "pcw_set_metrics")
    end subroutine

    subroutine device_set_bathymetry(b)
        !! bottom elevation b(ilow-2:ihigh+2, jlow-2:jhigh+2) of a bathymetry source
        double precision, intent(in), contiguous :: b(:,:)
        call device_check(pcw_set_bathymetry(handle, 0_c_int32_t, b), "pcw_set_bathymetry")
    end subroutine

    subroutine device_fill_ghosts()
        !! the physical ghost cells from the interior; after a step they are
        !! already filled, and the refill is skipped
        if (ghosts_current) return
        call device_check(pcw_fill_ghosts(handle), "pcw_fill_ghosts")
        ghosts_current = .true.
    end subroutine

    subroutine device_step(dt, smax_xi, smax_eta, eta_failed)
        !! one attempt of both sweeps (+ source + ghost fill) into the pending
        !! state; eta_failed: the eta sweep met a NaN Jacobian or complex
        !! eigenvalues (fatal only if the xi sweep alone accepts the step)
        double precision, intent(in) :: dt
        double precision, intent(out) :: smax_xi, smax_eta
        logical, intent(out) :: eta_failed
        real(c_double) :: dts(1), smax(2)
        integer(c_int) :: rc
        dts(1) = dt
        rc = pcw_step(handle, dts, smax)
        eta_failed = rc == PCW_ERR_STATE_ETA
        if (.not. eta_failed) call device_check(rc, "pcw_step")
        smax_xi = smax(1)
        smax_eta = smax(2)
    end subroutine

    subroutine device_commit(keep)
        logical, intent(in) :: keep
        if (keep) then
            call device_check(pcw_accept(handle), "pcw_accept")
        else
            call device_check(pcw_reject(handle), "pcw_reject")
        end if
        ghosts_current = .true.
    end subroutine

    subroutine device_autotune()
        !! Tile shape of the step kernels for this domain (pcw_autotune: a few
        !! timed attempts per candidate that are never kept; the state and the
        !! run are unchanged).  Only the fused engine has tiles.
        use MOD_solver_parameters, only: device_engine
        integer(c_int32_t) :: rows, waves
        if (device_engine /= PCW_ENGINE_FUSED) return
        call device_check(pcw_autotune(handle, 3_c_int32_t, rows, waves), "pcw_autotune")
        write(*, '(a,i0,a,i0,a)') ' device tiles: ', rows, ' rows, ', waves, ' waves per block'
    end subroutine

    subroutine device_loop_start(t_ctrl)
        !! The time loop on the device: a copy of the controller (pcw_time_init
        !! with control.in's settings, at the initial time) that stops the run
        !! after every step reaching a snapshot time.
        use MOD_time_control, only: time_control
        type(time_control), intent(inout) :: t_ctrl
        type(pcw_time) :: t
        double precision :: dt_max, dt_min
        integer :: n_snapshots
        call t_ctrl%get_limits(dt_max, dt_min, n_snapshots)
        t%t0 = t_ctrl%time_initial
        t%t_final = t_ctrl%time_final
        t%dt0 = t_ctrl%get_dt()
        t%cfl_target = t_ctrl%CFL_target
        t%dt_max = dt_max
        t%dt_min = dt_min
        t%growth = t_ctrl%timestep_growth_limit
        t%shrink = t_ctrl%timestep_shrink_limit
        t%n_snapshots = n_snapshots
        t%adaptive = merge(1, 0, t_ctrl%adaptive_timestepping)
        call device_check(pcw_time_init(handle, t), "pcw_time_init")
        call device_check(pcw_stop_at_snapshots(handle, 1_c_int32_t), "pcw_stop_at_snapshots")
    end subroutine

    subroutine device_loop_segment(recs, n)
        !! One device-controlled run: graph-launched attempts until the run
        !! stops (a snapshot or the final time reached) or segment_attempts
        !! were made; recs(1:n) are its attempts in order.
        type(pcw_step_record), intent(out) :: recs(segment_attempts)
        integer, intent(out) :: n
        integer(c_int32_t) :: m
        call device_check(pcw_enable_log(handle, int(segment_attempts, c_int32_t)), "pcw_enable_log")
        call device_check(pcw_run(handle, 32_c_int32_t, int(segment_attempts, c_int64_t), 0_c_int64_t), &
            "pcw_run")
        ghosts_current = .true.
        call device_check(pcw_read_log(handle, 0_c_int32_t, recs, int(segment_attempts, c_int32_t), m), &
            "pcw_read_log")
        n = m
    end subroutine

    function device_time() result(t)
        !! simulated time of the device controller
        double precision :: t
        type(pcw_status) :: st
        call device_check(pcw_get_status(handle, 0_c_int32_t, st), "pcw_get_status")
        t = st%time
    end function

    subroutine device_totals(out)
        double precision, intent(out) :: out(3)
        call device_check(pcw_totals(handle, 0_c_int32_t, out), "pcw_totals")
    end subroutine

    subroutine device_destroy()
        integer(c_int) :: rc
        if (c_associated(handle)) rc = pcw_destroy(handle)
        handle = c_null_ptr
    end subroutine

end module MOD_device
