! iso_c_binding interfaces of the gfx950 engine (include/pcw.h).
!
! The derived types below are interoperable mirrors of pcw_params,
! pcw_domain, pcw_time, pcw_step_record and pcw_status; every entry point returns the C status code
! (0 ok, 1 rejected, < 0 fatal).
module MOD_pcw
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: PCW_OK = 0, PCW_REJECTED = 1, PCW_ERR_STATE_ETA = -6
    integer(c_int), parameter :: PCW_FP64 = 0, PCW_FP32 = 1
    integer(c_int), parameter :: PCW_ENGINE_FUSED = 0, PCW_ENGINE_STAGED = 1
    integer(c_int), parameter :: PCW_SRC_NONE = 0, PCW_SRC_BATHY = 1, PCW_SRC_FRICTION = 2, &
        PCW_SRC_BATHY_FRICTION = 3

    type, bind(C) :: pcw_domain
        integer(c_int32_t) :: nx, ny
        integer(c_int32_t) :: bc(4)
        integer(c_int32_t) :: source
        integer(c_int32_t) :: pad_
        real(c_double) :: xi_min, xi_max, eta_min, eta_max
        real(c_double) :: manning_n
    end type

    type, bind(C) :: pcw_params
        integer(c_int32_t) :: precision, engine
        integer(c_int32_t) :: higher_order, use_limiter, limiter
        integer(c_int32_t) :: tile_rows, waves_per_block, pad_
        real(c_double) :: grav, eps, dry_tol
    end type

    type, bind(C) :: pcw_time
        real(c_double) :: t0, t_final, dt0, cfl_target, dt_max, dt_min, growth, shrink
        integer(c_int32_t) :: n_snapshots
        integer(c_int32_t) :: adaptive
    end type

    type, bind(C) :: pcw_step_record
        integer(c_int32_t) :: accepted
        integer(c_int32_t) :: snapshot
        real(c_double) :: dt_used, cfl, time, dt_next
    end type

    type, bind(C) :: pcw_status
        integer(c_int32_t) :: status
        integer(c_int32_t) :: error
        integer(c_int64_t) :: accepted, rejected
        real(c_double) :: time, dt, cfl_last
    end type

    interface
        function pcw_create(p, domains, n, h) bind(C, name="pcw_create") result(rc)
            import :: c_int, c_int32_t, c_ptr, pcw_params, pcw_domain
            type(pcw_params), intent(in) :: p
            type(pcw_domain), intent(in) :: domains(*)
            integer(c_int32_t), value :: n
            type(c_ptr), intent(out) :: h
            integer(c_int) :: rc
        end function

        function pcw_destroy(h) bind(C, name="pcw_destroy") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int) :: rc
        end function

        function pcw_last_error(h) bind(C, name="pcw_last_error") result(msg)
            import :: c_ptr
            type(c_ptr), value :: h
            type(c_ptr) :: msg
        end function

        function pcw_set_state(h, inst, vars) bind(C, name="pcw_set_state") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: inst
            real(c_double), intent(in) :: vars(*)
            integer(c_int) :: rc
        end function

        function pcw_set_bathymetry(h, inst, b) bind(C, name="pcw_set_bathymetry") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: inst
            real(c_double), intent(in) :: b(*)
            integer(c_int) :: rc
        end function

        function pcw_get_state(h, inst, vars) bind(C, name="pcw_get_state") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: inst
            real(c_double), intent(out) :: vars(*)
            integer(c_int) :: rc
        end function

        function pcw_set_metrics(h, inst, kappa, gamma_xi, gamma_eta) &
                bind(C, name="pcw_set_metrics") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: inst
            real(c_double), intent(in) :: kappa(*), gamma_xi(*), gamma_eta(*)
            integer(c_int) :: rc
        end function

        function pcw_step(h, dt, smax) bind(C, name="pcw_step") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: dt(*)
            real(c_double), intent(out) :: smax(*)
            integer(c_int) :: rc
        end function

        function pcw_accept(h) bind(C, name="pcw_accept") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int) :: rc
        end function

        function pcw_reject(h) bind(C, name="pcw_reject") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int) :: rc
        end function

        function pcw_fill_ghosts(h) bind(C, name="pcw_fill_ghosts") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int) :: rc
        end function

        function pcw_totals(h, inst, out) bind(C, name="pcw_totals") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: inst
            real(c_double), intent(out) :: out(3)
            integer(c_int) :: rc
        end function

        function pcw_time_init(h, t) bind(C, name="pcw_time_init") result(rc)
            import :: c_int, c_ptr, pcw_time
            type(c_ptr), value :: h
            type(pcw_time), intent(in) :: t
            integer(c_int) :: rc
        end function

        function pcw_enable_log(h, capacity) bind(C, name="pcw_enable_log") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: capacity
            integer(c_int) :: rc
        end function

        function pcw_read_log(h, inst, out, cap, n) bind(C, name="pcw_read_log") result(rc)
            import :: c_int, c_int32_t, c_ptr, pcw_step_record
            type(c_ptr), value :: h
            integer(c_int32_t), value :: inst
            type(pcw_step_record), intent(out) :: out(*)
            integer(c_int32_t), value :: cap
            integer(c_int32_t), intent(out) :: n
            integer(c_int) :: rc
        end function

        function pcw_run(h, steps_per_graph, max_attempts, max_accepted) bind(C, name="pcw_run") &
                result(rc)
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: steps_per_graph
            integer(c_int64_t), value :: max_attempts, max_accepted
            integer(c_int) :: rc
        end function

        function pcw_stop_at_snapshots(h, on) bind(C, name="pcw_stop_at_snapshots") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: on
            integer(c_int) :: rc
        end function

        function pcw_autotune(h, steps, tile_rows, waves_per_block) bind(C, name="pcw_autotune") &
                result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: steps
            integer(c_int32_t), intent(out) :: tile_rows, waves_per_block
            integer(c_int) :: rc
        end function

        function pcw_get_status(h, inst, out) bind(C, name="pcw_get_status") result(rc)
            import :: c_int, c_int32_t, c_ptr, pcw_status
            type(c_ptr), value :: h
            integer(c_int32_t), value :: inst
            type(pcw_status), intent(out) :: out
            integer(c_int) :: rc
        end function
    end interface

contains

    function pcw_error_message(h) result(msg)
        !! The handle's last error as a Fortran string.
        type(c_ptr), intent(in) :: h
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: n
        interface
            function c_strlen(s) bind(C, name="strlen") result(len)
                import :: c_ptr, c_size_t
                type(c_ptr), value :: s
                integer(c_size_t) :: len
            end function
        end interface
        p = pcw_last_error(h)
        if (.not. c_associated(p)) then
            msg = "unknown engine error"
            return
        end if
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        msg = transfer(chars(1:n), msg)
    end function

end module MOD_pcw
