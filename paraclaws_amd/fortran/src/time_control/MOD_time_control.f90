! The time controller of the reference (src/time_control/MOD_time_control.f90):
! a-posteriori CFL control with reject/retry, snapshot alignment, and time
! advanced by the next dt.  FP64 scalar logic on the host.
module MOD_time_control
    implicit none

    type :: time_control
        double precision :: time
        double precision :: time_initial
        double precision :: time_final

        double precision, private :: dt
        double precision, private :: dt_previous
        double precision, private :: dt_max
        double precision, private :: dt_min

        double precision, allocatable :: snapshot_intervals(:)
        integer, private :: current_snapshot_index
        integer, private :: number_of_snapshots

        double precision :: time_precision
        double precision, private :: CFL
        double precision :: CFL_target
        double precision :: CFL_previous_timestep
        double precision :: timestep_shrink_limit
        double precision :: timestep_growth_limit
        double precision, private :: emergency_shrink_ratio = 0.25

        logical :: reset_timestep
        logical :: adaptive_timestepping
        logical :: snapshot_exception = .false.
    contains
        procedure :: initialize
        procedure :: updateTime
        procedure :: update_max_CFL
        procedure :: advance_snapshot_index
        procedure :: check_timestep_validity
        procedure :: reset_CFL
        procedure :: get_CFL
        procedure :: get_dt
        procedure :: get_previous_dt
        procedure :: get_snapshot_index
        procedure :: get_limits
    end type time_control

contains

    subroutine initialize(this, time_initial_, time_final_, timestep_, n_output_times_, CFL_target_, &
            timestep_max_, timestep_min_, growth_limit_, shrink_limit_, adaptive_timestepping_)
        use MOD_error_handling, only: end_program
        class(time_control), intent(inout) :: this
        double precision, intent(in) :: time_initial_, time_final_, timestep_
        double precision, intent(in) :: timestep_max_, timestep_min_
        double precision, intent(in) :: growth_limit_, shrink_limit_, CFL_target_
        integer, intent(in) :: n_output_times_
        logical, intent(in), optional :: adaptive_timestepping_
        double precision :: interval
        integer :: i

        this%time_initial = time_initial_
        this%time_final = time_final_
        this%time = time_initial_
        this%dt = timestep_
        this%dt_previous = timestep_
        this%dt_max = timestep_max_
        this%dt_min = timestep_min_
        this%adaptive_timestepping = .false.
        if (present(adaptive_timestepping_)) this%adaptive_timestepping = adaptive_timestepping_

        if (this%dt_min > this%dt_max) then
            write(*,*) "WARNING: Minimum timestep is greater than maximum timestep!"
            call end_program("ABORT: Invalid timestep bounds!")
        end if

        if (n_output_times_ < 1) then
            this%number_of_snapshots = 1
            write(*,*) "WARNING: number of output timestamps < 1! => changed to 1"
        else
            this%number_of_snapshots = n_output_times_
        end if

        if (allocated(this%snapshot_intervals)) deallocate(this%snapshot_intervals)
        allocate(this%snapshot_intervals(this%number_of_snapshots))
        interval = (this%time_final - this%time_initial) / this%number_of_snapshots
        do i = 1, this%number_of_snapshots
            this%snapshot_intervals(i) = this%time_initial + i * interval
        end do
        this%snapshot_intervals(this%number_of_snapshots) = this%time_final
        this%current_snapshot_index = 1

        this%CFL = 0.d0
        this%CFL_previous_timestep = 0.d0
        this%CFL_target = CFL_target_
        this%time_precision = 0.01d0
        this%timestep_growth_limit = growth_limit_
        this%timestep_shrink_limit = shrink_limit_
    end subroutine initialize

    subroutine updateTime(this, step_valid)
        use MOD_error_handling, only: end_program
        class(time_control), intent(inout) :: this
        logical, intent(in) :: step_valid
        double precision :: change_factor, dt_step

        this%dt_previous = this%dt
        dt_step = this%dt

        if (step_valid) then
            if (this%adaptive_timestepping) then
                if (this%CFL > 0.d0) then
                    change_factor = this%CFL_target / this%CFL
                else
                    change_factor = 1.d0
                end if
                if (change_factor >= this%timestep_growth_limit) then
                    change_factor = this%timestep_growth_limit
                elseif (abs(change_factor - 1.d0) <= this%time_precision) then
                    change_factor = 1.d0
                end if
                dt_step = this%dt_previous * change_factor
            end if
            if (this%time + dt_step > this%snapshot_intervals(this%current_snapshot_index)) then
                dt_step = this%snapshot_intervals(this%current_snapshot_index) - this%time
                this%snapshot_exception = .true.
            else
                this%snapshot_exception = .false.
            end if
        else
            if (.not. this%adaptive_timestepping) then
                call end_program(msg='constant time steps do not allow for adaptive correction to severe CFL violation!')
            end if
            dt_step = this%emergency_shrink_ratio * this%dt
            print *, "BAD CFL number ", this%CFL
            print *, "-> retry with dt=", dt_step
        end if

        if (dt_step > this%dt_max) dt_step = this%dt_max
        if (dt_step < this%dt_min) then
            write(*,*) "WARNING: dt limited to dt_min =", this%dt_min
            if (.not. this%adaptive_timestepping .and. (.not. this%snapshot_exception)) then
                call end_program(msg='minimum time step size reached, but control still detects CFL violation - constant dt not compatible!')
            end if
            dt_step = this%dt_min
        end if

        if (step_valid) this%time = this%time + dt_step
        this%dt = dt_step
    end subroutine updateTime

    subroutine update_max_CFL(this, CFL_curr)
        class(time_control), intent(inout) :: this
        double precision, intent(in) :: CFL_curr
        this%CFL = max(this%CFL, CFL_curr)
    end subroutine

    subroutine advance_snapshot_index(this)
        class(time_control), intent(inout) :: this
        this%current_snapshot_index = this%current_snapshot_index + 1
    end subroutine

    function check_timestep_validity(this) result(valid_flag)
        class(time_control) :: this
        logical :: valid_flag
        valid_flag = this%CFL_target / this%CFL >= this%timestep_shrink_limit
    end function

    subroutine reset_CFL(this)
        class(time_control), intent(inout) :: this
        this%CFL_previous_timestep = this%CFL
        this%CFL = 0.d0
    end subroutine

    function get_CFL(this) result(v)
        class(time_control) :: this
        double precision :: v
        v = this%CFL
    end function

    function get_dt(this) result(v)
        class(time_control) :: this
        double precision :: v
        v = this%dt
    end function

    function get_previous_dt(this) result(v)
        class(time_control) :: this
        double precision :: v
        v = this%dt_previous
    end function

    function get_snapshot_index(this) result(v)
        class(time_control) :: this
        integer :: v
        v = this%current_snapshot_index
    end function

    subroutine get_limits(this, dt_max, dt_min, number_of_snapshots)
        !! the private settings a copy of the controller on the device needs
        class(time_control) :: this
        double precision, intent(out) :: dt_max, dt_min
        integer, intent(out) :: number_of_snapshots
        dt_max = this%dt_max
        dt_min = this%dt_min
        number_of_snapshots = this%number_of_snapshots
    end subroutine

end module MOD_time_control
