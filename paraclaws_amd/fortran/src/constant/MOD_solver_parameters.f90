! Solver parameters read from control.in (src/constant/MOD_solver_parameters.f90),
! plus the device settings this host reads from custom.in.
module MOD_solver_parameters
    implicit none

    integer, protected :: num_equations = 3
    integer, protected :: num_waves = 2
    integer, protected :: num_ghost = 2
    integer, protected :: num_aux = 10

    logical, protected :: integrate_source = .true.
    logical, protected :: add_higher_order_flux_correction = .false.
    logical, protected :: execute_post = .true.
    logical, protected :: use_limiters = .true.
    integer, protected :: limiter_method = 1

    double precision :: grav

    ! device settings (custom.in): arithmetic and engine of the GPU path
    integer, protected :: device_precision = 0   ! 0 FP64 (reference order), 1 FP32
    integer, protected :: device_engine = 0      ! 0 fused step, 1 one kernel per stage
    integer, protected :: device_loop = 0        ! 0 time loop on the host, 1 on the device between snapshots

contains

    subroutine set_solver_parameters(num_equations_, num_waves_, num_ghost_, num_aux_, &
            add_higher_order_flux_correction_, use_limiters_, limiter_method_)
        integer :: num_equations_, num_waves_, num_ghost_, num_aux_
        logical :: add_higher_order_flux_correction_, use_limiters_
        integer :: limiter_method_
        num_equations = num_equations_
        num_waves = num_waves_
        num_ghost = num_ghost_
        num_aux = num_aux_
        add_higher_order_flux_correction = add_higher_order_flux_correction_
        use_limiters = use_limiters_
        limiter_method = limiter_method_
    end subroutine

    subroutine set_device_parameters(precision_, engine_, loop_)
        integer, intent(in) :: precision_, engine_, loop_
        device_precision = precision_
        device_engine = engine_
        device_loop = loop_
    end subroutine

end module MOD_solver_parameters
