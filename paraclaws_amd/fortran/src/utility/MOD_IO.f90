! Input and snapshot output (src/utility/MOD_IO.f90): control.in is read in
! the reference's 28-line order; custom.in (written but unread by the
! reference) carries this host's device settings; snapshots use the
! reference's qout.dat/aout.dat format (50e26.16, j outer, i inner, ghosts
! included) under output/<time>/rank_0/.
module MOD_IO
    implicit none
contains

    subroutine read_solver_control(t_ctrl, mesh, solution)
        use MOD_solver_parameters, only: set_solver_parameters
        use MOD_time_control, only: time_control
        use MOD_domain, only: mesh_fields, solution_fields
        type(time_control), intent(inout) :: t_ctrl
        type(mesh_fields), intent(inout) :: mesh
        type(solution_fields), intent(inout) :: solution
        integer :: filestat
        integer :: num_equations, num_dimensions, num_ghost, num_aux, num_waves
        double precision :: t_start, t_final, dt_initial, CFL_limit, dt_min, dt_max
        double precision :: dt_max_growth_ratio, dt_min_shrink_ratio
        integer :: number_of_snapshots, limiter_method, m_xi, m_eta
        logical :: adaptive_timestepping, higher_order_flux_correction, use_limiter
        double precision :: xi_dimensions(2), eta_dimensions(2)
        integer :: bc_xi_lower, bc_xi_upper, bc_eta_lower, bc_eta_upper

        open(unit=1, file='control.in', status='old', action='read', iostat=filestat)
        if (filestat /= 0) then
            error stop "Error opening <control.in>. Please ensure the file exists and is in the correct location."
        end if
        read(1,*) num_equations
        read(1,*) num_dimensions
        read(1,*) num_ghost
        read(1,*) num_aux
        read(1,*) num_waves
        read(1,*) t_start
        read(1,*) t_final
        read(1,*) number_of_snapshots
        read(1,*) dt_initial
        read(1,*) adaptive_timestepping
        read(1,*) CFL_limit
        read(1,*) dt_min
        read(1,*) dt_max
        read(1,*) dt_max_growth_ratio
        read(1,*) dt_min_shrink_ratio
        read(1,*) higher_order_flux_correction
        read(1,*) use_limiter
        read(1,*) limiter_method
        read(1,*) m_xi
        read(1,*) m_eta
        read(1,*) xi_dimensions(1)
        read(1,*) xi_dimensions(2)
        read(1,*) eta_dimensions(1)
        read(1,*) eta_dimensions(2)
        read(1,*) bc_xi_lower
        read(1,*) bc_xi_upper
        read(1,*) bc_eta_lower
        read(1,*) bc_eta_upper
        close(1)

        if (num_equations /= 3 .or. num_waves /= 2 .or. num_ghost /= 2) then
            error stop "the GPU engine solves 3 equations with 2 waves and 2 ghost layers"
        end if
        call set_solver_parameters(num_equations, num_waves, num_ghost, num_aux, &
            higher_order_flux_correction, use_limiter, limiter_method)
        call t_ctrl%initialize(t_start, t_final, dt_initial, number_of_snapshots, CFL_limit, &
            dt_max, dt_min, dt_max_growth_ratio, dt_min_shrink_ratio, adaptive_timestepping)
        call mesh%setup_mesh_base(m_xi, m_eta, bc_xi_lower, bc_xi_upper, bc_eta_lower, bc_eta_upper, &
            xi_dimensions, eta_dimensions)
    end subroutine read_solver_control

    subroutine read_custom_settings()
        !! custom.in: line 1 device precision (0 FP64, 1 FP32), line 2 engine
        !! (0 fused step, 1 one kernel per reference stage).  Missing file or
        !! lines keep the defaults (FP64, fused).
        use MOD_solver_parameters, only: set_device_parameters
        integer :: filestat, p, e
        p = 0
        e = 0
        open(unit=2, file='custom.in', status='old', action='read', iostat=filestat)
        if (filestat == 0) then
            read(2, *, iostat=filestat) p
            if (filestat == 0) read(2, *, iostat=filestat) e
            close(2)
        end if
        call set_device_parameters(p, e)
    end subroutine

    subroutine writeData(decomposition, time, solution)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: solution_fields
        use MOD_solver_parameters, only: num_ghost, num_equations, num_aux
        type(decomp_info), intent(in) :: decomposition
        double precision, intent(in) :: time
        type(solution_fields), intent(in) :: solution
        integer :: i, j, m, ilow, ihigh, jlow, jhigh
        character(len=40) :: timestamp_str, dirname, subdirname
1001    format(50e26.16)
        call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
        write(timestamp_str, '(F20.10)') time
        dirname = 'output/'//trim(adjustl(timestamp_str))
        subdirname = trim(dirname)//'/rank_0'
        call execute_command_line('mkdir -p '//trim(subdirname))
        open(unit=1, file=trim(subdirname)//'/qout.dat', status='unknown', form='formatted')
        do j = jlow-num_ghost, jhigh+num_ghost
            do i = ilow-num_ghost, ihigh+num_ghost
                write(1,1001) (solution%vars(m,i,j), m=1,num_equations)
            end do
        end do
        close(unit=1)
        open(unit=2, file=trim(subdirname)//'/aout.dat', status='unknown', form='formatted')
        do j = jlow-num_ghost, jhigh+num_ghost
            do i = ilow-num_ghost, ihigh+num_ghost
                write(2,1001) (solution%aux(m,i,j), m=1,num_aux)
            end do
        end do
        close(unit=2)
    end subroutine

    subroutine write_mesh_to_disk(decomposition, mesh)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: mesh_fields
        use MOD_solver_parameters, only: num_ghost
        type(decomp_info), intent(in) :: decomposition
        type(mesh_fields), intent(in) :: mesh
        integer :: i, j, ilow, ihigh, jlow, jhigh
1001    format(50e26.16)
        call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
        open(unit=1, file='output/mesh_rank_0.dat', status='unknown', form='formatted')
        do j = jlow-num_ghost, jhigh+num_ghost
            do i = ilow-num_ghost, ihigh+num_ghost
                write(1,1001) mesh%physical_space%x(i,j), mesh%physical_space%y(i,j), &
                    mesh%computational_space%xi(i,j), mesh%computational_space%eta(i,j)
            end do
        end do
        close(unit=1)
    end subroutine

end module MOD_IO
