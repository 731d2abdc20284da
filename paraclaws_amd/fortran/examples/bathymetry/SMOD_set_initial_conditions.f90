! Config 3's initial state over the whole frame: still water h + b = 1 over
! the bathymetry of bathymetry.bin (h = max(0, 1 - b), dry below 1e-6), then
! a dam-break step h = h + 1 where x < 0, at rest.  Replaces the base
! submodule of the same name.
submodule(MOD_user_subroutines) SMOD_set_initial_conditions
    implicit none
contains
    module subroutine set_initial_conditions(decomposition, mesh, solution)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: solution_fields, mesh_fields
        use MOD_solver_parameters, only: num_ghost
        use MOD_error_handling, only: abort_program
        type(decomp_info), intent(in) :: decomposition
        type(mesh_fields), intent(in) :: mesh
        type(solution_fields), intent(inout) :: solution
        integer :: i, j, ilow, ihigh, jlow, jhigh, u, stat
        double precision, allocatable :: b(:,:)
        double precision :: h
        call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
        allocate(b(ilow-num_ghost:ihigh+num_ghost, jlow-num_ghost:jhigh+num_ghost))
        open(newunit=u, file='bathymetry.bin', access='stream', form='unformatted', status='old', &
            action='read', iostat=stat)
        if (stat /= 0) call abort_program("bathymetry.bin missing: run make_bathymetry.py")
        read(u, iostat=stat) b
        close(u)
        if (stat /= 0) call abort_program("bathymetry.bin does not match the grid of control.in")
        do j = jlow-num_ghost, jhigh+num_ghost
            do i = ilow-num_ghost, ihigh+num_ghost
                h = max(0.d0, 1.d0 - b(i,j))
                if (h < 1.d-6) h = 0.d0
                if (mesh%physical_space%x(i,j) < 0.d0) h = h + 1.d0
                solution%vars(1,i,j) = h
                solution%vars(2,i,j) = 0.d0
                solution%vars(3,i,j) = 0.d0
            end do
        end do
    end subroutine
end submodule
