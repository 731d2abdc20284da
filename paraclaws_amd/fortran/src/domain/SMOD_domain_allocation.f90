submodule(MOD_domain) SMOD_domain_allocation
    implicit none
contains

    module subroutine allocate_solution(this, decomposition)
        use MOD_decomposition, only: decomp_info
        use MOD_solver_parameters, only: num_equations, num_ghost, num_aux
        class(solution_fields), intent(inout) :: this
        type(decomp_info), intent(in) :: decomposition
        integer :: ilow, ihigh, jlow, jhigh
        call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
        allocate(this%vars(num_equations, ilow-num_ghost:ihigh+num_ghost, jlow-num_ghost:jhigh+num_ghost))
        allocate(this%aux(num_aux, ilow-num_ghost:ihigh+num_ghost, jlow-num_ghost:jhigh+num_ghost))
        this%vars = 0
        this%aux = 0
    end subroutine

    module subroutine deallocate_solution(this)
        class(solution_fields), intent(inout) :: this
        if (allocated(this%vars)) deallocate(this%vars)
        if (allocated(this%aux)) deallocate(this%aux)
    end subroutine

end submodule
