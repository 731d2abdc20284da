! Field containers and mesh of one rank (src/domain/MOD_domain.f90 of the
! reference): solution_fields holds vars(3, ilow-2:ihigh+2, jlow-2:jhigh+2)
! on the host; the engine keeps the device-resident copy.
module MOD_domain
    implicit none

    type mesh_base
        integer :: m_xi, m_eta
        integer :: bc_xi_low, bc_xi_high, bc_eta_low, bc_eta_high
        double precision :: xi_dimensions(2), eta_dimensions(2)
    contains
        procedure :: setup_mesh_base
    end type mesh_base

    type mesh_fields_computational_space
        double precision :: d_xi, d_eta
        double precision, allocatable :: xi(:,:), eta(:,:)
        double precision, allocatable :: xi_corner(:,:,:), eta_corner(:,:,:)
    end type

    type mesh_fields_physical_space
        double precision, allocatable :: x(:,:), y(:,:)
        double precision, allocatable :: x_corner(:,:,:), y_corner(:,:,:)
    end type

    type mesh_fields_quadrilateral_mapping
        double precision, allocatable :: capacity(:,:)
        double precision, allocatable :: length_ratio_xi(:,:,:), length_ratio_eta(:,:,:)
    end type

    type, extends(mesh_base) :: mesh_fields
        type(mesh_fields_computational_space) :: computational_space
        type(mesh_fields_physical_space) :: physical_space
        type(mesh_fields_quadrilateral_mapping) :: quadrilateral_mapping
    end type mesh_fields

    type solution_fields
        double precision, allocatable :: vars(:,:,:)
        double precision, allocatable :: aux(:,:,:)
    contains
        procedure :: allocate_solution
        procedure :: deallocate_solution
    end type solution_fields

    interface
        module subroutine allocate_solution(this, decomposition)
            use MOD_decomposition, only: decomp_info
            class(solution_fields), intent(inout) :: this
            type(decomp_info), intent(in) :: decomposition
        end subroutine

        module subroutine deallocate_solution(this)
            class(solution_fields), intent(inout) :: this
        end subroutine

        module subroutine setup_mesh(mesh, decomposition)
            !! coordinates, corners, capacity and length ratios
            !! (src/domain/SMOD_mesh_manipulation.f90:14-178)
            use MOD_decomposition, only: decomp_info
            type(mesh_fields), intent(inout) :: mesh
            type(decomp_info), intent(in) :: decomposition
        end subroutine

        module subroutine map_to_physical_coordinates(xp, yp, xi, eta)
            double precision, intent(out) :: xp, yp
            double precision, intent(in) :: xi, eta
        end subroutine

        module subroutine update_rank_boundaries(decomposition, mesh, solution)
            !! physical BC fill of the device state (one rank: no halo exchange)
            use MOD_decomposition, only: decomp_info
            type(decomp_info), intent(in) :: decomposition
            type(mesh_fields), intent(in) :: mesh
            type(solution_fields), intent(inout) :: solution
        end subroutine
    end interface

contains

    subroutine setup_mesh_base(this, m_xi, m_eta, bc_xi_lower, bc_xi_upper, bc_eta_lower, &
            bc_eta_upper, xi_dimensions, eta_dimensions)
        class(mesh_base), intent(inout) :: this
        integer, intent(in) :: m_xi, m_eta
        integer, intent(in) :: bc_xi_lower, bc_xi_upper, bc_eta_lower, bc_eta_upper
        double precision, intent(in) :: xi_dimensions(2), eta_dimensions(2)
        this%m_xi = m_xi
        this%m_eta = m_eta
        this%bc_xi_low = bc_xi_lower
        this%bc_xi_high = bc_xi_upper
        this%bc_eta_low = bc_eta_lower
        this%bc_eta_high = bc_eta_upper
        this%xi_dimensions = xi_dimensions
        this%eta_dimensions = eta_dimensions
    end subroutine

    subroutine write_base_info(mesh)
        use MOD_solver_parameters, only: num_ghost, num_aux, num_equations
        type(mesh_fields), intent(in) :: mesh
        open(unit=1, file="./output/domain.dat", status='unknown', form='formatted')
        write(1,*) 1
        write(1,*) mesh%m_xi
        write(1,*) mesh%m_eta
        write(1,*) num_equations
        write(1,*) num_aux
        write(1,*) num_ghost
        write(1,*) mesh%xi_dimensions(1)
        write(1,*) mesh%xi_dimensions(2)
        write(1,*) mesh%eta_dimensions(1)
        write(1,*) mesh%eta_dimensions(2)
        close(unit=1)
    end subroutine

end module MOD_domain
