! work_arrays of the reference (src/solver/MOD_work_arrays.f90).  The wave,
! fluctuation, limiter and correction arrays live on the device inside the
! engine (registers of the fused step, or HBM planes of the staged engine);
! this type keeps the type-bound interface and owns the engine's lifetime.
module MOD_work_arrays
    implicit none

    type work_arrays
        logical :: ready = .false.
    contains
        procedure :: allocate_work_arrays
        procedure :: deallocate_work_arrays
        procedure :: flash_work_to_zero
        procedure :: initialize_work
    end type

contains

    subroutine allocate_work_arrays(work, decomposition, mesh)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: mesh_fields
        use MOD_device, only: device_create, device_set_metrics, device_set_bathymetry
        use MOD_user_subroutines, only: device_physics, device_bathymetry
        use MOD_solver_parameters
        class(work_arrays), intent(inout) :: work
        type(decomp_info), intent(in) :: decomposition
        type(mesh_fields), intent(in) :: mesh
        integer :: source_kind, ilow, ihigh, jlow, jhigh
        double precision :: manning_n, dry_tol
        double precision, allocatable :: b(:,:)
        ! the device source that replaces source_function (SMOD_device_physics)
        call device_physics(source_kind, manning_n, dry_tol)
        call device_create(mesh%m_xi, mesh%m_eta, &
            [mesh%bc_xi_low, mesh%bc_xi_high, mesh%bc_eta_low, mesh%bc_eta_high], &
            mesh%xi_dimensions, mesh%eta_dimensions, grav, device_precision, device_engine, &
            add_higher_order_flux_correction, use_limiters, limiter_method, &
            source_kind, manning_n, dry_tol)
        if (source_kind /= 0 .or. dry_tol > 0.d0) write(*,'(a,i0,a,es12.5,a,es12.5)') &
            " device source kind ", source_kind, "  Manning n ", manning_n, "  wet/dry threshold ", dry_tol
        if (source_kind == 1 .or. source_kind == 3) then
            call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
            allocate(b(ilow-2:ihigh+2, jlow-2:jhigh+2))
            call device_bathymetry(decomposition, mesh, b)
            call device_set_bathymetry(b)
            deallocate(b)
        end if
        ! a user coordinate map other than the identity: hand the engine the
        ! capacity and length ratios computed from it (SMOD_mesh_manipulation)
        if (.not. identity_map(mesh)) then
            call device_set_metrics(mesh%quadrilateral_mapping%capacity, &
                mesh%quadrilateral_mapping%length_ratio_xi(1,:,:), &
                mesh%quadrilateral_mapping%length_ratio_eta(1,:,:))
            write(*,*) "mapped grid: capacity and length ratios uploaded to the engine"
        end if
        work%ready = .true.
    end subroutine

    logical function identity_map(mesh)
        use MOD_domain, only: mesh_fields
        type(mesh_fields), intent(in) :: mesh
        identity_map = all(mesh%physical_space%x_corner == mesh%computational_space%xi_corner) .and. &
            all(mesh%physical_space%y_corner == mesh%computational_space%eta_corner)
    end function

    subroutine deallocate_work_arrays(work)
        use MOD_device, only: device_destroy
        class(work_arrays), intent(inout) :: work
        call device_destroy()
        work%ready = .false.
    end subroutine

    subroutine flash_work_to_zero(work)
        !! the engine keeps no accumulating work arrays between sweeps
        class(work_arrays), intent(inout) :: work
    end subroutine

    subroutine initialize_work(work)
        class(work_arrays), intent(inout) :: work
        call work%flash_work_to_zero()
    end subroutine

end module MOD_work_arrays
