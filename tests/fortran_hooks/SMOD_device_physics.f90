! Test case of the host's hooks: the reference's physics, and an after_step
! that changes the state, so the host fetches the state before it and sends
! it back after it.
submodule(MOD_user_subroutines) SMOD_device_physics
    implicit none
contains
    module subroutine device_physics(source_kind, manning_n, dry_tol)
        integer, intent(out) :: source_kind
        double precision, intent(out) :: manning_n, dry_tol
        source_kind = 0
        manning_n = 0.d0
        dry_tol = 0.d0
    end subroutine

    module subroutine device_bathymetry(decomposition, mesh, b)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: mesh_fields
        type(decomp_info), intent(in) :: decomposition
        type(mesh_fields), intent(in) :: mesh
        double precision, intent(inout) :: b(-1:, -1:)
        b = 0.d0
    end subroutine

    module function host_state_hooks() result(needs)
        logical :: needs
        needs = .true.
    end function
end submodule
