! Settings of the GPU path that the reference's per-cell hooks cannot carry
! (see MOD_user_subroutines).  This base file selects the reference's physics:
! no device source (phi = 0, src/user_subroutines/SMOD_source_function.f90:15-28),
! no wet/dry cut-off, no bathymetry, and an after_step that does not touch the
! state.  A USER_DIR override of this submodule selects other settings.
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
        needs = .false.
    end function
end submodule
