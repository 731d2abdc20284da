! Config 3's physics on the GPU path: the bathymetry slope source with the
! wet/dry cut-off (h <= 1e-6 is dry), no friction.  The bottom elevation is
! read from bathymetry.bin (make_bathymetry.py writes the seeded sum of eight
! Gaussians for the grid of control.in): b(i, j), i = -1 .. nx+2 fastest, as
! doubles.  Replaces the base submodule of the same name.
submodule(MOD_user_subroutines) SMOD_device_physics
    implicit none
contains
    module subroutine device_physics(source_kind, manning_n, dry_tol)
        integer, intent(out) :: source_kind
        double precision, intent(out) :: manning_n, dry_tol
        source_kind = 1
        manning_n = 0.d0
        dry_tol = 1.d-6
    end subroutine

    module subroutine device_bathymetry(decomposition, mesh, b)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: mesh_fields
        use MOD_error_handling, only: abort_program
        type(decomp_info), intent(in) :: decomposition
        type(mesh_fields), intent(in) :: mesh
        double precision, intent(inout) :: b(-1:, -1:)
        integer :: u, stat
        open(newunit=u, file='bathymetry.bin', access='stream', form='unformatted', status='old', &
            action='read', iostat=stat)
        if (stat /= 0) call abort_program("bathymetry.bin missing: run make_bathymetry.py")
        read(u, iostat=stat) b
        close(u)
        if (stat /= 0) call abort_program("bathymetry.bin does not match the grid of control.in")
    end subroutine

    module function host_state_hooks() result(needs)
        logical :: needs
        needs = .false.
    end function
end submodule
