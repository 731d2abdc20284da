! Stretched Cartesian map x = 4 xi, y = 2 eta (the reference's
! examples/2D_dam_break/cartesian_stretching/SMOD_coordinate_mapping.f90:34-42);
! replaces the base identity map, so the host uploads the capacity (8) and
! length ratios (2 on xi faces, 4 on eta faces) to the engine.
submodule(MOD_domain) SMOD_coordinate_mapping
    implicit none
contains
    module subroutine map_to_physical_coordinates(xp, yp, xi, eta)
        double precision, intent(out) :: xp, yp
        double precision, intent(in) :: xi, eta
        xp = 4*xi
        yp = 2*eta
    end subroutine
end submodule
