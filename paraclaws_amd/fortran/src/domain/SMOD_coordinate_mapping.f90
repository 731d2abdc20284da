! Identity map from computational to physical coordinates
! (src/user_subroutines/SMOD_coordinate_mapping.f90:35-43).  The engine's
! metrics assume this map; a USER_DIR override of it is reported at start-up.
submodule(MOD_domain) SMOD_coordinate_mapping
    implicit none
contains
    module subroutine map_to_physical_coordinates(xp, yp, xi, eta)
        double precision, intent(out) :: xp, yp
        double precision, intent(in) :: xi, eta
        xp = xi
        yp = eta
    end subroutine
end submodule
