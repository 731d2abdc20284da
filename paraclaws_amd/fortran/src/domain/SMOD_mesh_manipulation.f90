! Mesh coordinates and metrics as the reference computes them
! (src/domain/SMOD_mesh_manipulation.f90:14-178): cell centres, corners, the
! shoelace capacity and the face length ratios, for the user's coordinate map.
submodule(MOD_domain) SMOD_mesh_manipulation
    implicit none
contains

    module subroutine setup_mesh(mesh, decomposition)
        use MOD_decomposition, only: decomp_info
        use MOD_solver_parameters, only: num_ghost
        type(mesh_fields), intent(inout) :: mesh
        type(decomp_info), intent(in) :: decomposition
        integer :: ilow, ihigh, jlow, jhigh, i, j, m
        double precision :: fdx, fdy, vnorm, area

        call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
        associate(c => mesh%computational_space, p => mesh%physical_space, &
                  q => mesh%quadrilateral_mapping)
        allocate(c%xi(ilow-num_ghost:ihigh+num_ghost, jlow-num_ghost:jhigh+num_ghost))
        allocate(c%eta, p%x, p%y, q%capacity, mold=c%xi)
        allocate(c%xi_corner(5, ilow-num_ghost:ihigh+num_ghost, jlow-num_ghost:jhigh+num_ghost))
        allocate(c%eta_corner, p%x_corner, p%y_corner, mold=c%xi_corner)
        allocate(q%length_ratio_xi(2, ilow-num_ghost:ihigh+num_ghost, jlow-num_ghost:jhigh+num_ghost))
        allocate(q%length_ratio_eta, mold=q%length_ratio_xi)

        c%d_xi = (mesh%xi_dimensions(2) - mesh%xi_dimensions(1)) / mesh%m_xi
        c%d_eta = (mesh%eta_dimensions(2) - mesh%eta_dimensions(1)) / mesh%m_eta
        do j = jlow-num_ghost, jhigh+num_ghost
            do i = ilow-num_ghost, ihigh+num_ghost
                c%xi_corner(1,i,j) = mesh%xi_dimensions(1) + (i-1)*c%d_xi
                c%xi_corner(2,i,j) = c%xi_corner(1,i,j) + c%d_xi
                c%xi_corner(3,i,j) = c%xi_corner(2,i,j)
                c%xi_corner(4,i,j) = c%xi_corner(1,i,j)
                c%xi_corner(5,i,j) = c%xi_corner(1,i,j)
                c%eta_corner(1,i,j) = mesh%eta_dimensions(1) + (j-1)*c%d_eta
                c%eta_corner(2,i,j) = c%eta_corner(1,i,j)
                c%eta_corner(3,i,j) = c%eta_corner(2,i,j) + c%d_eta
                c%eta_corner(4,i,j) = c%eta_corner(3,i,j)
                c%eta_corner(5,i,j) = c%eta_corner(1,i,j)
                c%xi(i,j) = mesh%xi_dimensions(1) + (i-1)*c%d_xi + c%d_xi/2
                c%eta(i,j) = mesh%eta_dimensions(1) + (j-1)*c%d_eta + c%d_eta/2
            end do
        end do
        do j = jlow-num_ghost, jhigh+num_ghost
            do i = ilow-num_ghost, ihigh+num_ghost
                call map_to_physical_coordinates(p%x(i,j), p%y(i,j), c%xi(i,j), c%eta(i,j))
                do m = 1, 5
                    call map_to_physical_coordinates(p%x_corner(m,i,j), p%y_corner(m,i,j), &
                        c%xi_corner(m,i,j), c%eta_corner(m,i,j))
                end do
                fdx = p%x_corner(4,i,j) - p%x_corner(1,i,j)
                fdy = p%y_corner(4,i,j) - p%y_corner(1,i,j)
                vnorm = norm2([fdx, fdy])
                q%length_ratio_xi(:,i,j) = vnorm / c%d_eta
                fdx = p%x_corner(2,i,j) - p%x_corner(1,i,j)
                fdy = p%y_corner(2,i,j) - p%y_corner(1,i,j)
                vnorm = norm2([fdx, fdy])
                q%length_ratio_eta(:,i,j) = vnorm / c%d_xi
                area = 0.d0
                do m = 1, 4
                    area = area + 0.5d0*(p%y_corner(m,i,j) + p%y_corner(m+1,i,j)) &
                        * (p%x_corner(m,i,j) - p%x_corner(m+1,i,j))
                end do
                q%capacity(i,j) = area / (c%d_xi*c%d_eta)
            end do
        end do
        end associate
    end subroutine

end submodule
