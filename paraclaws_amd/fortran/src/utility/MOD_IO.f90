! Input and snapshot output (src/utility/MOD_IO.f90): control.in is read in
! the reference's 28-line order; custom.in (written but unread by the
! reference) carries this host's device settings; snapshots use the
! reference's qout.dat/aout.dat format (50e26.16, j outer, i inner, ghosts
! included) under output/<time>/rank_0/.  Built with USE_VTK (the Makefile's
! use_vtk=1, the default as in the reference), each snapshot also goes to the
! reference's VTK layout (MOD_IO.f90:111-394 there): a StructuredGrid piece
! output/<time>/rank_0/solution_rank_0.vts, the parallel header
! output/<time>/solution.pvts and the collection output/solution.pvd.  The
! XML is written here directly (the reference links the VTKFortran library).
module MOD_IO
    use iso_fortran_env, only: int8, int32
    implicit none
    double precision, allocatable, private :: pvd_times(:)
    character(len=64), parameter, private :: b64_table = &
        'ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/'
contains

    subroutine read_solver_control(t_ctrl, mesh, solution)
        use MOD_solver_parameters, only: set_solver_parameters
        use MOD_time_control, only: time_control
        use MOD_domain, only: mesh_fields, solution_fields
        type(time_control), intent(inout) :: t_ctrl
        type(mesh_fields), intent(inout) :: mesh
        type(solution_fields), intent(inout) :: solution
        integer :: filestat
        integer :: num_equations, num_dimensions, num_ghost, num_aux, num_waves
        double precision :: t_start, t_final, dt_initial, CFL_limit, dt_min, dt_max
        double precision :: dt_max_growth_ratio, dt_min_shrink_ratio
        integer :: number_of_snapshots, limiter_method, m_xi, m_eta
        logical :: adaptive_timestepping, higher_order_flux_correction, use_limiter
        double precision :: xi_dimensions(2), eta_dimensions(2)
        integer :: bc_xi_lower, bc_xi_upper, bc_eta_lower, bc_eta_upper

        open(unit=1, file='control.in', status='old', action='read', iostat=filestat)
        if (filestat /= 0) then
            error stop "Error opening <control.in>. Please ensure the file exists and is in the correct location."
        end if
        read(1,*) num_equations
        read(1,*) num_dimensions
        read(1,*) num_ghost
        read(1,*) num_aux
        read(1,*) num_waves
        read(1,*) t_start
        read(1,*) t_final
        read(1,*) number_of_snapshots
        read(1,*) dt_initial
        read(1,*) adaptive_timestepping
        read(1,*) CFL_limit
        read(1,*) dt_min
        read(1,*) dt_max
        read(1,*) dt_max_growth_ratio
        read(1,*) dt_min_shrink_ratio
        read(1,*) higher_order_flux_correction
        read(1,*) use_limiter
        read(1,*) limiter_method
        read(1,*) m_xi
        read(1,*) m_eta
        read(1,*) xi_dimensions(1)
        read(1,*) xi_dimensions(2)
        read(1,*) eta_dimensions(1)
        read(1,*) eta_dimensions(2)
        read(1,*) bc_xi_lower
        read(1,*) bc_xi_upper
        read(1,*) bc_eta_lower
        read(1,*) bc_eta_upper
        close(1)

        if (num_equations /= 3 .or. num_waves /= 2 .or. num_ghost /= 2) then
            error stop "the GPU engine solves 3 equations with 2 waves and 2 ghost layers"
        end if
        call set_solver_parameters(num_equations, num_waves, num_ghost, num_aux, &
            higher_order_flux_correction, use_limiter, limiter_method)
        call t_ctrl%initialize(t_start, t_final, dt_initial, number_of_snapshots, CFL_limit, &
            dt_max, dt_min, dt_max_growth_ratio, dt_min_shrink_ratio, adaptive_timestepping)
        call mesh%setup_mesh_base(m_xi, m_eta, bc_xi_lower, bc_xi_upper, bc_eta_lower, bc_eta_upper, &
            xi_dimensions, eta_dimensions)
    end subroutine read_solver_control

    subroutine read_custom_settings()
        !! custom.in: line 1 device precision (0 FP64, 1 FP32), line 2 engine
        !! (0 fused step, 1 one kernel per reference stage), line 3 time loop
        !! (0 on the host, step by step as the reference's; 1 on the device,
        !! returning to the host at every snapshot).  Missing file or lines
        !! keep the defaults (FP64, fused, host loop).
        use MOD_solver_parameters, only: set_device_parameters
        integer :: filestat, p, e, l
        p = 0
        e = 0
        l = 0
        open(unit=2, file='custom.in', status='old', action='read', iostat=filestat)
        if (filestat == 0) then
            read(2, *, iostat=filestat) p
            if (filestat == 0) read(2, *, iostat=filestat) e
            if (filestat == 0) read(2, *, iostat=filestat) l
            close(2)
        end if
        call set_device_parameters(p, e, l)
    end subroutine

    subroutine writeData(decomposition, time, solution)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: solution_fields
        use MOD_solver_parameters, only: num_ghost, num_equations, num_aux
        use MOD_textio, only: write_e26_records
        type(decomp_info), intent(in) :: decomposition
        double precision, intent(in) :: time
        type(solution_fields), intent(in) :: solution
        integer :: ilow, ihigh, jlow, jhigh, i0, i1, j0, j1
        character(len=40) :: timestamp_str, dirname, subdirname
        call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
        i0 = ilow - num_ghost
        i1 = ihigh + num_ghost
        j0 = jlow - num_ghost
        j1 = jhigh + num_ghost
        write(timestamp_str, '(F20.10)') time
        dirname = 'output/'//trim(adjustl(timestamp_str))
        subdirname = trim(dirname)//'/rank_0'
        call execute_command_line('mkdir -p '//trim(subdirname))
        ! qout.dat and aout.dat: one line of (50e26.16) fields per frame cell, j
        ! outer, i inner, ghosts included
        call write_e26_records(trim(subdirname)//'/qout.dat', solution%vars(1:num_equations, i0:i1, j0:j1))
        call write_e26_records(trim(subdirname)//'/aout.dat', solution%aux(1:num_aux, i0:i1, j0:j1))
    end subroutine

    subroutine write_mesh_to_disk(decomposition, mesh)
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: mesh_fields
        use MOD_solver_parameters, only: num_ghost
        type(decomp_info), intent(in) :: decomposition
        type(mesh_fields), intent(in) :: mesh
        integer :: i, j, ilow, ihigh, jlow, jhigh
1001    format(50e26.16)
        call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
        open(unit=1, file='output/mesh_rank_0.dat', status='unknown', form='formatted')
        do j = jlow-num_ghost, jhigh+num_ghost
            do i = ilow-num_ghost, ihigh+num_ghost
                write(1,1001) mesh%physical_space%x(i,j), mesh%physical_space%y(i,j), &
                    mesh%computational_space%xi(i,j), mesh%computational_space%eta(i,j)
            end do
        end do
        close(unit=1)
    end subroutine

    subroutine write_vtk_rank(decomposition, mesh, solution, time)
        !! StructuredGrid piece of this rank's interior cells (reference
        !! write_vtk_rank): points are the cell centres (x, y, 0); point data
        !! q1..q(num_equations) in Float64; VTK "binary" DataArrays (base64 of
        !! a UInt32 byte count, then base64 of the little-endian values).
        use MOD_decomposition, only: decomp_info
        use MOD_domain, only: mesh_fields, solution_fields
        use MOD_solver_parameters, only: num_equations
        type(decomp_info), intent(in) :: decomposition
        type(mesh_fields), intent(in) :: mesh
        type(solution_fields), intent(in) :: solution
        double precision, intent(in) :: time
#ifdef USE_VTK
        integer :: ilow, ihigh, jlow, jhigh, i, j, k, m, npts, u
        double precision, allocatable :: buf(:)
        character(len=40) :: timestamp_str, qname
        character(len=200) :: dirname, filename
        character(len=96) :: ext
        character(len=1), parameter :: nl = new_line('a')

        call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
        npts = (ihigh - ilow + 1) * (jhigh - jlow + 1)
        write(ext, '(I0,1X,I0,1X,I0,1X,I0,A)') ilow, ihigh, jlow, jhigh, ' 1 1'
        write(timestamp_str, '(F20.10)') time
        dirname = 'output/'//trim(adjustl(timestamp_str))//'/rank_0'
        call execute_command_line('mkdir -p '//trim(dirname))
        filename = trim(dirname)//'/solution_rank_0.vts'
        open(newunit=u, file=trim(filename), status='replace', access='stream', form='unformatted')
        write(u) '<?xml version="1.0"?>'//nl
        write(u) '<VTKFile type="StructuredGrid" version="1.0" byte_order="LittleEndian" ' // &
            'header_type="UInt32">'//nl
        write(u) '  <StructuredGrid WholeExtent="'//trim(ext)//'">'//nl
        write(u) '    <Piece Extent="'//trim(ext)//'">'//nl
        write(u) '      <Points>'//nl
        write(u) '        <DataArray type="Float64" NumberOfComponents="3" Name="Points" ' // &
            'format="binary">'//nl
        allocate(buf(3 * npts))
        k = 0
        do j = jlow, jhigh
            do i = ilow, ihigh
                buf(k + 1) = mesh%physical_space%x(i, j)
                buf(k + 2) = mesh%physical_space%y(i, j)
                buf(k + 3) = 0.0d0
                k = k + 3
            end do
        end do
        call write_b64_array(u, buf)
        deallocate(buf)
        write(u) nl//'        </DataArray>'//nl
        write(u) '      </Points>'//nl
        write(u) '      <PointData>'//nl
        allocate(buf(npts))
        do m = 1, num_equations
            k = 0
            do j = jlow, jhigh
                do i = ilow, ihigh
                    k = k + 1
                    buf(k) = solution%vars(m, i, j)
                end do
            end do
            write(qname, '(A,I0)') 'q', m
            write(u) '        <DataArray type="Float64" NumberOfComponents="1" Name="' // &
                trim(qname)//'" format="binary">'//nl
            call write_b64_array(u, buf)
            write(u) nl//'        </DataArray>'//nl
        end do
        deallocate(buf)
        write(u) '      </PointData>'//nl
        write(u) '    </Piece>'//nl
        write(u) '  </StructuredGrid>'//nl
        write(u) '</VTKFile>'//nl
        close(u)
#endif
    end subroutine

    subroutine write_pvts_rank(decomposition, time)
        !! Parallel header of the snapshot (one piece per rank; one rank here)
        !! and the time-series collection output/solution.pvd (reference
        !! write_pvts_rank / write_pvd_collection).
        use MOD_decomposition, only: decomp_info
        use MOD_solver_parameters, only: num_equations
        type(decomp_info), intent(in) :: decomposition
        double precision, intent(in) :: time
#ifdef USE_VTK
        integer :: ilow, ihigh, jlow, jhigh, m, u, count, idx
        double precision, allocatable :: times(:)
        character(len=40) :: timestamp_str, qname
        character(len=96) :: ext

        call decomposition%get_local_block_bounds(ilow, ihigh, jlow, jhigh)
        write(ext, '(I0,1X,I0,1X,I0,1X,I0,A)') ilow, ihigh, jlow, jhigh, ' 1 1'
        write(timestamp_str, '(F20.10)') time
        open(newunit=u, file='output/'//trim(adjustl(timestamp_str))//'/solution.pvts', &
            status='replace', form='formatted')
        write(u, '(A)') '<?xml version="1.0"?>'
        write(u, '(A)') '<VTKFile type="PStructuredGrid" version="1.0" byte_order="LittleEndian">'
        write(u, '(A)') '  <PStructuredGrid WholeExtent="'//trim(ext)//'" GhostLevel="0">'
        write(u, '(A)') '    <PPointData>'
        do m = 1, num_equations
            write(qname, '(A,I0)') 'q', m
            write(u, '(A)') '      <PDataArray type="Float64" Name="'//trim(qname)// &
                '" NumberOfComponents="1"/>'
        end do
        write(u, '(A)') '    </PPointData>'
        write(u, '(A)') '    <PPoints>'
        write(u, '(A)') '      <PDataArray type="Float64" NumberOfComponents="3" Name="Points"/>'
        write(u, '(A)') '    </PPoints>'
        write(u, '(A)') '    <Piece Extent="'//trim(ext)//'" Source="rank_0/solution_rank_0.vts"/>'
        write(u, '(A)') '  </PStructuredGrid>'
        write(u, '(A)') '</VTKFile>'
        close(u)

        ! collection: one DataSet per snapshot time, rewritten each time
        count = 0
        if (allocated(pvd_times)) count = size(pvd_times)
        if (count > 0) then
            if (pvd_times(count) == time) return
        end if
        allocate(times(count + 1))
        if (count > 0) times(1:count) = pvd_times
        times(count + 1) = time
        call move_alloc(times, pvd_times)
        open(newunit=u, file='output/solution.pvd', status='replace', form='formatted')
        write(u, '(A)') '<?xml version="1.0"?>'
        write(u, '(A)') '<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">'
        write(u, '(A)') '  <Collection>'
        do idx = 1, count + 1
            write(timestamp_str, '(F20.10)') pvd_times(idx)
            write(u, '(A)') '    <DataSet timestep="'//trim(adjustl(timestamp_str))// &
                '" group="" part="0" file="'//trim(adjustl(timestamp_str))//'/solution.pvts"/>'
        end do
        write(u, '(A)') '  </Collection>'
        write(u, '(A)') '</VTKFile>'
        close(u)
#endif
    end subroutine

    subroutine write_b64_array(u, a)
        !! One VTK inline binary DataArray body: base64 of the UInt32 byte
        !! count, then base64 of the values' bytes (encoded in chunks of a
        !! multiple of three bytes, so the chunks concatenate to one encoding).
        integer, intent(in) :: u
        double precision, intent(in) :: a(:)
        integer(int8) :: hdr(4)
        integer(int8), allocatable :: bytes(:)
        integer :: n, first, last
        integer, parameter :: chunk = 3 * 131072
        n = 8 * size(a)
        hdr = transfer(int(n, int32), hdr)
        call write_b64(u, hdr)
        allocate(bytes(n))
        bytes = transfer(a, bytes)
        do first = 1, n, chunk
            last = min(n, first + chunk - 1)
            call write_b64(u, bytes(first:last))
        end do
        deallocate(bytes)
    end subroutine

    subroutine write_b64(u, bytes)
        !! Base64 (RFC 4648, with '=' padding) of a byte string.
        integer, intent(in) :: u
        integer(int8), intent(in) :: bytes(:)
        character(len=:), allocatable :: out
        integer :: n, i, k, b0, b1, b2
        n = size(bytes)
        allocate(character(len=4 * ((n + 2) / 3)) :: out)
        k = 0
        do i = 1, n, 3
            b0 = iand(int(bytes(i)), 255)
            b1 = 0
            b2 = 0
            if (i + 1 <= n) b1 = iand(int(bytes(i + 1)), 255)
            if (i + 2 <= n) b2 = iand(int(bytes(i + 2)), 255)
            out(k+1:k+1) = b64_table(ishft(b0, -2) + 1:ishft(b0, -2) + 1)
            out(k+2:k+2) = b64_table(ior(ishft(iand(b0, 3), 4), ishft(b1, -4)) + 1: &
                                     ior(ishft(iand(b0, 3), 4), ishft(b1, -4)) + 1)
            if (i + 1 <= n) then
                out(k+3:k+3) = b64_table(ior(ishft(iand(b1, 15), 2), ishft(b2, -6)) + 1: &
                                         ior(ishft(iand(b1, 15), 2), ishft(b2, -6)) + 1)
            else
                out(k+3:k+3) = '='
            end if
            if (i + 2 <= n) then
                out(k+4:k+4) = b64_table(iand(b2, 63) + 1:iand(b2, 63) + 1)
            else
                out(k+4:k+4) = '='
            end if
            k = k + 4
        end do
        write(u) out
    end subroutine

end module MOD_IO
