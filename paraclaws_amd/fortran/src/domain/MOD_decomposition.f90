! Single-rank block decomposition: the whole m_xi x m_eta mesh is the local
! block (the reference takes its bounds from the mpidcl library; on one GPU
! there is one block and no neighbours).
module MOD_decomposition
    implicit none

    type :: decomp_info
        integer :: m_xi = 0, m_eta = 0
    contains
        procedure :: get_local_block_bounds
    end type

contains

    subroutine initialize_decomposition(d, m_xi, m_eta)
        type(decomp_info), intent(out) :: d
        integer, intent(in) :: m_xi, m_eta
        d%m_xi = m_xi
        d%m_eta = m_eta
    end subroutine

    subroutine get_local_block_bounds(this, ilow, ihigh, jlow, jhigh)
        class(decomp_info), intent(in) :: this
        integer, intent(out) :: ilow, ihigh, jlow, jhigh
        ilow = 1
        ihigh = this%m_xi
        jlow = 1
        jhigh = this%m_eta
    end subroutine

end module MOD_decomposition
