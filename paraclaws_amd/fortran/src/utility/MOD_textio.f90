! Text output of the reference's snapshot formats, formatted in parallel.
!
! write_e26_records writes the records a(:, i, j) (i inner, j outer) as the
! lines the reference's write(u, '(50e26.16)') a(:, i, j) makes -- one line
! of E26.16 fields per record (src/utility/MOD_IO.f90:397-476) -- with the
! same edit descriptors: blocks of rows are formatted into a line buffer by
! internal writes on OpenMP threads, and each block goes to the file as one
! stream write (the line terminator of a formatted sequential record is a
! newline).  The bytes are those of the serial formatted write.
module MOD_textio
    implicit none
    private
    public :: write_e26_records

    integer, parameter :: rows_per_block = 64

contains

    subroutine write_e26_records(path, a)
        character(len=*), intent(in) :: path
        double precision, intent(in) :: a(:, :, :)
        character(len=:), allocatable :: buf(:, :)
        integer :: m, ni, nj, w, i, j, j0, j1, u
        m = size(a, 1)
        ni = size(a, 2)
        nj = size(a, 3)
        if (m > 50) then  ! a record of the reference's format would span several lines
            open(newunit=u, file=path, status='replace', form='formatted')
            do j = 1, nj
                do i = 1, ni
                    write(u, '(50e26.16)') a(:, i, j)
                end do
            end do
            close(u)
            return
        end if
        w = 26 * m + 1
        allocate(character(len=w) :: buf(ni, rows_per_block))
        open(newunit=u, file=path, access='stream', form='unformatted', status='replace')
        do j0 = 1, nj, rows_per_block
            j1 = min(nj, j0 + rows_per_block - 1)
            !$omp parallel do private(i)
            do j = j0, j1
                do i = 1, ni
                    write(buf(i, j - j0 + 1)(1:w - 1), '(50e26.16)') a(:, i, j)
                    buf(i, j - j0 + 1)(w:w) = new_line('a')
                end do
            end do
            !$omp end parallel do
            write(u) buf(:, 1:j1 - j0 + 1)
        end do
        close(u)
    end subroutine

end module MOD_textio
