! Error convention of the reference (src/utility/MOD_error_handling.f90):
! end_program for a graceful premature stop, abort_program for fatal errors.
! This host runs a single rank, so there is no MPI communicator to abort.
module MOD_error_handling
    implicit none
contains

    subroutine end_program(msg)
        character(len=*), intent(in) :: msg
        print *, "***************************************************"
        print *, "PROGRAM FINISHED PREMATURELY"
        print *, "REASON: "
        print *, msg
        print *, "***************************************************"
        stop
    end subroutine

    subroutine abort_program(msg)
        character(len=*), intent(in) :: msg
        write(*,*) 'Rank ', 0, ' encountered error: ', trim(msg)
        error stop 1
    end subroutine abort_program

end module MOD_error_handling
