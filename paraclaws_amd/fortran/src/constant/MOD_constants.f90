! Numerical constants of the solver (src/constant/MOD_constants.f90 of the
! reference: ud_precision is the epsilon of the HLL denominator, the
! zero-speed skip and the limiter's small-wave test).
module MOD_constants
    implicit none
    integer, parameter :: axis_xi = 1
    integer, parameter :: axis_eta = 2
    double precision, parameter :: ud_precision = 1d-10
    double precision, parameter :: PI = 4.D0*DATAN(1.D0)
end module MOD_constants
