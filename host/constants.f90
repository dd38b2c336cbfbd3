!=======================================================================
! Named constants of the problem interface.  The names and values are the
! ones problem files already use (src/constants.f90:31-98 of the
! reference: solver, boundary, limiter, EOS, cooling and conduction
! selectors), so an unchanged user parameters.f90 compiles against it.
!=======================================================================
module constants
  implicit none
  real, parameter :: pi = acos(-1.)
  real, parameter :: deg = pi / 180.
  real, parameter :: amh = 1.66e-24
  real, parameter :: Kb = 1.38e-16
  real, parameter :: Rg = 8.3145e7
  real, parameter :: Ggrav = 6.67259e-8
  real, parameter :: clight = 2.99e10
  real, parameter :: Msun = 1.99e33
  real, parameter :: Rsun = 6.955e10
  real, parameter :: gsun = 274.e2
  real, parameter :: Mjup = 1.898e30
  real, parameter :: Rjup = 7.1492e9
  real, parameter :: AU = 1.496e13
  real, parameter :: pc = 3.0857e18
  real, parameter :: kpc = 3.0857e21
  real, parameter :: hr = 3600.
  real, parameter :: day = 86400.
  real, parameter :: yr = 3.1536e7
  real, parameter :: Myr = 3.1536e13

  ! Riemann solvers
  integer, parameter :: SOLVER_HLL = 1, SOLVER_HLLC = 2, SOLVER_HLLE = 3, SOLVER_HLLD = 4
  integer, parameter :: SOLVER_HLLE_SPLIT_B = 5, SOLVER_HLLD_SPLIT_B = 6
  integer, parameter :: SOLVER_HLLE_SPLIT_ALL = 7, SOLVER_HLLD_SPLIT_ALL = 8
  ! equations of state
  integer, parameter :: EOS_ADIABATIC = 1, EOS_SINGLE_SPECIE = 2, EOS_H_RATE = 3, EOS_CHEM = 4
  ! cooling
  integer, parameter :: COOL_NONE = 0, COOL_H = 1, COOL_BBC = 2, COOL_DMC = 3, COOL_CHI = 4, COOL_CHEM = 5
  ! boundaries
  integer, parameter :: BC_OUTFLOW = 1, BC_CLOSED = 2, BC_PERIODIC = 3, BC_OTHER = 4
  ! slope limiters
  integer, parameter :: LIMITER_NO_AVERAGE = -1, LIMITER_NO_LIMIT = 0, LIMITER_MINMOD = 1
  integer, parameter :: LIMITER_VAN_LEER = 2, LIMITER_VAN_ALBADA = 3, LIMITER_UMIST = 4
  integer, parameter :: LIMITER_WOODWARD = 5, LIMITER_SUPERBEE = 6
  ! thermal conduction
  integer, parameter :: TC_OFF = 0, TC_ISOTROPIC = 1, TC_ANISOTROPIC = 2
end module constants
