!=======================================================================
! Test-problem configuration for building the reference's own physics
! modules (src/hydro_core.f90, src/hll*.f90, src/boundaries.f90) into the
! reference build oracle/_ref.  Every Guacho problem supplies a module named
! `parameters` (OT/parameters.f90 and EXO/parameters.f90 are the two the
! reference ships); this one is the oracle's problem, with the
! compile-time choices taken from REF_* macros so one source builds each
! configuration variant listed in oracle/Makefile.
!
! Serial build (no MPIP): the reference's hydro_core.f90:625 imports
! `mpi_real_kind` from this module unconditionally, so it is declared here
! as an unused integer; no MPI routine is called on the serial path.
!
! REF_FLUX_CD (default 1) and REF_ETA (default 0.) select enable_flux_cd
! and the artificial viscosity; the step builds (REF_FLUX_CD = 0) link the
! reference's hydro_solver.f90 and call its tstep.
!=======================================================================
#ifndef REF_FLUX_CD
#define REF_FLUX_CD 1
#endif
#ifndef REF_ETA
#define REF_ETA 0.0
#endif
module parameters
  use constants
  implicit none

  logical, parameter :: pmhd = (REF_PMHD /= 0)
  logical, parameter :: mhd  = (REF_MHD /= 0)
  integer, parameter :: riemann_solver = REF_SOLVER
  logical, parameter :: eight_wave = (REF_EIGHT_WAVE /= 0)
  logical, parameter :: enable_flux_cd = (REF_FLUX_CD /= 0)
  logical, parameter :: dump_divb = .false.
  logical, parameter :: out_bin = .false.
  logical, parameter :: out_vtk = .false.
  integer, parameter :: eq_of_state = REF_EOS
  integer, parameter :: cooling = COOL_NONE

  integer, parameter :: bc_left   = REF_BC_LEFT
  integer, parameter :: bc_right  = REF_BC_RIGHT
  integer, parameter :: bc_bottom = REF_BC_BOTTOM
  integer, parameter :: bc_top    = REF_BC_TOP
  integer, parameter :: bc_out    = REF_BC_OUT
  integer, parameter :: bc_in     = REF_BC_IN
  logical, parameter :: bc_user   = (REF_USER_BC /= 0)

  integer, parameter :: slope_limiter = REF_LIMITER
  integer, parameter :: th_cond = REF_TH_COND
  logical, parameter :: tc_saturation = (REF_TC_SAT /= 0)
  logical, parameter :: dif_rad = .false.
  logical, parameter :: user_source_terms = (REF_USER_SRC /= 0)
  logical, parameter :: radiation_pressure = .false.

  integer, parameter :: npas = REF_NPAS
  integer, parameter :: nxtot = REF_NX, nytot = REF_NY, nztot = REF_NZ

  real, parameter :: xmax = 1.0, ymax = 1.0, zmax = 1.0
  real, parameter :: xphys = 1.0
  real, parameter :: cv = REF_CV
  real, parameter :: gamma = (cv + 1.) / cv
  real, parameter :: mu = 1.
  real, parameter :: rsc = xphys / xmax
  real, parameter :: rhosc = 1.
  real, parameter :: Tempsc = REF_TEMPSC
  real, parameter :: vsc2 = 1.
  real, parameter :: vsc = sqrt(vsc2)
  real, parameter :: Psc = rhosc * vsc2
  real, parameter :: tsc = rsc / sqrt(vsc2)
  real, parameter :: bsc = sqrt(4.0 * pi * Psc)
  real, parameter :: tmax = 1.0
  real, parameter :: dtprint = 0.1
  real, parameter :: cfl = REF_CFL
  real, parameter :: eta = REF_ETA
  logical, parameter :: iwarm = .false.
  integer :: itprint0 = 0

  logical, parameter :: passives = (REF_NPAS > 0)
  integer, parameter :: ndim = 3
  integer, parameter :: nghost = 2
  integer, parameter :: neqdyn = 8
  integer, parameter :: neq = neqdyn + npas
  logical, parameter :: out_silo = .false.

  integer, parameter :: nx = nxtot, ny = nytot, nz = nztot
  integer, parameter :: np = 1, MPI_NBX = 1, MPI_NBY = 1, MPI_NBZ = 1
  integer, parameter :: nxmin = 1 - nghost, nxmax = nx + nghost
  integer, parameter :: nymin = 1 - nghost, nymax = ny + nghost
  integer, parameter :: nzmin = 1 - nghost, nzmax = nz + nghost
  integer, parameter :: master = 0
  ! the cooling modules read src/<lib>/<table> as workdir//'../src/...';
  ! the build writes workdir (the reference's src/) into refpaths.inc
#include "refpaths.inc"
  integer, parameter :: mpi_real_kind = 0

end module parameters
