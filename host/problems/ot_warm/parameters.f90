!=======================================================================
! Orszag-Tang vortex, 128 x 128 x 4, as host/problems/ot, warm-started
! from its dump 3 (iwarm, itprint0 = 3: initflow reads
! OT128/BIN/points003.bin, time = 3 dtprint, the CFL ramp restarts at
! currentIteration = 1; src/init.f90:124-142,432-477).
!=======================================================================
module parameters
  use constants
  implicit none
  character(len=128), parameter :: outputpath = './OT128/'
  character(len=128), parameter :: workdir = './'
  logical, parameter :: pmhd = .false.
  logical, parameter :: mhd = .true.
  integer, parameter :: riemann_solver = SOLVER_HLLD
  logical, parameter :: eight_wave = .false.
  logical, parameter :: enable_flux_cd = .true.
  logical, parameter :: dump_divb = .true.
  logical, parameter :: out_bin = .true.
  logical, parameter :: out_vtk = .false.
  integer, parameter :: eq_of_state = EOS_ADIABATIC
  integer, parameter :: cooling = COOL_NONE
  integer, parameter :: bc_left = BC_PERIODIC, bc_right = BC_PERIODIC
  integer, parameter :: bc_bottom = BC_PERIODIC, bc_top = BC_PERIODIC
  integer, parameter :: bc_out = BC_PERIODIC, bc_in = BC_PERIODIC
  logical, parameter :: bc_user = .false.
  integer, parameter :: slope_limiter = LIMITER_MINMOD
  integer, parameter :: th_cond = TC_OFF
  logical, parameter :: tc_saturation = .false.
  logical, parameter :: dif_rad = .false.
  logical, parameter :: user_source_terms = .false.
  logical, parameter :: radiation_pressure = .false.
  integer, parameter :: npas = 0
  integer, parameter :: nxtot = 128, nytot = 128, nztot = 4
  real, parameter :: xmax = 1., ymax = 1., zmax = 4. / 128.
  real, parameter :: xphys = 1.
  real, parameter :: cv = 1.5
  real, parameter :: gamma = (cv + 1.) / cv
  real, parameter :: mu = 1.
  real, parameter :: rsc = xphys / xmax
  real, parameter :: rhosc = 1.
  real, parameter :: Tempsc = 1.
  real, parameter :: vsc2 = 1.
  real, parameter :: vsc = sqrt(vsc2)
  real, parameter :: Psc = rhosc * vsc2
  real, parameter :: tsc = rsc / sqrt(vsc2)
  real, parameter :: bsc = sqrt(4.0 * pi * Psc)
  real, parameter :: tmax = 0.5
  real, parameter :: dtprint = .1 / tsc
  real, parameter :: cfl = 0.4
  real, parameter :: eta = 0.0
  logical, parameter :: iwarm = .true.
  integer :: itprint0 = 3
  logical, parameter :: passives = .false.
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
end module parameters
