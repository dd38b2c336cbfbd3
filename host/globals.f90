!=======================================================================
! Global state of the host, in the layout problem files expect
! (src/globals.f90:29-69 of the reference): u(neq, nxmin:nxmax,
! nymin:nymax, nzmin:nzmax) with neq fastest.  On this host the device
! holds the evolving state; u and primit are the host copies that initial
! conditions fill and dumps read.  up, f, g, h stay unallocated: their
! device counterparts never leave the GPU.
!=======================================================================
module globals
  implicit none
  real, allocatable :: u(:,:,:,:)       !< conserved variables
  real, allocatable :: up(:,:,:,:)      !< (device only)
  real, allocatable :: primit(:,:,:,:)  !< primitives (filled for dumps)
  real, allocatable :: f(:,:,:,:), g(:,:,:,:), h(:,:,:,:)   !< (device only)
  real, allocatable :: Temp(:,:,:)      !< temperature
  real, allocatable :: primit0(:,:,:,:) !< split-solver background (unused)
  real :: dx, dy, dz
  integer, dimension(0:2) :: coords = 0
  integer :: left = -1, right = -1, top = -1, bottom = -1, out = -1, in = -1
  integer :: rank = 0
  integer :: comm3d = 0
  real :: time
  real :: dt_CFL
  integer :: currentIteration
end module globals
