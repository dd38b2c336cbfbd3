!=======================================================================
! Orszag-Tang vortex initial conditions through the user_mod hooks
! (interface of OT/user_mod.f90:43-122): rho = 25/(36 pi), p = 5/(12 pi),
! v = (-sin 2 pi y, sin 2 pi x, 0), B = (-sin 2 pi y, sin 4 pi x, 0)/sqrt(4 pi),
! cell centres at x = (i + 1/2) dx, every cell including ghosts.
!=======================================================================
module user_mod
  use parameters
  use constants
  implicit none
  real :: rho0, p0
contains

  subroutine init_user_mod()
    rho0 = 25. / (36. * pi)
    p0 = 5. / (12. * pi)
  end subroutine init_user_mod

  subroutine initial_conditions(u)
    use globals, only : coords, dx, dy
    real, intent(out) :: u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax)
    real :: x, y, vx, vy, vz, bx, by, bz
    real, parameter :: twopi = 2. * pi
    integer :: i, j, k
    do i = nxmin, nxmax
      do j = nymin, nymax
        do k = nzmin, nzmax
          x = (float(i + coords(0) * nx) + 0.5) * dx * rsc
          y = (float(j + coords(1) * ny) + 0.5) * dy * rsc
          vx = -sin(y * twopi)
          vy = sin(x * twopi)
          vz = 0.
          bx = -sin(y * twopi) / sqrt(4 * pi)
          by = sin(2. * x * twopi) / sqrt(4 * pi)
          bz = 0.
          u(1,i,j,k) = rho0
          u(2,i,j,k) = rho0 * vx
          u(3,i,j,k) = rho0 * vy
          u(4,i,j,k) = rho0 * vz
          u(5,i,j,k) = 0.5 * rho0 * (vx**2 + vy**2 + vz**2) + cv * p0 + 0.5 * (bx**2 + by**2 + bz**2)
          u(6,i,j,k) = bx
          u(7,i,j,k) = by
          u(8,i,j,k) = bz
        end do
      end do
    end do
  end subroutine initial_conditions

  subroutine impose_user_bc(u, order)
    real :: u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax)
    integer, intent(in) :: order
  end subroutine impose_user_bc

  subroutine get_user_source_terms(pp, s, i, j, k)
    integer, intent(in) :: i, j, k
    real, intent(in) :: pp(neq)
    real, intent(inout) :: s(neq)
  end subroutine get_user_source_terms

end module user_mod
