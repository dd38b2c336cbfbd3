!=======================================================================
! Problem plug-in of the oracle's test problem: the four hooks every
! Guacho problem provides (interface as OT/user_mod.f90:43-122), all empty
! because the oracle sets the state through oracle/ref/ref_kat.f90.
!=======================================================================
module user_mod
  implicit none
contains

  subroutine init_user_mod()
  end subroutine init_user_mod

  subroutine initial_conditions(u)
    use parameters, only : neq, nxmin, nxmax, nymin, nymax, nzmin, nzmax
    real, intent(out) :: u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax)
    u = 0.
  end subroutine initial_conditions

  ! The oracle's user boundary (active in builds with bc_user), in the
  ! style of EXO/user_mod.f90:131-146: an inflow jet through the left x face
  ! (ghost layers 1-order..0 where (y-1/2)^2 + (z-1/2)^2 < 1/16) with
  ! vx = 1 + time, and an obstacle of interior cells within 0.15 of
  ! (0.7, 0.5, 0.5) reset to rest at density 5, pressure 1, keeping B.
  subroutine impose_user_bc(u, order)
    use parameters, only : neq, nxmin, nxmax, nymin, nymax, nzmin, nzmax, nx, ny, nz, cv
    use globals, only : dx, dy, dz, time
    real :: u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax)
    integer, intent(in) :: order
    integer :: i, j, k
    real :: xc, yc, zc, rho, vx, pr, bx, by, bz
    do k = nzmin, nzmax
      do j = nymin, nymax
        yc = (real(j) - 0.5) * dy - 0.5
        zc = (real(k) - 0.5) * dz - 0.5
        if (yc * yc + zc * zc < 0.0625) then
          do i = 1 - order, 0
            rho = 2.0
            vx = 1.0 + time
            pr = 1.0
            bx = 0.5
            u(1, i, j, k) = rho
            u(2, i, j, k) = rho * vx
            u(3, i, j, k) = 0.0
            u(4, i, j, k) = 0.0
            u(5, i, j, k) = ((0.5 * rho) * (vx * vx) + cv * pr) + 0.5 * (bx * bx)
            u(6, i, j, k) = bx
            u(7, i, j, k) = 0.0
            u(8, i, j, k) = 0.0
          end do
        end if
      end do
    end do
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          xc = (real(i) - 0.5) * dx - 0.7
          yc = (real(j) - 0.5) * dy - 0.5
          zc = (real(k) - 0.5) * dz - 0.5
          if ((xc * xc + yc * yc) + zc * zc < 0.0225) then
            bx = u(6, i, j, k)
            by = u(7, i, j, k)
            bz = u(8, i, j, k)
            u(1, i, j, k) = 5.0
            u(2, i, j, k) = 0.0
            u(3, i, j, k) = 0.0
            u(4, i, j, k) = 0.0
            u(5, i, j, k) = cv * 1.0 + 0.5 * ((bx * bx + by * by) + bz * bz)
          end if
        end do
      end do
    end do
  end subroutine impose_user_bc

  ! The oracle's user source (active in builds with user_source_terms):
  ! uniform gravity along -y plus a softened point mass at the box centre,
  ! in the style of EXO/user_mod.f90:159-200 (r^3 written as r2*sqrt(r2)).
  subroutine get_user_source_terms(pp, s, i, j, k)
    use parameters, only : neq
    use globals, only : dx, dy, dz
    integer, intent(in) :: i, j, k
    real, intent(in) :: pp(neq)
    real, intent(inout) :: s(neq)
    real, parameter :: g0 = 0.5, gm = 0.02, soft = 0.01
    real :: xc, yc, zc, r2, r3
    xc = (real(i) - 0.5) * dx - 0.5
    yc = (real(j) - 0.5) * dy - 0.5
    zc = (real(k) - 0.5) * dz - 0.5
    r2 = xc * xc + yc * yc + zc * zc + soft
    r3 = r2 * sqrt(r2)
    s(3) = s(3) - pp(1) * g0
    s(5) = s(5) - pp(1) * pp(3) * g0
    s(2) = s(2) - pp(1) * gm * xc / r3
    s(3) = s(3) - pp(1) * gm * yc / r3
    s(4) = s(4) - pp(1) * gm * zc / r3
    s(5) = s(5) - pp(1) * gm * (pp(2) * xc + pp(3) * yc + pp(4) * zc) / r3
  end subroutine get_user_source_terms

end module user_mod
